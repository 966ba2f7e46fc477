"""Shows that the projection pins in test_oracle_projection.py catch plausible restatement mistakes: each case
compiles a copy of the oracle with one planted mistake (in the forward stage and its VJP together, so the
finite-difference VJP checks cannot see it), swaps it in for the FP64 oracle, and expects the pins to fail.
CPU only; the planted mistakes live in temporary copies, never in the oracle itself."""
import ctypes
import os
import subprocess

import pytest

import test_oracle_projection as T
from oracle import oracle as O

SRC = os.path.join(os.path.dirname(os.path.abspath(O.__file__)), "pict_oracle.c")

# (name, [(text in the oracle, replacement)])
MUTATIONS = [
    ("pressure matrix face coefficient without the 1/2",
     [("real coef = (real)0.5 * (alphaP[dim] * raP + alphaN * raN);", "real coef = (alphaP[dim] * raP + alphaN * raN);"),
      ("raP_grad += (real)0.5 * alphaP * cg;", "raP_grad += alphaP * cg;"),
      ("Ag[gn] += -((real)0.5 * alphaN * cg) * raN * raN;", "Ag[gn] += -(alphaN * cg) * raN * raN;")]),
    ("pressure-RHS divergence with the sign flipped",
     [("for (int dd = 0; dd < c.D; ++dd) div += fl[2 * dd + 1] - fl[2 * dd];\n            if (time_step_norm)",
       "for (int dd = 0; dd < c.D; ++dd) div -= fl[2 * dd + 1] - fl[2 * dd];\n            if (time_step_norm)"),
      ("fg[2 * dd + 1] = dg; fg[2 * dd] = -dg;", "fg[2 * dd + 1] = -dg; fg[2 * dd] = dg;")]),
    ("central pressure gradient without the 1/2",
     [("real fac = (real)0.5;\n        if (!(pos.a[dim] == 0 && prescribed(2 * dim, b)))",
       "real fac = (real)1.0;\n        if (!(pos.a[dim] == 0 && prescribed(2 * dim, b)))")]),
]

PINS = {
    "continuum": T.test_pressure_matrix_continuum_limit,
    "fourier": lambda: T.test_projection_of_fourier_mode_exact_fraction(1, 1),
    "linear_div": lambda: T.test_velocity_divergence_of_linear_field(True),
    "closed_box": T.test_projection_reduces_divergence_in_closed_box,
}


def _run_pins():
    failed = []
    for name, pin in PINS.items():
        try:
            pin()
        except AssertionError:
            failed.append(name)
    return failed


def test_pins_pass_on_the_oracle():
    assert _run_pins() == []


@pytest.mark.parametrize("name,edits", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_mistake_fails_projection_pins(name, edits, tmp_path):
    src = open(SRC).read()
    for old, new in edits:
        assert src.count(old) == 1, (name, old)
        src = src.replace(old, new)
    mut = tmp_path / "pict_oracle_mut.c"
    mut.write_text(src)
    so = tmp_path / "liboracle_mut.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c99", "-fopenmp", "-DORACLE_REAL=double", "-I",
                    os.path.dirname(SRC), "-shared", str(mut), "-o", str(so), "-lm"], check=True,
                   cwd=os.path.dirname(SRC))
    good = O.lib(64)
    O._libs[64] = O.declare(ctypes.CDLL(str(so)))
    try:
        failed = _run_pins()
    finally:
        O._libs[64] = good
    assert "fourier" in failed, (name, failed)
    assert len(failed) >= 2, (name, failed)
