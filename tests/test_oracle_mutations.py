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


# Planted mistakes in the non-orthogonal terms, each in the forward stage and its VJP where both have it; the
# operator pins of test_oracle_nonortho.py (exactness for quadratics on sheared grids, one operator whatever the
# split between matrix and right-hand side on a bent grid) must fail for each.  On a uniform affine grid the
# matrix's corner entries cancel over the faces (the cross derivative comes from the diagonal cells alone), so the
# matrix part is pinned by the split check, whose bent grid varies the cross coefficients.
NON_ORTHO_MUTATIONS = [
    ("interpolated cross coefficient without the 1/2",
     [(": (aP[lap_ci(axis, oa)] * raP + aN[oa] * raN) * (real)0.5;",
       ": (aP[lap_ci(axis, oa)] * raP + aN[oa] * raN);"),
      ("raP_grad += aP[lap_ci(axis, oa)] * (real)0.5 * g;", "raP_grad += aP[lap_ci(axis, oa)] * g;"),
      ("raN_grad += aN[oa] * (real)0.5 * g;", "raN_grad += aN[oa] * g;")]),
    ("matrix corner terms without the tangential face sign",
     [("const real v = (real)(fs * tfs) * alpha * norm;", "const real v = (real)(fs) * alpha * norm;"),
      ("ag += (real)(fs * tfs) * norm * (sign * vg);", "ag += (real)(fs) * norm * (sign * vg);")]),
    ("corner value averaged over one cell too many",
     [("r.data = data / (real)num; r.num = num; r.btype = btype;",
       "r.data = data / (real)(num + 1); r.num = num + 1; r.btype = btype;")]),
    ("explicit non-orthogonal terms with the sign flipped",
     [("            S -= (real)fs * fa * tdg;\n", "            S += (real)fs * fa * tdg;\n")]),
    ("centre cross coefficient of the wrong axis pair",
     [(": (aP[lap_ci(axis, ta)] * raP + aN[ta] * raN) * (real)0.5;",
       ": (aP[lap_ci(axis, axis)] * raP + aN[ta] * raN) * (real)0.5;")]),
]


def _non_ortho_pins():
    import test_oracle_nonortho as N
    failed = []
    for name, pin in (("pressure2d", lambda: N.test_pressure_operator_exact_for_quadratics(*N.CASES[0])),
                      ("pressure3d", lambda: N.test_pressure_operator_exact_for_quadratics(*N.CASES[2])),
                      ("velocity2d", lambda: N.test_velocity_operator_exact_for_quadratics(*N.CASES[1])),
                      ("split", lambda: N.test_flag_split_gives_one_operator(("x",)))):
        try:
            pin()
        except AssertionError:
            failed.append(name)
    return failed


def _with_mutant(edits, tmp_path, fn):
    src = open(SRC).read()
    for old, new in edits:
        assert src.count(old) == 1, old
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
        return fn()
    finally:
        O._libs[64] = good


def test_non_ortho_pins_pass_on_the_oracle():
    assert _non_ortho_pins() == []


@pytest.mark.parametrize("name,edits", NON_ORTHO_MUTATIONS, ids=[m[0] for m in NON_ORTHO_MUTATIONS])
def test_planted_mistake_fails_non_ortho_pins(name, edits, tmp_path):
    failed = _with_mutant(edits, tmp_path, _non_ortho_pins)
    assert len(failed) >= 1, (name, failed)


# Planted mistakes in the passive-scalar stages (forward and VJP together); the scalar pins of
# test_oracle_scalar.py (Fourier-mode closed form, wall profiles, equality with a velocity component, exactness for
# a linear field with Dirichlet and Neumann walls) must fail for each.
SCALAR_MUTATIONS = [
    ("Neumann wall term with the Dirichlet factor 2 alpha",
     [("                    else r += sv * visc;\n", "                    else r += sv * visc * 2 * alpha;\n"),
      ("                        s_grad += visc * G;\n                        vg += sv * G;\n",
       "                        s_grad += visc * 2 * alpha * G;\n                        vg += sv * 2 * alpha * G;\n")]),
    ("scalar matrix with the velocity viscosity",
     [("const real visc_p = sc ? scalar_viscosity(&c, ch) : viscosity(&c, b, pos);",
       "const real visc_p = viscosity(&c, b, pos);"),
      ("real visc_n = sc ? scalar_viscosity(&c, ch) : viscosity(&c, blk(&c, nb), t);",
       "real visc_n = viscosity(&c, blk(&c, nb), t);")]),
    ("boundary flux term with the sign flipped",
     [("                    r -= sv * flux;\n", "                    r += sv * flux;\n"),
      ("real s_grad = -flux * G;", "real s_grad = flux * G;"),
      ("scatter_fixed_contravariant(&c, fb, bp, -sv * G * normal);",
       "scatter_fixed_contravariant(&c, fb, bp, sv * G * normal);")]),
    ("scalar walls always Dirichlet in the corner walks",
     [("        case GD_SCALAR_RESULT: return fixed_scalar_bc(fb, ch);",
       "        case GD_SCALAR_RESULT: return PICT_BC_DIRICHLET;")]),
    ("scalar viscosity of channel 0 for every channel",
     [("return v[c->d->scalar_viscosity_static ? 0 : ch];", "return v[0];")]),
]


def _scalar_pins():
    import test_oracle_scalar as S
    failed = []
    pins = (("fourier", S.test_fourier_mode_step_closed_form),
            ("walls", S.test_wall_profiles_dirichlet_and_neumann),
            ("velocity_inflow", lambda: S.test_scalar_channel_equals_velocity_component("outflow_channel", 0)),
            ("velocity_sheared", lambda: S.test_scalar_channel_equals_velocity_component("sheared_walls", 26)),
            ("linear_field", lambda: S.test_linear_field_is_exact_with_dirichlet_and_neumann_walls(26)))
    for name, pin in pins:
        try:
            pin()
        except AssertionError:
            failed.append(name)
    return failed


def test_scalar_pins_pass_on_the_oracle():
    assert _scalar_pins() == []


@pytest.mark.parametrize("name,edits", SCALAR_MUTATIONS, ids=[m[0] for m in SCALAR_MUTATIONS])
def test_planted_mistake_fails_scalar_pins(name, edits, tmp_path):
    failed = _with_mutant(edits, tmp_path, _scalar_pins)
    assert len(failed) >= 1, (name, failed)


# Planted mistakes in the velocity correctors of versions 4, 5 and 6; the pins of test_oracle_correctors.py
# (exactness for a linear pressure on an affine grid, the rotated cut, the closed forms of version 4) must fail for
# each.
CORRECTOR_MUTATIONS = [
    ("finite-volume wall extrapolation over a whole cell",
     [("p = pP + (pP - pin) * (real)0.5;", "p = pP + (pP - pin);")]),
    ("face metrics with a column of Minv for its row",
     [("acc[j] += T_MINV(c, F, dim, j) * (T_DET(c, F) * (p + pP) * (real)0.5 * sgn);",
       "acc[j] += T_MINV(c, F, j, dim) * (T_DET(c, F) * (p + pP) * (real)0.5 * sgn);")]),
    ("connection sign of the neighbour's flux inverted",
     [("if ((fb->axes[0] & 1) == up) tsign = -1;", "if ((fb->axes[0] & 1) != up) tsign = -1;")]),
    ("version 4 centre metric per component",
     [("            const real t = metric_orth(&c, b, pos);\n            const real cc = det * t;\n"
       "            for (int dim = 0; dim < c.D; ++dim) {\n",
       "            for (int dim = 0; dim < c.D; ++dim) {\n                I4 pw = pos; pw.a[3] = dim;\n"
       "                const real t = metric_orth(&c, b, pw);\n                const real cc = det * t;\n")]),
    ("version 4 with both sides at a wall",
     [("                    if (at_bound(pos, face, b) && prescribed(face, b)) continue;\n"
       "                    const pict_block_desc *nb = b;\n",
       "                    const pict_block_desc *nb = b;\n")]),
]


def _corrector_pins():
    import test_oracle_correctors as C
    failed = []
    pins = (("linear5", lambda: C.test_fvm_gradient_of_a_linear_pressure_is_exact_on_an_affine_grid(2, 5)),
            ("linear6_3d", lambda: C.test_fvm_gradient_of_a_linear_pressure_is_exact_on_an_affine_grid(3, 6)),
            ("rotated5", lambda: C.test_fvm_on_a_rotated_cut_matches_the_uncut_grid(5)),
            ("rotated6", lambda: C.test_fvm_on_a_rotated_cut_matches_the_uncut_grid(6)),
            ("v4_aniso", C.test_v4_centre_metric_on_an_anisotropic_grid),
            ("v4_walls", C.test_v4_wall_cells_take_only_the_inner_side))
    for name, pin in pins:
        try:
            pin()
        except (AssertionError, IndexError):
            failed.append(name)
    return failed


def test_corrector_pins_pass_on_the_oracle():
    assert _corrector_pins() == []


@pytest.mark.parametrize("name,edits", CORRECTOR_MUTATIONS, ids=[m[0] for m in CORRECTOR_MUTATIONS])
def test_planted_mistake_fails_corrector_pins(name, edits, tmp_path):
    failed = _with_mutant(edits, tmp_path, _corrector_pins)
    assert len(failed) >= 1, (name, failed)


# Planted mistakes in the ILU(0) factorization and the preconditioned iteration; the pins of test_oracle_ilu.py
# must fail for each.
ILU_MUTATIONS = [
    ("ILU(0) without the pivot division", [("            lu[kk] /= lu[diag[k]];\n", "")]),
    ("ILU(0) without the row update",
     [("if (q < B->row[k + 1] && B->col[q] == j) lu[jj] -= lu[kk] * lu[q];",
       "if (q < B->row[k + 1] && B->col[q] == j) lu[jj] -= 0 * lu[q];")]),
    ("preconditioned step along p instead of M^-1 p",
     [("for (int i = 0; i < n; ++i) { r[i] -= (real)alpha * v[i]; x[i] += (real)alpha * pw[i]; }",
       "for (int i = 0; i < n; ++i) { r[i] -= (real)alpha * v[i]; x[i] += (real)alpha * p[i]; }")]),
]


def _ilu_pins():
    import test_oracle_ilu as I
    failed = []
    pins = (("identity_P", lambda: I.test_ilu0_reproduces_the_matrix_on_its_pattern("rotated_connection", "P")),
            ("tridiagonal", I.test_tridiagonal_ilu_is_exact_and_solves_in_one_half_step),
            ("criterion", lambda: I.test_preconditioned_solution_meets_the_criterion("cavity", False)))
    for name, pin in pins:
        try:
            pin()
        except AssertionError:
            failed.append(name)
    return failed


def test_ilu_pins_pass_on_the_oracle():
    assert _ilu_pins() == []


@pytest.mark.parametrize("name,edits", ILU_MUTATIONS, ids=[m[0] for m in ILU_MUTATIONS])
def test_planted_mistake_fails_ilu_pins(name, edits, tmp_path):
    failed = _with_mutant(edits, tmp_path, _ilu_pins)
    assert len(failed) >= 1, (name, failed)


# Planted mistakes in the face-transform pressure setup; the pins of test_oracle_face_transform.py must fail.
FACE_MUTATIONS = [
    ("face flux with a column of Minv for its row",
     [("for (int j = 0; j < D; ++j) s += T_MINV(c, F, dim, j) * v[j];",
       "for (int j = 0; j < D; ++j) s += T_MINV(c, F, j, dim) * v[j];")]),
    ("face matrix coefficient with the cell's metric",
     [("if (use_face) coef = laplace_orth_face(&c, b, pos, face) * ((real)0.5 * (raP + raN));",
       "if (use_face) coef = alphaP[dim] * ((real)0.5 * (raP + raN));")]),
    ("fixed-face velocity taken as it is",
     [("        real v[3];\n        slerp2_half(vc, vn, D, v);\n",
       "        real v[3];\n        if (at_bound(pos, face, b) && b->bounds[face].type == PICT_BOUNDARY_FIXED)\n"
       "            for (int k = 0; k < D; ++k) v[k] = vn[k];\n        else slerp2_half(vc, vn, D, v);\n")]),
]


def _face_pins():
    import test_oracle_face_transform as F
    failed = []
    pins = (("constant2d", lambda: F.test_constant_field_has_no_divergence_with_face_metrics(2)),
            ("linear", lambda: F.test_linear_field_divergence_closed_form(False)),
            ("stretched", F.test_face_matrix_coefficient_is_area_over_centre_distance))
    for name, pin in pins:
        try:
            pin()
        except AssertionError:
            failed.append(name)
    return failed


def test_face_pins_pass_on_the_oracle():
    assert _face_pins() == []


@pytest.mark.parametrize("name,edits", FACE_MUTATIONS, ids=[m[0] for m in FACE_MUTATIONS])
def test_planted_mistake_fails_face_pins(name, edits, tmp_path):
    failed = _with_mutant(edits, tmp_path, _face_pins)
    assert len(failed) >= 1, (name, failed)
