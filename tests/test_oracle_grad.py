"""Pins the oracle's backward stages (the VJPs restated from the reference's _GRAD kernels) against central finite
differences of the oracle's own forward stages, in FP64 on the small scenarios: for every tracked input X of a stage
and a random direction V, <dL/dX, V> must equal (L(X + eps V) - L(X - eps V)) / (2 eps) with L = <Ybar, F(X)> for
random output cotangents Ybar.  A dropped term, a wrong sign, a transposed index or a wrong neighbour across a
connection changes the left side but not the right.

The tracked inputs per stage follow the reference's PISOtorch_diff.py wrappers (:380, :461, :548, :809, :907, :987,
:1085, :1174)."""
import pytest
import torch

import scenarios
from grad_util import OUTPUTS, create_grads, input_tensors
from oracle import oracle as O

CPU = dict(dtype=torch.float64, device="cpu")
EPS = 1e-6


def _gen(seed):
    return torch.Generator().manual_seed(seed)


def _rand_like(t, g, scale=1.0):
    return (torch.rand(t.shape, generator=g, dtype=torch.float64) * 2 - 1) * scale


def _prepare(d, dt):
    """Realistic intermediate state: C, A from the advection stage, a random velocity result, the pressure RHS."""
    g = _gen(7)
    O.SetupAdvectionMatrix(d, dt)
    O.SetupAdvectionVelocity(d, dt)
    d.setVelocityResult(_rand_like(d.velocityResult, g, 0.5))
    O.SetupPressureMatrix(d, dt)
    O.SetupPressureRHS(d, dt)
    d._set_changed(False)


DT = 0.05
# stage: (forward, backward, tracked inputs, outputs, grad tensors to create)
STAGES = {
    "SetupAdvectionMatrix": (
        lambda d, dt: O.SetupAdvectionMatrix(d, dt), lambda d, dt: O.SetupAdvectionMatrixGrad(d, dt),
        ["VELOCITY", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK"], ["A", "C"]),
    "SetupAdvectionVelocity": (
        lambda d, dt: O.SetupAdvectionVelocity(d, dt), lambda d, dt: O.SetupAdvectionVelocityGrad(d, dt),
        ["VELOCITY", "VELOCITY_SOURCE", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK"], ["VELOCITY_RHS"]),
    "SetupPressureMatrix": (
        lambda d, dt: O.SetupPressureMatrix(d, dt), lambda d, dt: O.SetupPressureMatrixGrad(d, dt),
        ["A"], ["P"]),
    "SetupPressureRHSdiv": (
        lambda d, dt: O.SetupPressureRHSdiv(d, dt), lambda d, dt: O.SetupPressureRHSdivGrad(d, dt),
        ["PRESSURE_RHS", "BOUNDARY_VELOCITY"], ["PRESSURE_RHS_DIV"]),
    "SetupPressureRHS": (
        lambda d, dt: O.SetupPressureRHS(d, dt), lambda d, dt: O.SetupPressureRHSGrad(d, dt),
        ["VELOCITY", "VELOCITY_SOURCE", "VELOCITY_RESULT", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK", "C",
         "A"], ["PRESSURE_RHS", "PRESSURE_RHS_DIV"]),
    "SetupPressureCorrection": (
        lambda d, dt: O.SetupPressureCorrection(d, dt), lambda d, dt: O.SetupPressureCorrectionGrad(d, dt),
        ["VELOCITY", "VELOCITY_SOURCE", "VELOCITY_RESULT", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK", "C",
         "A"], ["P", "PRESSURE_RHS", "PRESSURE_RHS_DIV"]),
    "CorrectVelocity": (
        lambda d, dt: O.CorrectVelocity(d, dt, 1), lambda d, dt: O.CorrectVelocityGrad(d, dt),
        ["A", "PRESSURE", "PRESSURE_RHS"], ["VELOCITY_RESULT"]),
}


def run_check(d, stage, seed=0):
    fwd, bwd, inputs, outputs = STAGES[stage]
    g = _gen(100 + seed)
    _prepare(d, DT)
    ybar = {o: _rand_like(OUTPUTS[o][0](d), g) for o in outputs}
    # analytic VJP
    create_grads(d)
    for o in outputs:
        OUTPUTS[o][1](d, ybar[o].clone())
    d._set_changed(False)
    bwd(d, DT)

    def loss():
        fwd(d, DT)
        return sum(float((ybar[o] * OUTPUTS[o][0](d)).sum()) for o in outputs)

    checked = 0
    for name in inputs:
        for t, grad_of in input_tensors(d, name):
            gt = grad_of().clone()
            V = _rand_like(t, g)
            base = t.clone()
            t.copy_(base + EPS * V)
            lp = loss()
            t.copy_(base - EPS * V)
            lm = loss()
            t.copy_(base)
            fd = (lp - lm) / (2 * EPS)
            an = float((gt * V).sum())
            scale = max(abs(fd), abs(an), 1e-8)
            assert abs(fd - an) <= 2e-6 * scale + 1e-9, (stage, name, tuple(t.shape), fd, an)
            checked += 1
    return checked


@pytest.mark.parametrize("stage", list(STAGES))
@pytest.mark.parametrize("name", list(scenarios.ALL))
def test_oracle_vjp_matches_finite_differences(stage, name):
    d = scenarios.ALL[name](**CPU)
    O.build_csr_pattern(d)
    assert run_check(d, stage) > 0


def test_fd_check_detects_a_planted_sign_error(monkeypatch):
    """The check above is not vacuous: flipping the sign of one term of the oracle's VJP makes it fail."""
    d = scenarios.ALL["periodic_two_block"](**CPU)
    O.build_csr_pattern(d)
    orig = O.SetupAdvectionVelocityGrad

    def broken(dm, dt):
        orig(dm, dt)
        for b in dm.blocks:
            b.velocityGrad.mul_(-1.0)
    monkeypatch.setitem(STAGES, "SetupAdvectionVelocity", (STAGES["SetupAdvectionVelocity"][0], broken,
                                                            *STAGES["SetupAdvectionVelocity"][2:]))
    with pytest.raises(AssertionError):
        run_check(d, "SetupAdvectionVelocity")


@pytest.mark.parametrize("name", ["periodic_two_block", "cavity", "rotated_connection", "box3d"])
@pytest.mark.parametrize("correctors", [0, 1, 2])
def test_oracle_step_adjoint_matches_finite_differences(name, correctors):
    """oracle.piso_step_vjp (the reverse of a whole PISO step: stage VJPs, adjoint solves on A^T, -lambda x^T on the
    pattern, the pressure mean subtraction) against central differences of oracle.piso_step."""
    d = scenarios.ALL[name](**CPU)
    O.build_csr_pattern(d)
    u0 = [b.velocity.clone() for b in d.blocks]
    g = _gen(3)
    gout = [_rand_like(u, g) for u in u0]
    gu, _, _ = O.piso_step_vjp(d, DT, gout, corrector_steps=correctors)

    def loss(us):
        dd = scenarios.ALL[name](**CPU)
        O.build_csr_pattern(dd)
        for b, u in zip(dd.blocks, us):
            b.setVelocity(u.clone())
        O.piso_step(dd, DT, correctors, 1e-13, 1e-13)
        return sum(float((b.velocity * go).sum()) for b, go in zip(dd.blocks, gout))

    V = [_rand_like(u, g) for u in u0]
    fd = (loss([u + EPS * v for u, v in zip(u0, V)]) - loss([u - EPS * v for u, v in zip(u0, V)])) / (2 * EPS)
    an = sum(float((a * v).sum()) for a, v in zip(gu, V))
    assert abs(fd - an) <= 1e-5 * abs(fd) + 1e-9, (fd, an)


# Non-orthogonal terms: the stages with the driver's flags (25: centre and direct matrix terms, diagonal RHS
# terms) and with the RHS carrying the direct terms too (26 = 16|2|8), on sheared and bent grids, across rotated
# connections and periodic wraps.  The RHS stages then also depend on the previous iterate (velocity result,
# pressure result) and the pressure RHS on A (PISOtorch_diff.py:546-560, :985-999).
def _no_stages(flags):
    return {
        "SetupAdvectionMatrix": (
            lambda d, dt: O.SetupAdvectionMatrix(d, dt, flags), lambda d, dt: O.SetupAdvectionMatrixGrad(d, dt, flags),
            ["VELOCITY", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK"], ["A", "C"]),
        "SetupAdvectionVelocity": (
            lambda d, dt: O.SetupAdvectionVelocity(d, dt, flags),
            lambda d, dt: O.SetupAdvectionVelocityGrad(d, dt, flags),
            ["VELOCITY", "VELOCITY_SOURCE", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK", "VELOCITY_RESULT"],
            ["VELOCITY_RHS"]),
        "SetupPressureMatrix": (
            lambda d, dt: O.SetupPressureMatrix(d, dt, flags), lambda d, dt: O.SetupPressureMatrixGrad(d, dt, flags),
            ["A"], ["P"]),
        "SetupPressureRHSdiv": (
            lambda d, dt: O.SetupPressureRHSdiv(d, dt, flags), lambda d, dt: O.SetupPressureRHSdivGrad(d, dt, flags),
            ["PRESSURE_RHS", "BOUNDARY_VELOCITY", "PRESSURE_RESULT", "A"], ["PRESSURE_RHS_DIV"]),
        "SetupPressureCorrection": (
            lambda d, dt: O.SetupPressureCorrection(d, dt, flags),
            lambda d, dt: O.SetupPressureCorrectionGrad(d, dt, flags),
            ["VELOCITY", "VELOCITY_SOURCE", "VELOCITY_RESULT", "BOUNDARY_VELOCITY", "VISCOSITY", "VISCOSITY_BLOCK",
             "C", "A", "PRESSURE_RESULT"], ["P", "PRESSURE_RHS", "PRESSURE_RHS_DIV"]),
    }


@pytest.mark.parametrize("flags", [25, 26])
@pytest.mark.parametrize("stage", ["SetupAdvectionMatrix", "SetupAdvectionVelocity", "SetupPressureMatrix",
                                   "SetupPressureRHSdiv", "SetupPressureCorrection"])
@pytest.mark.parametrize("name", list(scenarios.NON_ORTHO))
def test_oracle_non_ortho_vjp_matches_finite_differences(name, stage, flags, monkeypatch):
    d = scenarios.NON_ORTHO[name](**CPU)
    g = _gen(11)
    d.setPressureResult(_rand_like(d.pressureResult, g, 0.3))
    for k, v in _no_stages(flags).items():
        monkeypatch.setitem(STAGES, k, v)
    assert run_check(d, stage) > 0
