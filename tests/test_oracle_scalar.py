"""Pins the oracle's passive-scalar stages (advection matrix per channel, scalar RHS, their VJPs; restated from
PISO_multiblock_cuda_kernel.cu:3601-3865, :3869-4075, :4080-4278) without the reference:
- closed form: one implicit step of a Fourier mode advected by a uniform flow on a periodic grid is the mode times
  1 / (1 + dt (i (U sin kx + V sin ky) + 2 nu (2 - cos kx - cos ky))), per channel with its own viscosity;
- boundaries: the steady profile between a Dirichlet and a Dirichlet (linear) or a Neumann (gradient) wall, with
  per-channel boundary types;
- consistency: with the domain viscosity and Dirichlet walls, channel 0 of a scalar equal to a velocity component
  has exactly that component's advection matrix and right-hand side (the velocity stages are pinned by their own
  tests), including the non-orthogonal terms on sheared grids;
- VJPs against central differences of the oracle's forward stages.
CPU only."""
import math

import pytest
import torch

import scenarios
from grad_util import OUTPUTS, create_grads, input_tensors
from oracle import oracle as O
from pict_amd import PISOtorch as P

CPU = dict(dtype=torch.float64, device="cpu")


def _periodic_box(nx, ny, channels, visc, velocity, scalar_viscosity=None):
    d = P.Domain(2, torch.tensor([visc], dtype=torch.float64), "box", torch.float64, "cpu",
                 passiveScalarChannels=channels, _allow_cpu=True)
    b = d.CreateBlockWithSize(P.Int4(nx, ny, 1))
    b.MakePeriodic("x")
    b.MakePeriodic("y")
    v = torch.zeros(1, 2, ny, nx, dtype=torch.float64)
    v[0, 0] = velocity[0]
    v[0, 1] = velocity[1]
    b.setVelocity(v)
    if scalar_viscosity is not None:
        d.setScalarViscosity(torch.tensor(scalar_viscosity, dtype=torch.float64))
    return scenarios._prepare(d)


def test_fourier_mode_step_closed_form():
    nx, ny, dt, U, V = 16, 12, 0.3, 0.7, -0.4
    nus = [0.05, 0.013]
    modes = [(2, 1, 0.3), (1, 3, -1.1)]  # (mx, my, phase) per channel
    d = _periodic_box(nx, ny, 2, 0.2, (U, V), scalar_viscosity=nus)
    b = d.blocks[0]
    j, i = torch.meshgrid(torch.arange(ny, dtype=torch.float64), torch.arange(nx, dtype=torch.float64),
                          indexing="ij")
    s0 = torch.zeros(1, 2, ny, nx, dtype=torch.float64)
    for c, (mx, my, ph) in enumerate(modes):
        s0[0, c] = torch.cos(2 * math.pi * (mx * i / nx + my * j / ny) + ph)
    b.setPassiveScalar(s0.clone())
    O.advect_scalar_step(d, torch.tensor([dt], dtype=torch.float64), 0, False, 1e-14, 5000)
    for c, (mx, my, ph) in enumerate(modes):
        kx, ky = 2 * math.pi * mx / nx, 2 * math.pi * my / ny
        lam = 1 / (1 + dt * complex(2 * nus[c] * (2 - math.cos(kx) - math.cos(ky)), U * math.sin(kx) + V * math.sin(ky)))
        theta = 2 * math.pi * (mx * i / nx + my * j / ny) + ph
        exact = lam.real * torch.cos(theta) - lam.imag * torch.sin(theta)
        err = float((b.passiveScalar[0, c] - exact).abs().max())
        assert err < 1e-10, (c, err)


def test_wall_profiles_dirichlet_and_neumann():
    """Walls at -y and +y (periodic in x), no flow: channel 0 Dirichlet at both walls gives the linear profile
    through the wall values; channel 1 Neumann at -y (value g, the outward normal derivative) gives slope -g."""
    nx, ny, dt = 6, 10, 1e12
    sB, sT, g = 0.7, -0.4, 0.15
    d = P.Domain(2, torch.tensor([0.1], dtype=torch.float64), "ch", torch.float64, "cpu", passiveScalarChannels=2,
                 _allow_cpu=True)
    b = d.CreateBlockWithSize(P.Int4(nx, ny, 1))
    b.MakePeriodic("x")
    b.CloseBoundary("-y")
    b.CloseBoundary("+y")
    gen = torch.Generator().manual_seed(3)
    b.setPassiveScalar(torch.rand(1, 2, ny, nx, generator=gen, dtype=torch.float64))
    bottom, top = b.getBoundary("-y"), b.getBoundary("+y")
    bottom.setPassiveScalar(torch.tensor([[sB, g]], dtype=torch.float64))
    bottom.setPassiveScalarType([P.BoundaryConditionType.DIRICHLET, P.BoundaryConditionType.NEUMANN])
    top.setPassiveScalar(torch.tensor([[sT, sT]], dtype=torch.float64))
    scenarios._prepare(d)
    O.advect_scalar_step(d, torch.tensor([dt], dtype=torch.float64), 0, False, 1e-15, 5000)
    y = torch.arange(ny, dtype=torch.float64) + 0.5
    lin = sB + (sT - sB) * y / ny
    neu = sT + g * (ny - y)
    s = b.passiveScalar[0]
    assert float((s[0] - lin[:, None]).abs().max()) < 1e-8
    assert float((s[1] - neu[:, None]).abs().max()) < 1e-8


def _scalar_as_velocity(d, comp, bc=P.BoundaryConditionType.DIRICHLET):
    """Channel 0 of the block and boundary scalars := velocity component `comp`; every wall of type `bc` for the
    velocity and the scalar, no scalar viscosity and no block viscosity (the domain's for both), no velocity
    source."""
    for b in d.blocks:
        if b.viscosity is not None:
            b.clearViscosity()
        s = b.passiveScalar.clone()
        s[:, 0] = b.velocity[:, comp]
        b.setPassiveScalar(s)
        b.clearVelocitySource()
        for _, fb in b.getFixedBoundaries():
            v = fb.velocity
            if v.dim() == 2:
                s = torch.zeros(1, d.getPassiveScalarChannels(), dtype=v.dtype)
                s[0, 0] = v[0, comp]
            else:
                s = torch.zeros([1, d.getPassiveScalarChannels()] + list(v.shape[2:]), dtype=v.dtype)
                s[:, 0] = v[:, comp]
            fb.setPassiveScalar(s)
            fb.setPassiveScalarType(bc)
            fb.setVelocityType(bc)
    d.clearScalarViscosity()


BUILDERS = dict(scenarios.ALL, **scenarios.NON_ORTHO, outflow_channel=scenarios.outflow_channel)
CONSISTENCY = [("cavity", 0), ("periodic_two_block", 0), ("rotated_connection", 0), ("box3d", 0),
               ("outflow_channel", 0), ("sheared_walls", 25), ("sheared_walls", 26), ("sheared_periodic_x", 25),
               ("sheared_box3d", 26)]


@pytest.mark.parametrize("name,flags", CONSISTENCY)
def test_scalar_channel_equals_velocity_component(name, flags):
    d = scenarios.with_scalars(BUILDERS[name], **CPU)
    O.build_csr_pattern(d)
    dt = 0.05
    g = torch.Generator().manual_seed(9)
    for comp in range(d.getSpatialDims()):
        _scalar_as_velocity(d, comp)
        vr = (torch.rand(d.velocityResult.shape, generator=g, dtype=torch.float64) * 2 - 1) * 0.5
        d.setVelocityResult(vr)
        sr = torch.zeros_like(d.scalarResult)
        n = d.getTotalSize()
        sr[:n] = vr[comp * n:(comp + 1) * n]
        d.setScalarResult(sr)
        O.SetupAdvectionMatrix(d, dt, flags)
        cv, av = d.C.value.clone(), d.A.clone()
        O.SetupAdvectionVelocity(d, dt, flags)
        rv = d.velocityRHS[comp * n:(comp + 1) * n].clone()
        O.SetupAdvectionMatrix(d, dt, flags, True, 0)
        assert float((d.C.value - cv).abs().max()) <= 1e-12 * float(cv.abs().max())
        assert float((d.A - av).abs().max()) <= 1e-12 * float(av.abs().max())
        O.SetupAdvectionScalar(d, dt, flags)
        assert float((d.scalarRHS[:n] - rv).abs().max()) <= 1e-12 * float(rv.abs().max())


@pytest.mark.parametrize("flags", [0, 25, 26])
def test_linear_field_is_exact_with_dirichlet_and_neumann_walls(flags):
    """A linear field s = a x + b y on a uniformly sheared box closed on all four sides has zero Laplacian, and the
    scheme is exact for it: with the wall values (channel 0, Dirichlet) or the walls' outward fluxes
    sum_c alpha_nc ds/dxi_c (channel 1, Neumann; alpha_nn ds/dxi_n without the non-orthogonal terms) prescribed,
    the full operator C s - rhs (matrix terms plus the
    explicit non-orthogonal terms read from s itself, no flow) vanishes in every cell that does not touch two
    walls (with the direct terms in the matrix, flags 25: in every cell off the walls).  This pins the channels' boundary types in the corner walks (a Neumann wall's corners are one-sided
    differences, a Dirichlet wall's corners its values) and the Neumann wall term (the prescribed flux times the
    scalar viscosity)."""
    from scenarios import sheared_box
    nx, ny, h, sh = 9, 8, 0.1, 0.4
    scenarios._SCALAR_CHANNELS[0] = 2
    try:
        d = sheared_box(device="cpu", size=(nx, ny), shear=(sh,), walls_moving=False)
    finally:
        scenarios._SCALAR_CHANNELS[0] = 0
    b = d.blocks[0]
    b.setVelocity(torch.zeros_like(b.velocity))
    a_, b_ = 0.7, -1.3
    j, i = torch.meshgrid(torch.arange(ny, dtype=torch.float64) + 0.5, torch.arange(nx, dtype=torch.float64) + 0.5,
                          indexing="ij")
    field = lambda ii, jj: a_ * h * (ii + sh * jj) + b_ * h * jj  # noqa: E731
    s = field(i, j)
    b.setPassiveScalar(torch.stack([s, s])[None].contiguous())
    T = b.transform[0, 0, 0]
    Minv = T[4:8].reshape(2, 2)
    alpha = T[8] * Minv @ Minv.T
    grad_xi = torch.tensor([a_ * h, a_ * h * sh + b_ * h], dtype=torch.float64)  # ds/dxi, ds/deta
    # computational-space flux through x and y faces per unit viscosity; without the non-orthogonal terms the
    # scheme's faces carry the orthogonal part only
    flux = alpha @ grad_xi if flags else torch.diagonal(alpha) * grad_xi
    for face in range(4):
        ax, up = face >> 1, face & 1
        fb = b.getBoundary(face)
        if ax == 0:
            jj = torch.arange(ny, dtype=torch.float64) + 0.5
            ii = torch.full_like(jj, float(nx) if up else 0.0)
            vals = field(ii, jj)[:, None]  # [y, 1]
        else:
            ii = torch.arange(nx, dtype=torch.float64) + 0.5
            jj = torch.full_like(ii, float(ny) if up else 0.0)
            vals = field(ii, jj)[None, :]  # [1, x]
        g = (flux[ax] if up else -flux[ax]) * torch.ones_like(vals)
        fb.setPassiveScalar(torch.stack([vals, g])[None].contiguous())
        fb.setPassiveScalarType([P.BoundaryConditionType.DIRICHLET, P.BoundaryConditionType.NEUMANN])
    d.setScalarViscosity(torch.tensor([0.9], dtype=torch.float64))
    O.build_csr_pattern(d)
    n = d.getTotalSize()
    d.setScalarResult(torch.cat([s.reshape(-1), s.reshape(-1)]).contiguous())
    O.SetupAdvectionScalar(d, 0.05, flags)
    rows = d.C.row.long()
    cols = d.C.index.long()
    inner = torch.ones(ny, nx, dtype=torch.bool)
    for jj in (0, ny - 1):
        for ii in (0, nx - 1):
            inner[jj, ii] = False
    if flags & 1:  # NON_ORTHO_DIRECT_MATRIX
        # with the direct neighbours' corner terms in the matrix the wall rows are not exact (the reference's
        # matrix and RHS forms of those terms differ at walls; the flag-split pin holds in the interior only)
        inner[0, :] = inner[-1, :] = False
        inner[:, 0] = inner[:, -1] = False
    for ch in range(2):
        O.SetupAdvectionMatrix(d, 0.05, flags, True, ch)
        M = torch.sparse_csr_tensor(rows, cols, d.C.value, (n, n)).to_dense()
        res = (M @ s.reshape(-1) - d.scalarRHS[ch * n:(ch + 1) * n]).reshape(ny, nx)
        scale = float(d.scalarRHS[ch * n:(ch + 1) * n].abs().max())
        assert float(res[inner].abs().max()) < 1e-11 * scale, (ch, flags, res)


# ---------------------------------------------------------------------------
# VJPs against central differences
# ---------------------------------------------------------------------------

EPS = 1e-6
SCALAR_STAGES = {
    "matrix_ch0": (lambda d, dt, f: O.SetupAdvectionMatrix(d, dt, f, True, 0),
                   lambda d, dt, f: O.SetupAdvectionMatrixGrad(d, dt, f, True, 0),
                   ["VELOCITY", "BOUNDARY_VELOCITY", "VISCOSITY", "PASSIVE_SCALAR_VISCOSITY"], ["A", "C"]),
    "matrix_ch1": (lambda d, dt, f: O.SetupAdvectionMatrix(d, dt, f, True, 1),
                   lambda d, dt, f: O.SetupAdvectionMatrixGrad(d, dt, f, True, 1),
                   ["VELOCITY", "BOUNDARY_VELOCITY", "VISCOSITY", "PASSIVE_SCALAR_VISCOSITY"], ["A", "C"]),
    "rhs": (lambda d, dt, f: O.SetupAdvectionScalar(d, dt, f), lambda d, dt, f: O.SetupAdvectionScalarGrad(d, dt, f),
            ["PASSIVE_SCALAR", "BOUNDARY_PASSIVE_SCALAR", "BOUNDARY_VELOCITY", "VISCOSITY", "PASSIVE_SCALAR_VISCOSITY",
             "PASSIVE_SCALAR_RESULT"], ["PASSIVE_SCALAR_RHS"]),
}
FD_CASES = [("cavity", 0, "per_channel"), ("periodic_two_block", 0, None), ("rotated_connection", 0, "static"),
            ("box3d", 0, "per_channel"), ("outflow_channel", 0, "per_channel"), ("sheared_walls", 25, "per_channel"),
            ("sheared_two_block_rotated", 26, None), ("sheared_box3d", 25, "static")]


def scalar_vjp_check(d, stage, flags, seed=0):
    fwd, bwd, inputs, outputs = SCALAR_STAGES[stage]
    g = torch.Generator().manual_seed(200 + seed)
    dt = 0.05
    d.setScalarResult((torch.rand(d.scalarResult.shape, generator=g, dtype=torch.float64) * 2 - 1) * 0.5)
    fwd(d, dt, flags)
    ybar = {o: (torch.rand(OUTPUTS[o][0](d).shape, generator=g, dtype=torch.float64) * 2 - 1) for o in outputs}
    create_grads(d)
    d.CreateViscosityGrad()
    for o in outputs:
        OUTPUTS[o][1](d, ybar[o].clone())
    d._set_changed(False)
    bwd(d, dt, flags)

    def loss():
        fwd(d, dt, flags)
        return sum(float((ybar[o] * OUTPUTS[o][0](d)).sum()) for o in outputs)

    checked = 0
    for name in inputs:
        for t, grad_of in input_tensors(d, name):
            gt = grad_of().clone()
            V = torch.rand(t.shape, generator=g, dtype=torch.float64) * 2 - 1
            base = t.clone()
            t.copy_(base + EPS * V)
            lp = loss()
            t.copy_(base - EPS * V)
            lm = loss()
            t.copy_(base)
            fd = (lp - lm) / (2 * EPS)
            an = float((gt * V).sum())
            scale = max(abs(fd), abs(an), 1e-8)
            noise = 1e-14 * (abs(lp) + abs(lm)) / EPS  # rounding of the two losses in the difference quotient
            assert abs(fd - an) <= 2e-6 * scale + noise + 1e-9, (stage, name, tuple(t.shape), fd, an)
            checked += 1
    return checked


@pytest.mark.parametrize("stage", list(SCALAR_STAGES))
@pytest.mark.parametrize("name,flags,visc", FD_CASES)
def test_scalar_vjp_matches_finite_differences(stage, name, flags, visc):
    d = scenarios.with_scalars(BUILDERS[name], viscosity=visc, **CPU)
    O.build_csr_pattern(d)
    assert scalar_vjp_check(d, stage, flags) > 0
