"""Workload builders for the benchmark configurations and tests.

Each builder returns a prepared PISOtorch.Domain (PrepareSolve called).  Grids
are given as vertex coordinates, so the transformation-metric path of the
kernels is exercised, as in the reference's samples
(minimal_lid_sample.py:25-86, learning_sample.py:34-60).
"""
import math

import torch

from . import PISOtorch


def uniform_coords(nx, ny, nz=None, lx=1.0, ly=1.0, lz=1.0, dtype=torch.float32, device="cuda", x0=0.0, y0=0.0,
                   z0=0.0):
    """Vertex coordinates [1, D, (nz+1), ny+1, nx+1] of a uniform box."""
    xs = torch.linspace(x0, x0 + lx, nx + 1, dtype=torch.float64)
    ys = torch.linspace(y0, y0 + ly, ny + 1, dtype=torch.float64)
    if nz is None:
        gy, gx = torch.meshgrid(ys, xs, indexing="ij")
        c = torch.stack([gx, gy])[None]
    else:
        zs = torch.linspace(z0, z0 + lz, nz + 1, dtype=torch.float64)
        gz, gy, gx = torch.meshgrid(zs, ys, xs, indexing="ij")
        c = torch.stack([gx, gy, gz])[None]
    return c.to(dtype=dtype, device=device).contiguous()


def lid_cavity(n, viscosity=0.01, lid_velocity=1.0, dtype=torch.float32, device="cuda", coords=True, name="cavity"):
    """2D lid-driven cavity on the unit square: all walls closed, the +y wall moves with (lid, 0)."""
    dom = PISOtorch.Domain(2, torch.tensor([viscosity], dtype=dtype), name, dtype, device, passiveScalarChannels=0)
    if coords:
        blk = dom.CreateBlock(vertexCoordinates=uniform_coords(n, n, dtype=dtype, device=device), name="cavity")
    else:
        blk = dom.CreateBlockWithSize(PISOtorch.Int4(n, n, 1), name="cavity")
    blk.CloseAllBoundaries()
    blk.CloseBoundary("+y", velocity=torch.tensor([[lid_velocity, 0.0]], dtype=dtype, device=device))
    dom.PrepareSolve()
    return dom


def periodic_box_2block(nx=256, ny_block=128, length=2 * math.pi, viscosity=1e-3, dtype=torch.float32, device="cuda",
                        coords=True, name="box2"):
    """2D periodic box of 2 blocks (nx x ny_block each) stacked along y and connected on both y faces, periodic in x."""
    dom = PISOtorch.Domain(2, torch.tensor([viscosity], dtype=dtype), name, dtype, device, passiveScalarChannels=0)
    h = length / nx
    blocks = []
    for i in range(2):
        if coords:
            c = uniform_coords(nx, ny_block, lx=length, ly=ny_block * h, y0=i * ny_block * h, dtype=dtype,
                               device=device)
            b = dom.CreateBlock(vertexCoordinates=c, name="b%d" % i)
        else:
            b = dom.CreateBlockWithSize(PISOtorch.Int4(nx, ny_block, 1), name="b%d" % i)
        b.MakePeriodic("x")
        blocks.append(b)
    blocks[0].ConnectBlock("+y", blocks[1], "-y", "-x")
    blocks[1].ConnectBlock("+y", blocks[0], "-y", "-x")
    dom.PrepareSolve()
    return dom


def spectral_velocity_2d(n=256, kmin=1, kmax=32, slope=-5.0 / 3.0, rms=1.0, seed=0):
    """Divergence-free random 2D velocity on an n x n periodic grid from a streamfunction with random phases and
    energy spectrum E(k) ~ k^slope for kmin <= |k| <= kmax, scaled to the given RMS speed.
    Returns u [2, n, n] (float64, [component, y, x])."""
    g = torch.Generator().manual_seed(seed)
    k = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64)
    ky, kx = torch.meshgrid(k, k, indexing="ij")
    kk = torch.sqrt(kx ** 2 + ky ** 2)
    band = (kk >= kmin) & (kk <= kmax)
    amp = torch.zeros_like(kk)
    # |u_k|^2 ~ E(k)/k (shell density), |psi_k| = |u_k| / k
    amp[band] = torch.sqrt(kk[band] ** slope / kk[band]) / kk[band]
    phase = torch.rand((n, n), generator=g, dtype=torch.float64) * 2 * math.pi
    psi_h = amp * torch.exp(1j * phase)
    u = torch.fft.ifft2(1j * ky * psi_h).real
    v = torch.fft.ifft2(-1j * kx * psi_h).real
    vel = torch.stack([u, v])
    vel *= rms / torch.sqrt((vel ** 2).sum(0).mean())
    return vel


def periodic_box_c3(n=256, viscosity=1e-3, dtype=torch.float32, device="cuda", seed=0):
    """Differentiation configuration: 2-block periodic box [0,2pi]^2 of n x n/2 blocks, spectral initial velocity
    (k^-5/3, k in 1..32, RMS 1).  Returns (domain, time_step) with dt = 0.25 h / max|u0|."""
    dom = periodic_box_2block(n, n // 2, 2 * math.pi, viscosity, dtype, device, coords=True, name="box_c3")
    vel = spectral_velocity_2d(n, seed=seed)
    half = n // 2
    for i, b in enumerate(dom.blocks):
        b.setVelocity(vel[None, :, i * half:(i + 1) * half, :].to(dtype=dtype, device=device).contiguous())
    dom.UpdateDomainData()
    h = 2 * math.pi / n
    dt = 0.25 * h / torch.linalg.vector_norm(vel, dim=0).max().item()
    return dom, dt


C4_SIDES = (32, 48, 64, 96, 128, 192, 256)


def cavity_batch_params(count=64, seed=0, sides=C4_SIDES, visc_range=(1e-3, 1e-1)):
    """Side lengths (uniform over `sides`) and viscosities (log-uniform over visc_range) of a batch of lid
    cavities, drawn from torch.Generator seeded with `seed`: sides first, then viscosities."""
    g = torch.Generator().manual_seed(seed)
    idx = torch.randint(0, len(sides), (count,), generator=g)
    lo, hi = math.log10(visc_range[0]), math.log10(visc_range[1])
    visc = 10.0 ** (lo + (hi - lo) * torch.rand(count, generator=g, dtype=torch.float64))
    return [int(sides[i]) for i in idx.tolist()], [float(v) for v in visc.tolist()]


def cavity_batch_c4(count=64, seed=0, dtype=torch.float32, device="cuda", sides=C4_SIDES):
    """Batching configuration: `count` independent lid-driven cavities on the unit square (lid velocity 1 on +y,
    zero initial state), sides and viscosities from cavity_batch_params.  Returns (domains, sides, viscosities,
    time_step) with dt = 0.4 / max(sides) (CFL <= 0.4 at the lid speed)."""
    ns, visc = cavity_batch_params(count, seed, sides)
    doms = [lid_cavity(n, viscosity=v, dtype=dtype, device=device, name="cavity%d" % i)
            for i, (n, v) in enumerate(zip(ns, visc))]
    return doms, ns, visc, 0.4 / max(sides)


def channel_coords(nx=192, ny=96, nz=96, stretch=2.0, dtype=torch.float32, device="cuda"):
    """Vertex coordinates [1, 3, nz+1, ny+1, nx+1] of a channel [-pi, pi] x [-1, 1] x [-pi/2, pi/2] (half-height 1,
    the proportions of the reference's channel grids, channel_flow_sim.py:110-141) with a wall-normal tanh stretch:
    y_j = tanh(stretch (2j/ny - 1)) / tanh(stretch), cells cluster at the walls y = +-1."""
    xs = torch.linspace(-math.pi, math.pi, nx + 1, dtype=torch.float64)
    eta = torch.linspace(-1.0, 1.0, ny + 1, dtype=torch.float64)
    ys = torch.tanh(stretch * eta) / math.tanh(stretch)
    zs = torch.linspace(-math.pi / 2, math.pi / 2, nz + 1, dtype=torch.float64)
    gz, gy, gx = torch.meshgrid(zs, ys, xs, indexing="ij")
    return torch.stack([gx, gy, gz])[None].to(dtype=dtype, device=device).contiguous()


def channel_c5(nx=192, ny=96, nz=96, stretch=2.0, re_tau=180.0, noise=0.05, bulk=15.7, seed=0,
               dtype=torch.float32, device="cuda"):
    """3-D channel configuration: one block of nx x ny x nz cells on channel_coords, periodic in x and z, no-slip walls at
    y = +-1 (channel_flow_sim.py:385-398).  With half-height and friction velocity 1: viscosity 1/re_tau and a static
    streamwise forcing (u_tau^2 / delta = 1).  Initial state (amplitudes not fixed by the reference, fixed here): the
    laminar parabola with the bulk velocity of a turbulent channel at this Re_tau (u = 1.5 bulk (1 - y^2)) plus seeded
    Gaussian noise of `noise` times its centre-line velocity in every component.  Returns (domain, time_step, forcing)
    with dt = 0.4 min(dx) / (1.5 bulk), forcing the static source tensor [1, 3]."""
    coords = channel_coords(nx, ny, nz, stretch, dtype, device)
    dom = PISOtorch.Domain(3, torch.tensor([1.0 / re_tau], dtype=dtype), "channel", dtype, device,
                           passiveScalarChannels=0)
    blk = dom.CreateBlock(vertexCoordinates=coords, name="channel")
    blk.MakePeriodic("x")
    blk.MakePeriodic("z")
    blk.CloseBoundary("-y")
    blk.CloseBoundary("+y")
    # cell-centre y of every cell [nz, ny, nx] (the grid is a tensor product: the mean of two y-adjacent vertices)
    yc = 0.5 * (coords[0, 1, :-1, :-1, :-1] + coords[0, 1, :-1, 1:, :-1]).to(torch.float64).cpu()
    ucl = 1.5 * bulk
    g = torch.Generator().manual_seed(seed)
    vel = torch.randn([1, 3, nz, ny, nx], generator=g, dtype=torch.float64) * (noise * ucl)
    vel[0, 0] += ucl * (1.0 - yc ** 2)
    blk.setVelocity(vel.to(dtype=dtype, device=device).contiguous())
    forcing = torch.tensor([[1.0, 0.0, 0.0]], dtype=dtype, device=device)
    blk.setVelocitySource(forcing)
    dom.PrepareSolve()
    dx = 2 * math.pi / nx
    return dom, 0.4 * dx / ucl, forcing


def sgs_cnn(channels=16, seed=0, dtype=torch.float32, device="cuda"):
    """The SGS forcing model of the 3-D channel configuration: three 3^3 convolutions 3 -> channels -> channels -> 3
    with 'same' padding and ReLU between them (channel_flow_learning.py:276-297), random initialisation from `seed`
    (torch.nn.Conv3d's); the convolutions use sgs.Conv3dSame's backward."""
    from .sgs import Conv3dSame
    torch.manual_seed(seed)
    net = torch.nn.Sequential(Conv3dSame(3, channels), torch.nn.ReLU(), Conv3dSame(channels, channels),
                              torch.nn.ReLU(), Conv3dSame(channels, 3))
    return net.to(dtype=dtype, device=device)


def tanh_radii(r1, r2, n, ratio=1.05):
    """n+1 radii from r1 to r2 with a one-sided tanh stretch, r_j = r1 + (r2 - r1) (1 - tanh(s (1 - j/n)) / tanh(s)),
    cells clustered at r1; the stretch s is chosen (bisection) so that the second cell is `ratio` times the first."""
    def radii(s):
        eta = torch.linspace(0.0, 1.0, n + 1, dtype=torch.float64)
        return r1 + (r2 - r1) * (1.0 - torch.tanh(s * (1.0 - eta)) / math.tanh(s))

    def first_ratio(s):
        r = radii(s)
        return float((r[2] - r[1]) / (r[1] - r[0]))

    lo, hi = 1e-6, 20.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first_ratio(mid) < ratio:
            lo = mid
        else:
            hi = mid
    return radii(0.5 * (lo + hi))


def ogrid_coords(n_angle, radii, start_deg, span_deg=-90.0, dtype=torch.float64, device="cpu"):
    """Vertex coordinates [1, 2, len(radii), n_angle+1] of one O-grid block around the origin: x along the angle
    (from start_deg, spanning span_deg; negative = clockwise), y along the radius (the layout of the reference's
    make_torus_2D blocks, lib/data/shapes.py:554-600)."""
    ang = torch.deg2rad(torch.tensor(start_deg, dtype=torch.float64)
                        + torch.linspace(0.0, span_deg, n_angle + 1, dtype=torch.float64))
    r = radii.to(torch.float64)
    x = r[:, None] * torch.cos(ang)[None, :]
    y = r[:, None] * torch.sin(ang)[None, :]
    return torch.stack([x, y])[None].to(dtype=dtype, device=device).contiguous()


def cylinder_ogrid_c2(n_angle=128, n_radius=128, r1=0.5, r2=5.0, ratio=1.05, viscosity=0.005, inflow=1.0,
                      noise=0.01, seed=0, dtype=torch.float32, device="cuda"):
    """Precision configuration: flow past a cylinder of radius r1 on an O-grid of 4 blocks (n_angle x n_radius cells
    each, the quadrants clockwise from the top as in the reference's make4BlockTorus_VortexStreet,
    tests/test_setups.py:204-350), blocks joined +x -> -x around the cylinder, radial spacing from tanh_radii.
    No-slip wall on the cylinder (-y of every block); on the outer boundary (+y) uniform inflow (inflow, 0) on the
    two upstream quadrants and an advective outflow, initially (inflow, 0), on the two downstream ones.  Initial
    velocity: (inflow, 0) plus seeded Gaussian noise of std `noise` per cell and component.
    Returns (domain, outflow boundaries, inflow velocity [1, 2]); step with
    PISOtorch_simulation.update_advective_boundaries(domain, outflow, inflow_velocity, dt) as the PRE hook."""
    radii = tanh_radii(r1, r2, n_radius, ratio)
    dom = PISOtorch.Domain(2, torch.tensor([viscosity], dtype=dtype), "cylinder_ogrid", dtype, device,
                           passiveScalarChannels=0)
    names = ("top_right", "bottom_right", "bottom_left", "top_left")
    blocks = []
    for k, start in enumerate((90.0, 0.0, -90.0, -180.0)):
        c = ogrid_coords(n_angle, radii, start, -90.0, dtype=dtype, device=device)
        b = dom.CreateBlock(vertexCoordinates=c, name=names[k])
        b.CloseAllBoundaries()
        blocks.append(b)
    for k in range(4):
        blocks[k].ConnectBlock("+x", blocks[(k + 1) % 4], "-x", "-y")
    vin = torch.tensor([[inflow, 0.0]], dtype=dtype, device=device)
    for b in (blocks[2], blocks[3]):
        b.getBoundary("+y").setVelocity(vin)
    outflow = []
    for b in (blocks[0], blocks[1]):
        bd = b.getBoundary("+y")
        bd.setVelocity(vin)
        bd.makeVelocityVarying()
        outflow.append(bd)
    g = torch.Generator().manual_seed(seed)
    for b in blocks:
        u = torch.randn(b.velocity.shape, generator=g, dtype=torch.float64) * noise
        u[:, 0] += inflow
        b.setVelocity(u.to(dtype=dtype, device=device).contiguous())
    dom.PrepareSolve()
    # one update with zero advection speed balances the boundary fluxes before the first step (test_setups.py:341-343)
    from . import PISOtorch_simulation
    PISOtorch_simulation.update_advective_boundaries(dom, outflow, torch.zeros_like(vin),
                                                     torch.ones([1], dtype=dtype, device=device))
    return dom, outflow, vin
