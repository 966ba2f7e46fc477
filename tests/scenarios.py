"""Small seeded domains exercising every boundary/connection/transform path.

Restates the kinds of setups in the reference's tests/test_setups.py:30-496
(1/2/4-block, periodic/closed, uniform/stretched grids, 2D and 3D, connections
with permuted and inverted axes) at sizes the oracle finishes in milliseconds.
"""
import math

import torch

from pict_amd import PISOtorch as P
from pict_amd.configs import uniform_coords


# passive-scalar channels of the domains the builders create (set by with_scalars around a builder call)
_SCALAR_CHANNELS = [0]


def _domain(dims, visc, dtype, device):
    return P.Domain(dims, torch.tensor([visc], dtype=dtype), "test", dtype, device,
                    passiveScalarChannels=_SCALAR_CHANNELS[0], _allow_cpu=(str(device) == "cpu"))


def _coords_block(d, coords, name="Block"):
    """Block from vertex coordinates; on a host (oracle) domain the metrics come from the oracle."""
    if str(d.getDevice()) == "cpu":
        from oracle import oracle as O
        c64 = coords.double()
        b = d.CreateBlock(velocity=torch.zeros([1, coords.size(1)] + [s - 1 for s in coords.shape[2:]],
                                               dtype=coords.dtype), name=name)
        b.setTransform(O.CoordsToTransforms(c64).to(coords.dtype),
                       O.CoordsToFaceTransforms(c64).to(coords.dtype) if coords.size(1) > 1 else None)
        return b
    return d.CreateBlock(vertexCoordinates=coords, name=name)


def _rand(shape, gen, dtype, device, scale=1.0):
    return (torch.rand(shape, generator=gen, dtype=torch.float64) * 2 - 1).mul(scale).to(dtype=dtype,
                                                                                     device=device).contiguous()


def stretched_coords(nx, ny, dtype, device, ratio=1.1, x0=0.0, y0=0.0):
    """Orthogonal, non-uniform 2D grid: geometric spacing in x and y."""
    def axis(n, r, o):
        w = torch.tensor([r ** i for i in range(n)], dtype=torch.float64)
        return torch.cat([torch.zeros(1, dtype=torch.float64), torch.cumsum(w / w.sum(), 0)]) + o
    xs, ys = axis(nx, ratio, x0), axis(ny, 1.0 / ratio, y0)
    gy, gx = torch.meshgrid(ys, xs, indexing="ij")
    return torch.stack([gx, gy])[None].to(dtype=dtype, device=device).contiguous()


def cavity(dtype=torch.float64, device="cuda", n=8, coords=True, seed=0, random_init=True):
    d = _domain(2, 0.05, dtype, device)
    if coords:
        b = _coords_block(d, uniform_coords(n, n, dtype=dtype, device=device))
    else:
        b = d.CreateBlockWithSize(P.Int4(n, n, 1))
    b.CloseAllBoundaries()
    b.CloseBoundary("+y", velocity=torch.tensor([[1.0, 0.0]], dtype=dtype, device=device))
    if random_init:
        g = torch.Generator().manual_seed(seed)
        b.setVelocity(_rand(b.velocity.shape, g, dtype, device, 0.3))
    d.PrepareSolve()
    return d


def periodic_two_block(dtype=torch.float64, device="cuda", nx=10, ny=6, coords=False, seed=1):
    """2 blocks connected on both y faces, periodic in x (the 2-block periodic box)."""
    d = _domain(2, 0.02, dtype, device)
    g = torch.Generator().manual_seed(seed)
    bs = []
    for i in range(2):
        if coords:
            b = _coords_block(d, uniform_coords(nx, ny, lx=1.0, ly=ny / nx, y0=i * ny / nx, dtype=dtype,
                                                device=device))
        else:
            b = d.CreateBlockWithSize(P.Int4(nx, ny, 1))
        b.MakePeriodic("x")
        b.setVelocity(_rand(b.velocity.shape, g, dtype, device, 0.5))
        bs.append(b)
    bs[0].ConnectBlock("+y", bs[1], "-y", "-x")
    bs[1].ConnectBlock("+y", bs[0], "-y", "-x")
    d.PrepareSolve()
    return d


def rotated_connection(dtype=torch.float64, device="cuda", seed=2):
    """Two blocks of different sizes joined with permuted and inverted axes: block0 +x <-> block1 -y, with block0's y
    axis running against block1's x axis; stretched orthogonal metrics; a moving wall, a varying wall velocity, a
    per-cell viscosity and a velocity source."""
    d = _domain(2, 0.03, dtype, device)
    g = torch.Generator().manual_seed(seed)
    b0 = _coords_block(d, stretched_coords(7, 5, dtype, device, 1.15))
    b1 = _coords_block(d, stretched_coords(5, 6, dtype, device, 0.9, x0=2.0))
    b0.CloseAllBoundaries()
    b1.CloseAllBoundaries()
    b0.ConnectBlock("+x", b1, "-y", "+x")  # inverted axis
    b0.CloseBoundary("-y", velocity=torch.tensor([[0.7, 0.0]], dtype=dtype, device=device))
    vvar = _rand([1, 2, 1, 5], g, dtype, device, 0.5)
    vvar[:, 1] = 0  # tangential wall motion: zero net boundary flux (compatible pressure system)
    b1.CloseBoundary("+y", velocity=vvar)
    for b in (b0, b1):
        b.setVelocity(_rand(b.velocity.shape, g, dtype, device, 0.4))
        b.setPressure(_rand(b.pressure.shape, g, dtype, device, 0.2))
    b0.setViscosity((torch.rand([1, 1, 5, 7], generator=g, dtype=torch.float64) * 0.05 + 0.01).to(dtype=dtype,
                                                                                                   device=device))
    b1.setVelocitySource(_rand([1, 2], g, dtype, device, 0.3))
    d.PrepareSolve()
    return d


def box3d(dtype=torch.float64, device="cuda", seed=3, size=(5, 4, 6)):
    """3D block: periodic in z, closed in x and y with a moving +y wall, a varying velocity source, transforms."""
    d = _domain(3, 0.04, dtype, device)
    g = torch.Generator().manual_seed(seed)
    nx, ny, nz = size
    b = _coords_block(d, uniform_coords(nx, ny, nz, lx=1.0, ly=0.8, lz=1.2, dtype=dtype, device=device))
    b.CloseAllBoundaries()
    b.MakePeriodic("z")
    b.CloseBoundary("+y", velocity=torch.tensor([[0.5, 0.0, 0.2]], dtype=dtype, device=device))
    b.setVelocity(_rand(b.velocity.shape, g, dtype, device, 0.3))
    b.setVelocitySource(_rand(b.velocity.shape, g, dtype, device, 0.1))
    d.PrepareSolve()
    return d


def three_d_connected(dtype=torch.float64, device="cuda", seed=4):
    """Two 3D blocks joined x-face to z-face with swapped tangential axes."""
    d = _domain(3, 0.05, dtype, device)
    g = torch.Generator().manual_seed(seed)
    b0 = d.CreateBlockWithSize(P.Int4(4, 3, 5))
    b1 = d.CreateBlockWithSize(P.Int4(5, 3, 4))
    b0.CloseAllBoundaries()
    b1.CloseAllBoundaries()
    # b0 +x face (y:3, z:5) <-> b1 -z face (x:5, y:3); b0's y aligns to b1's y, b0's z to b1's x (inverted)
    b0.ConnectBlock("+x", b1, "-z", "-y", "+x")
    for b in (b0, b1):
        b.setVelocity(_rand(b.velocity.shape, g, dtype, device, 0.3))
    d.PrepareSolve()
    return d


def cavity_static_fields(dtype=torch.float64, device="cuda", n=6, seed=5):
    """Cavity with a static block viscosity, a static velocity source and a varying lid velocity."""
    d = _domain(2, 0.05, dtype, device)
    b = _coords_block(d, uniform_coords(n, n + 1, dtype=dtype, device=device))
    g = torch.Generator().manual_seed(seed)
    b.CloseAllBoundaries()
    lid = _rand([1, 2, 1, n], g, dtype, device, 0.5)
    lid[:, 1] = 0
    b.CloseBoundary("+y", velocity=lid)
    b.setVelocity(_rand(b.velocity.shape, g, dtype, device, 0.3))
    b.setPressure(_rand(b.pressure.shape, g, dtype, device, 0.2))
    b.setViscosity(torch.tensor([[0.07]], dtype=dtype, device=device))
    b.setVelocitySource(_rand([1, 2], g, dtype, device, 0.2))
    d.PrepareSolve()
    return d


def outflow_channel(dtype=torch.float64, device="cuda", coords=False, seed=6, nx=10, ny=8):
    """Two blocks in a row (block0 +x joined to block1 -x), walls on +-y, inflow u = 1 on block0 -x and an
    advective outflow on block1 +x (varying, initially u = 1).  With coords the grid is stretched in x and y, so
    the outflow boundary carries a non-uniform transform."""
    d = _domain(2, 0.02, dtype, device)
    g = torch.Generator().manual_seed(seed)
    bs = []
    for i in range(2):
        if coords:
            b = _coords_block(d, stretched_coords(nx, ny, dtype, device, ratio=1.08, x0=float(i)), "B%d" % i)
        else:
            b = d.CreateBlockWithSize(P.Int4(nx, ny, 1), name="B%d" % i)
        vel = _rand(b.velocity.shape, g, dtype, device, 0.2)
        vel[:, 0] += 1.0
        b.setVelocity(vel.contiguous())
        b.CloseAllBoundaries()
        bs.append(b)
    bs[0].ConnectBlock("+x", bs[1], "-x", "+y")
    bs[0].CloseBoundary("-x", velocity=torch.tensor([[1.0, 0.0]], dtype=dtype, device=device))
    bs[1].CloseBoundary("+x", velocity=torch.tensor([[1.0, 0.0]], dtype=dtype, device=device))
    bs[1].getBoundary("+x").makeVelocityVarying()
    d.PrepareSolve()
    return d


ALL = {
    "cavity": cavity,
    "cavity_nocoords": lambda **kw: cavity(coords=False, **kw),
    "periodic_two_block": periodic_two_block,
    "periodic_two_block_coords": lambda **kw: periodic_two_block(coords=True, **kw),
    "rotated_connection": rotated_connection,
    "box3d": box3d,
    "three_d_connected": three_d_connected,
    "cavity_static_fields": cavity_static_fields,
}


def sheared_coords(size, shear, h=0.1, dtype=torch.float64, device="cpu", bend=0.0):
    """Vertex coordinates of a non-orthogonal grid.  2-D: x = h (i + s0 j), y = h j; 3-D: x = h (i + s0 j + s1 k),
    y = h (j + s2 k), z = h k (an affine map of the unit lattice: every cell has the same metrics).  `bend` adds
    a smooth curvature, y += bend h sin(pi i / nx), so the metrics vary from cell to cell."""
    vs = [n + 1 for n in size]
    axes = [torch.arange(n, dtype=torch.float64) for n in vs]
    if len(size) == 2:
        j, i = torch.meshgrid(axes[1], axes[0], indexing="ij")
        x = h * (i + shear[0] * j)
        y = h * j + bend * h * torch.sin(math.pi * i / size[0])
        c = torch.stack([x, y])
    else:
        k, j, i = torch.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        x = h * (i + shear[0] * j + shear[1] * k)
        y = h * (j + shear[2] * k) + bend * h * torch.sin(math.pi * i / size[0])
        z = h * k
        c = torch.stack([x, y, z])
    return c[None].to(dtype=dtype, device=device).contiguous()


def sheared_box(dtype=torch.float64, device="cuda", size=(9, 8), shear=(0.4,), periodic=(), walls_moving=True,
                bend=0.0, seed=7, visc=0.04):
    """One block on a sheared grid: closed walls on every face not listed in `periodic` (axis names), the +y
    wall moving tangentially; random velocity and pressure."""
    dims = len(size)
    d = _domain(dims, visc, dtype, device)
    g = torch.Generator().manual_seed(seed)
    b = _coords_block(d, sheared_coords(size, shear, dtype=dtype, device=device, bend=bend))
    b.CloseAllBoundaries()
    for ax in periodic:
        b.MakePeriodic(ax)
    if walls_moving and "y" not in periodic:
        v = [0.0] * dims
        v[0] = 0.6
        b.CloseBoundary("+y", velocity=torch.tensor([v], dtype=dtype, device=device))
    b.setVelocity(_rand(b.velocity.shape, g, dtype, device, 0.4))
    b.setPressure(_rand(b.pressure.shape, g, dtype, device, 0.2))
    return _prepare(d)


def _prepare(d):
    d.PrepareSolve()
    if str(d.getDevice()) == "cpu":
        from oracle import oracle as O
        O.build_csr_pattern(d)
    return d


def sheared_two_block(dtype=torch.float64, device="cuda", seed=8, rotate=True):
    """Two sheared, bent blocks joined along x (block0 +x <-> block1, with block1's axes permuted and inverted
    when `rotate`), walls elsewhere, one moving; random velocity and pressure, a per-cell viscosity."""
    d = _domain(2, 0.03, dtype, device)
    g = torch.Generator().manual_seed(seed)
    b0 = _coords_block(d, sheared_coords((7, 6), (0.3,), dtype=dtype, device=device, bend=0.3))
    if rotate:
        b1 = _coords_block(d, sheared_coords((6, 5), (-0.25,), dtype=dtype, device=device, bend=0.2))
    else:
        b1 = _coords_block(d, sheared_coords((5, 6), (0.3,), dtype=dtype, device=device, bend=0.2))
    b0.CloseAllBoundaries()
    b1.CloseAllBoundaries()
    if rotate:
        b0.ConnectBlock("+x", b1, "-y", "+x")  # inverted axis
    else:
        b0.ConnectBlock("+x", b1, "-x", "+y")
    b0.CloseBoundary("-y", velocity=torch.tensor([[0.7, 0.0]], dtype=dtype, device=device))
    for b in (b0, b1):
        b.setVelocity(_rand(b.velocity.shape, g, dtype, device, 0.4))
        b.setPressure(_rand(b.pressure.shape, g, dtype, device, 0.2))
    b0.setViscosity((torch.rand([1, 1, 6, 7], generator=g, dtype=torch.float64) * 0.05 + 0.01).to(dtype=dtype,
                                                                                                   device=device))
    return _prepare(d)


NON_ORTHO = {
    "sheared_walls": lambda **kw: sheared_box(bend=0.3, **kw),
    "sheared_periodic_x": lambda **kw: sheared_box(periodic=("x",), **kw),
    "sheared_two_block_rotated": sheared_two_block,
    "sheared_two_block": lambda **kw: sheared_two_block(rotate=False, **kw),
    "sheared_box3d": lambda **kw: sheared_box(size=(5, 4, 6), shear=(0.3, -0.2, 0.25), periodic=("z",), bend=0.2,
                                              **kw),
}


def with_scalars(build, channels=2, seed=41, viscosity="per_channel", neumann=True, **kw):
    """`build(**kw)` with `channels` passive-scalar channels: random block scalars; a scalar viscosity per channel
    (`viscosity` "per_channel"), one for all ("static") or none (the domain viscosity, None); random boundary
    scalars on every FIXED face, per cell on the first, static on the others; with `neumann`, channel 1 is Neumann
    on every second FIXED face (per-channel types)."""
    _SCALAR_CHANNELS[0] = channels
    try:
        d = build(**kw)
    finally:
        _SCALAR_CHANNELS[0] = 0
    g = torch.Generator().manual_seed(seed)
    dtype, device = d.getDtype(), d.getDevice()
    dims = d.getSpatialDims()
    for b in d.blocks:
        b.setPassiveScalar(_rand(b.passiveScalar.shape, g, dtype, device, 1.0))
    if viscosity == "per_channel":
        d.setScalarViscosity(torch.tensor([0.02 + 0.03 * c for c in range(channels)], dtype=dtype, device=device))
    elif viscosity == "static":
        d.setScalarViscosity(torch.tensor([0.035], dtype=dtype, device=device))
    k = 0
    for b in d.blocks:
        for _, fb in b.getFixedBoundaries():
            sz = fb.getSizes()
            shape = [1, channels] + [sz[a] for a in reversed(range(dims))] if k == 0 else [1, channels]
            fb.setPassiveScalar(_rand(shape, g, dtype, device, 0.8))
            if neumann and channels > 1 and k % 2 == 1:
                types = [P.BoundaryConditionType.DIRICHLET] * channels
                types[1] = P.BoundaryConditionType.NEUMANN
                fb.setPassiveScalarType(types)
            k += 1
    if str(device) != "cpu":
        d.UpdateDomainData()
    return d
