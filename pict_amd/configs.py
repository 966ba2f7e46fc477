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
