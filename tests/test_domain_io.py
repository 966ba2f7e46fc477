"""Domain save/load (the reference's lib/util/domain_io.py:44-236 file format): a saved and reloaded domain has the
same structure and tensors, steps identically (oracle on the CPU, the library on the GPU), and the JSON carries the
reference's keys.  Shared tensors are stored once."""
import json
import os

import numpy as np
import pytest
import torch

import scenarios
from oracle import oracle as O
from pict_amd import PISOtorch as P
from pict_amd import domain_io

CASES = ["cavity", "periodic_two_block", "rotated_connection", "box3d", "outflow_channel", "sheared_two_block_rotated"]
BUILDERS = dict(scenarios.ALL, **scenarios.NON_ORTHO, outflow_channel=scenarios.outflow_channel)


def _same(a, b):
    if a is None or b is None:
        return a is None and b is None
    return a.shape == b.shape and torch.equal(a.detach().cpu(), b.detach().cpu())


def _check_same_domain(d, e):
    assert e.name == d.name and e.getSpatialDims() == d.getSpatialDims()
    assert e.getPassiveScalarChannels() == d.getPassiveScalarChannels()
    assert _same(e.viscosity, d.viscosity) and _same(e.passiveScalarViscosity, d.passiveScalarViscosity)
    assert len(e.blocks) == len(d.blocks)
    for bd, be in zip(d.blocks, e.blocks):
        assert be.name == bd.name
        for f in ("velocity", "pressure", "passiveScalar", "velocitySource", "viscosity", "transform",
                  "faceTransform"):
            assert _same(getattr(be, f), getattr(bd, f)), f
        for x, y in zip(bd.boundaries, be.boundaries):
            assert x.type == y.type
            if x.type == P.BoundaryType.FIXED:
                assert _same(y.velocity, x.velocity) and _same(y.passiveScalar, x.passiveScalar)
                assert list(y.passiveScalarTypes or []) == list(x.passiveScalarTypes or [])
            elif x.type == P.BoundaryType.CONNECTED:
                assert list(y.axes) == list(x.axes)
                assert d.blocks.index(x.getConnectedBlock()) == e.blocks.index(y.getConnectedBlock())


@pytest.mark.parametrize("name", CASES)
def test_round_trip_cpu(name, tmp_path):
    d = scenarios.with_scalars(BUILDERS[name], viscosity="per_channel", dtype=torch.float64, device="cpu")
    path = str(tmp_path / "dom")
    domain_io.save_domain(d, path)
    e = domain_io.load_domain(path, device="cpu")
    _check_same_domain(d, e)
    scenarios._prepare(e)
    O.build_csr_pattern(d)
    for dom in (d, e):
        O.piso_step(dom, 0.05, advection_tol=1e-12, pressure_tol=1e-12)
    for bd, be in zip(d.blocks, e.blocks):
        assert torch.equal(be.velocity, bd.velocity) and torch.equal(be.passiveScalar, bd.passiveScalar)


def test_file_format(tmp_path):
    d = scenarios.with_scalars(BUILDERS["rotated_connection"], viscosity="static", dtype=torch.float32, device="cpu")
    path = str(tmp_path / "dom")
    domain_io.save_domain(d, path)
    with open(path + ".json") as fh:
        j = json.load(fh)
    assert set(j) >= {"name", "spatialDims", "viscosity", "passiveScalarChannels", "passiveScalarViscosity", "blocks",
                      "data_info"}
    kinds = {b["type"] for blk in j["blocks"] for b in blk["boundaries"]}
    assert kinds <= {"FIXED", "CONNECTED", "PERIODIC"} and "CONNECTED" in kinds
    fixed = [b for blk in j["blocks"] for b in blk["boundaries"] if b["type"] == "FIXED"]
    assert all(set(b) >= {"velocityType", "velocity", "passiveScalarType", "scalar"} for b in fixed)
    with np.load(path + ".npz") as z:
        assert sorted(z.files, key=int) == [str(i) for i in range(len(j["data_info"]))]
        for k, info in j["data_info"].items():
            assert list(z[k].shape) == info["shape"] and info["dtype"] == "float32"
    # one tensor shared by two blocks is stored once
    shared = torch.ones(1, 1, dtype=torch.float32) * 0.01
    d.blocks[0].setViscosity(shared)
    d.blocks[1].setViscosity(shared)
    domain_io.save_domain(d, path)
    with open(path + ".json") as fh:
        j = json.load(fh)
    assert j["blocks"][0]["viscosity"] == j["blocks"][1]["viscosity"]
    e = domain_io.load_domain(path, device="cpu", with_scalar=False)
    assert not e.hasPassiveScalar() and e.blocks[0].passiveScalar is None


@pytest.mark.gpu
def test_loaded_domain_steps_bitwise_like_the_original(tmp_path):
    from pict_amd import PISOtorch_simulation as S
    d = scenarios.with_scalars(BUILDERS["outflow_channel"], viscosity="per_channel", dtype=torch.float32,
                               device="cuda")
    path = str(tmp_path / "dom")
    domain_io.save_domain(d, path)
    e = domain_io.load_domain(path)
    assert e.getDevice().type == "cuda"
    e.PrepareSolve()
    for dom in (d, e):
        sim = S.Simulation(dom, time_step=0.05, corrector_steps=2, non_orthogonal=False, advection_tol=1e-6,
                           pressure_tol=1e-6)
        assert sim.run(3)
    torch.cuda.synchronize()
    for bd, be in zip(d.blocks, e.blocks):
        for f in ("velocity", "pressure", "passiveScalar"):
            assert torch.equal(getattr(be, f), getattr(bd, f)), f
    os.remove(path + ".npz")
