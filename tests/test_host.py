"""Host-side object model (PISOtorch.Domain/Block/boundaries): the behaviour of the reference's
extensions/domain_structs.cpp that the PISO step depends on, checked without a GPU."""
import pytest
import torch

from pict_amd import PISOtorch as P


def _dom(dims=2, dtype=torch.float64):
    return P.Domain(dims, torch.tensor([0.1], dtype=dtype), "t", dtype, "cpu", passiveScalarChannels=0,
                    _allow_cpu=True)


def test_default_boundaries_are_periodic_and_csr_size():
    d = _dom()
    b = d.CreateBlockWithSize(P.Int4(5, 4, 1))
    assert all(bd.type == P.BoundaryType.PERIODIC for bd in b.getBoundaries())
    assert b.ComputeCSRSize() == 20 * 5
    b.CloseAllBoundaries()
    # domain_structs.cpp:2151-2161: minus one entry per cell face on a closed boundary
    assert b.ComputeCSRSize() == 20 * 5 - 2 * 4 - 2 * 5
    assert all(bd.type == P.BoundaryType.FIXED for bd in b.getBoundaries())
    assert b.getBoundary("+y").getSizes()[1] == 1 and b.getBoundary("+y").getSizes()[0] == 5


def test_connect_block_axes_2d():
    """ConnectBlocks (domain_structs.cpp:1064-1097): axes of both sides, with inversion bits."""
    d = _dom()
    b0 = d.CreateBlockWithSize(P.Int4(6, 4, 1))
    b1 = d.CreateBlockWithSize(P.Int4(4, 5, 1))
    b0.ConnectBlock("+x", b1, "-y", "+x")
    c0, c1 = b0.getBoundary("+x"), b1.getBoundary("-y")
    assert c0.type == P.BoundaryType.CONNECTED and c1.type == P.BoundaryType.CONNECTED
    assert c0.axes == [2, 1] and c1.axes == [1, 3]
    assert c0.getConnectedBlock() is b1 and c1.getConnectedBlock() is b0
    # the faces on the other side of previous periodic boundaries were closed
    assert b0.getBoundary("-x").type == P.BoundaryType.FIXED
    assert b1.getBoundary("+y").type == P.BoundaryType.FIXED


def test_connect_block_axes_3d_swapped():
    d = _dom(3)
    b0 = d.CreateBlockWithSize(P.Int4(4, 3, 5))
    b1 = d.CreateBlockWithSize(P.Int4(5, 3, 4))
    b0.ConnectBlock("+x", b1, "-z", "-y", "+x")
    assert b0.getBoundary("+x").axes == [4, 2, 1]
    assert b1.getBoundary("-z").axes == [1, 5, 2]


def test_connect_size_mismatch_raises():
    d = _dom()
    b0 = d.CreateBlockWithSize(P.Int4(6, 4, 1))
    b1 = d.CreateBlockWithSize(P.Int4(5, 5, 1))
    with pytest.raises(RuntimeError):
        b0.ConnectBlock("+x", b1, "-y", "+x")


def test_prepare_solve_offsets_and_validation():
    d = _dom()
    b0 = d.CreateBlockWithSize(P.Int4(6, 4, 1))
    b1 = d.CreateBlockWithSize(P.Int4(6, 3, 1))
    b0.ConnectBlock("+y", b1, "-y", "-x")
    d.PrepareSolve()
    assert b0.globalOffset == 0 and b1.globalOffset == 24
    assert b1.csrOffset == b0.ComputeCSRSize()
    assert d.getTotalSize() == 42
    assert d.C.getSize() == b0.ComputeCSRSize() + b1.ComputeCSRSize() and d.C.getRows() == 42
    assert d.P.index is d.C.index
    assert d.velocityResult.shape == (84,) and d.pressureRHSdiv.shape == (42,)
    # a one-sided periodic boundary is rejected (domain_structs.cpp:2630-2633)
    d2 = _dom()
    b = d2.CreateBlockWithSize(P.Int4(4, 4, 1))
    b.CloseBoundary("-x")
    b.boundaries[1] = P.PeriodicBoundary(d2)
    with pytest.raises(RuntimeError):
        d2.PrepareSolve()


def test_tensor_changed_tracking():
    d = _dom()
    b = d.CreateBlockWithSize(P.Int4(4, 4, 1))
    d.PrepareSolve()
    assert not d.IsTensorChanged()
    d.setVelocityResult(torch.zeros_like(d.velocityResult))
    assert d.IsTensorChanged()
    d.UpdateDomainData()
    assert not d.IsTensorChanged()
    b.setVelocity(torch.ones_like(b.velocity))
    assert d.IsTensorChanged()


def test_field_validation_messages():
    d = _dom()
    b = d.CreateBlockWithSize(P.Int4(4, 4, 1))
    with pytest.raises(RuntimeError):
        b.setVelocity(torch.zeros(1, 2, 4, 5, dtype=torch.float64))
    with pytest.raises(RuntimeError):
        b.setVelocity(torch.zeros(1, 2, 4, 4, dtype=torch.float32))
    with pytest.raises(RuntimeError):
        d.CreateBlockWithSize(P.Int4(2, 4, 1))
    with pytest.raises(RuntimeError):
        P.Domain(4, torch.tensor([0.1]), "x", torch.float32, "cpu", _allow_cpu=True)


def test_flux_balance():
    d = _dom()
    b = d.CreateBlockWithSize(P.Int4(4, 4, 1))
    b.CloseAllBoundaries()
    b.CloseBoundary("-x", velocity=torch.tensor([[1.0, 0.0]], dtype=torch.float64))
    d.PrepareSolve()
    # inflow on -x without outflow: the balance is -(sum of the inflow flux) (domain_structs.cpp:2460-2493)
    assert d.GetBoundaryFluxBalance().item() == pytest.approx(-4.0)
    assert not d.CheckBoundaryFluxBalance()
    b.CloseBoundary("+x", velocity=torch.tensor([[1.0, 0.0]], dtype=torch.float64))
    d.PrepareSolve()
    assert d.CheckBoundaryFluxBalance()


def test_csr_matrix_helpers():
    m = P.CSRmatrix(5, 3, torch.float64, "cpu")
    assert m.getSize() == 5 and m.getRows() == 3
    z = m.WithZeroValue()
    assert z.index is m.index and z.row is m.row and z.value is not m.value
    c = m.clone()
    assert c.index is not m.index
    with pytest.raises(RuntimeError):
        m.setValue(torch.zeros(4, dtype=torch.float64))
