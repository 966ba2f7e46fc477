"""PISOtorch -- the Python surface of PICT's compiled `PISOtorch` module, on libpict_hip.

Mirrors the classes, methods, argument names, defaults and error behaviour of
the reference binding (tum-pbs/PICT extensions/PISOtorch.cpp:25-655) and its
object model (extensions/domain_structs.h:57-787, domain_structs.cpp), so that a
PICT-style driver (`PISOtorch_simulation`, `PISOtorch_diff`) runs unchanged.

All numerics run in hand-written HIP kernels (libpict_hip.so, C ABI in
include/pict_hip.h).  This module only owns tensors and packs the descriptor
the kernels read; there is no CPU fallback: a missing library or a non-GPU
domain raises.

Differences from the reference visible from Python:
  * Only FixedBoundary / ConnectedBoundary / PeriodicBoundary exist (the
    reference's Static/VaryingDirichletBoundary are refused by its own
    MakeClosedBoundary, domain_structs.cpp:1866-1868).
  * useFaceTransform=True in the pressure setup raises NotImplementedError
    (PictUnsupportedError); the reference's driver never sets it.
  * The ILU(0)-preconditioned BiCGStab applies M = L U (the factors' roles
    bicgstab_solver_kernel.cu:73-77 declares); a missing diagonal entry is
    an error.
  * Per-channel passive-scalar boundary types are limited to 32 channels
    (one type for all channels has no limit).
"""
import ctypes
import enum
import math

import torch

from . import _abi
from ._abi import PictError, PictUnsupportedError  # noqa: F401  (re-exported)

_INDEX = torch.int32


# ----------------------------------------------------------------------------
# enums and small types (PISOtorch.cpp:29-58, :544-550, :637-640)
# ----------------------------------------------------------------------------

class BoundaryType(enum.IntEnum):
    DIRICHLET = 0
    DIRICHLET_VARYING = 1
    NEUMANN = 10
    FIXED = 30
    CONNECTED = 20
    PERIODIC = 21


class BoundaryConditionType(enum.IntEnum):
    DIRICHLET = 0
    NEUMANN = 1


class ConvergenceCriterion(enum.IntEnum):
    NORM2 = 0
    NORM2_NORMALIZED = 1
    ABS_SUM = 2
    ABS_MEAN = 3
    ABS_MAX = 4


DIRICHLET = BoundaryType.DIRICHLET
DIRICHLET_VARYING = BoundaryType.DIRICHLET_VARYING
NEUMANN = BoundaryType.NEUMANN
FIXED = BoundaryType.FIXED
CONNECTED = BoundaryType.CONNECTED
PERIODIC = BoundaryType.PERIODIC
NORM2 = ConvergenceCriterion.NORM2
NORM2_NORMALIZED = ConvergenceCriterion.NORM2_NORMALIZED
ABS_SUM = ConvergenceCriterion.ABS_SUM
ABS_MEAN = ConvergenceCriterion.ABS_MEAN
ABS_MAX = ConvergenceCriterion.ABS_MAX


class Int4:
    """PISOtorch.Int4 (PISOtorch.cpp:29-43)."""

    def __init__(self, x=0, y=0, z=0, w=0):
        self._a = (int(x), int(y), int(z), int(w))

    x = property(lambda self: self._a[0])
    y = property(lambda self: self._a[1])
    z = property(lambda self: self._a[2])
    w = property(lambda self: self._a[3])

    def __getitem__(self, i):
        if 0 <= i < 4:
            return self._a[i]
        raise IndexError(i)

    def __len__(self):
        return 4

    def __str__(self):
        return "I4(%d,%d,%d,%d)" % self._a

    __repr__ = __str__


class LinearSolverResultInfo:
    """bicgstab_solver.h:20-25."""

    def __init__(self, finalResidual, usedIterations, converged, isFiniteResidual):
        self.finalResidual = float(finalResidual)
        self.usedIterations = int(usedIterations)
        self.converged = bool(converged)
        self.isFiniteResidual = bool(isFiniteResidual)

    def __str__(self):
        return ("LinearSolverResultInfo(finalResidual=%g, usedIterations=%d, converged=%d, isFiniteResidual=%d)"
                % (self.finalResidual, self.usedIterations, self.converged, self.isFiniteResidual))

    __repr__ = __str__


_FACES = {"-x": 0, "+x": 1, "-y": 2, "+y": 3, "-z": 4, "+z": 5}
_AXES = {"x": 0, "y": 1, "z": 2}


def _face_index(face):
    if isinstance(face, str):
        if face not in _FACES:
            raise RuntimeError("Invalid boundary location specified: %s" % face)
        return _FACES[face]
    return int(face)


def _axis_index(axis):
    if isinstance(axis, str):
        return _AXES.get(axis, -1)
    return int(axis)


def _check(cond, msg):
    if not cond:
        raise RuntimeError(msg)


def _ptr(t):
    return 0 if t is None else t.data_ptr()


def _dtype_code(dtype):
    if dtype == torch.float32:
        return _abi.PICT_DTYPE_F32
    if dtype == torch.float64:
        return _abi.PICT_DTYPE_F64
    raise RuntimeError("dtype must be float32 or float64, got %s" % dtype)


def _stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _spatial_size_xyz(t, dims, channels_first=True):
    """Spatial sizes (x,y,z) of an NC<spatial> (or N<spatial>C) tensor."""
    off = 1 if channels_first else 0
    size = [1, 1, 1]
    for d in range(dims):
        size[d] = t.size(dims + off - d)
    return size


# ----------------------------------------------------------------------------
# CSRmatrix (domain_structs.cpp:176-243)
# ----------------------------------------------------------------------------

class CSRmatrix:
    def __init__(self, *args):
        if len(args) == 3 and all(isinstance(a, torch.Tensor) for a in args):
            value, index, row = args
            _check(value.dim() == 1, "value must be 1D.")
            _check(index.dim() == 1, "index must be 1D.")
            _check(row.dim() == 1, "row must be 1D.")
            _check(value.size(0) == index.size(0), "value and index must have same size.")
            _check(index.dtype == _INDEX, "index must have integer type.")
            _check(row.dtype == _INDEX, "row must have integer type.")
            _check(index.device == value.device and row.device == value.device,
                   "all tensors must be on the same device.")
            self._value, self._index, self._row = value, index, row
        elif len(args) == 4:
            num_values, num_rows, dtype, device = args
            self._value = torch.zeros(int(num_values), dtype=dtype, device=device)
            self._index = torch.zeros(int(num_values), dtype=_INDEX, device=device)
            self._row = torch.zeros(int(num_rows) + 1, dtype=_INDEX, device=device)
        else:
            raise TypeError("CSRmatrix(value, index, row) or CSRmatrix(numValues, numRows, dtype, device)")
        self._changed = False

    value = property(lambda self: self._value)
    index = property(lambda self: self._index)
    row = property(lambda self: self._row)

    def setValue(self, tensor):
        _check(tensor.dim() == 1, "value must be 1D.")
        _check(self._value.size(0) == tensor.size(0), "value must have the correct size.")
        _check(self._value.dtype == tensor.dtype, "value must have the correct type.")
        _check(self._value.device == tensor.device, "all tensors must be on the same device.")
        self._value = tensor
        self._changed = True

    def CreateValue(self):
        self._value = torch.zeros_like(self._value)
        self._changed = True

    def getDevice(self):
        return self._value.device

    def getDtype(self):
        return self._value.dtype

    def copy(self):
        return CSRmatrix(self._value, self._index, self._row)

    def clone(self):
        return CSRmatrix(self._value.clone(), self._index.clone(), self._row.clone())

    def detach(self):
        self._value = self._value.detach()

    def WithZeroValue(self):
        return CSRmatrix(torch.zeros_like(self._value), self._index, self._row)

    def toType(self, dtype):
        return CSRmatrix(self._value.to(dtype), self._index.clone(), self._row.clone())

    def getSize(self):
        return self._value.size(0)

    def getRows(self):
        return self._row.size(0) - 1

    def IsTensorChanged(self):
        return self._changed

    def __str__(self):
        return "CSRmatrix( size=%d, rows=%d )" % (self.getSize(), self.getRows())

    def _desc(self):
        d = _abi.CsrDesc()
        d.n = self.getRows()
        d.nnz = self.getSize()
        d.dtype = _dtype_code(self._value.dtype)
        d.row = self._row.data_ptr()
        d.index = self._index.data_ptr()
        d.value = self._value.data_ptr()
        perm = getattr(self._index, "_pict_transpose_perm", None)
        d.transpose_perm = perm.data_ptr() if perm is not None else 0
        return d

    def _ensure_transpose_perm(self):
        """The A^T map of the pattern, built once and kept with the (immutable, shared) index tensor."""
        if getattr(self._index, "_pict_transpose_perm", None) is None and self._index.is_cuda:
            perm = torch.empty_like(self._index)
            d = self._desc()
            _abi.check(_abi.load_library().pict_csr_transpose_permutation(
                ctypes.byref(d), ctypes.c_void_p(perm.data_ptr()), _stream(self._index.device)), "transpose")
            self._index._pict_transpose_perm = perm


# ----------------------------------------------------------------------------
# boundaries (domain_structs.h:85-385)
# ----------------------------------------------------------------------------

class Boundary:
    def __init__(self, btype, domain):
        self._type = btype
        self._domain = domain
        self._changed = True

    type = property(lambda self: self._type)

    def getParentDomain(self):
        return self._domain

    def IsTensorChanged(self):
        return self._changed

    def getSpatialDims(self):
        return self._domain.getSpatialDims()

    def getDevice(self):
        return self._domain.getDevice()

    def __str__(self):
        return self._type.name


class PeriodicBoundary(Boundary):
    def __init__(self, domain):
        super().__init__(BoundaryType.PERIODIC, domain)

    def Copy(self):
        return PeriodicBoundary(self._domain)

    Clone = Copy


class ConnectedBoundary(Boundary):
    """ConnectedBoundary (domain_structs.cpp:995-1048).  axes[0] is the connected
    face of the other block, axes[1..] the other block's axes (with inversion bit)
    aligned to this block's axes (face/2+1)%D, (face/2+2)%D."""

    def __init__(self, connected_block, axes, domain):
        super().__init__(BoundaryType.CONNECTED, domain)
        dims = domain.getSpatialDims()
        _check(len(axes) == dims, "axes must match spatial dimensions of connectedBlock.")
        _check(0 <= axes[0] < 2 * dims, "axes[0] must be in [0,dim*2].")
        if dims > 1:
            _check(0 <= axes[1] < 2 * dims, "axes[1] must be in [0,dim*2].")
            _check(axes[0] >> 1 != axes[1] >> 1, "axes[1] must be a different axis than axes[0].")
            if dims > 2:
                _check(0 <= axes[2] < 2 * dims, "axes[2] must be in [0,dim*2].")
                _check(axes[0] >> 1 != axes[2] >> 1, "axes[2] must be a different axis than axes[0].")
                _check(axes[1] >> 1 != axes[2] >> 1, "axes[2] must be a different axis than axes[1].")
        self._block = connected_block
        self.axes = list(int(a) for a in axes)

    def getConnectedBlock(self):
        return self._block

    def getConnectedFace(self):
        return self.axes[0] >> 1

    def getConnectionAxis(self, axis):
        return self.axes[axis] >> 1

    def getConnectionAxisDirection(self, axis):
        return self.axes[axis] & 1

    def Copy(self):
        raise RuntimeError("ConnectedBoundary can't be copied")

    Clone = Copy

    def __str__(self):
        names = ["-x", "+x", "-y", "+y", "-z", "+z"]
        s = 'CONNECTED(to="%s", face=%s' % (self._block.name, names[self.axes[0]])
        for d in range(1, len(self.axes)):
            s += ", axis%d=%s" % (d, names[self.axes[d]])
        return s + ")"


class FixedBoundary(Boundary):
    """FixedBoundary (domain_structs.cpp:305-795): prescribed (Dirichlet) velocity,
    static [1,D] or varying [1,D,(z),y,x] with the face axis of size 1, and an
    optional face-transform slice [1,(z),y,x,2D^2+1]."""

    def __init__(self, velocity, velocityType, passiveScalar, passiveScalarType, transform, domain):
        super().__init__(BoundaryType.FIXED, domain)
        self._size = None
        dims = domain.getSpatialDims()
        if velocity is not None and velocity.dim() > 2:
            self._setSizeFromTensor(velocity, True)
        elif passiveScalar is not None and passiveScalar.dim() > 2:
            self._setSizeFromTensor(passiveScalar, True)
        elif transform is not None:
            self._setSizeFromTensor(transform, False)
        elif velocity is not None:
            _check(velocity.dim() > 1, "Velocity must be 1-3D and either static (shape NC) or varying (shape NCDHW).")
        elif passiveScalar is not None:
            raise RuntimeError("Can't create FixedBoundary from static passive scalar alone.")
        else:
            raise RuntimeError("Any field or 'size' is required to create FixedBoundary.")
        self._velocity_grad = None
        self._passiveScalar = None
        self._passiveScalar_grad = None
        self._passiveScalarTypes = None
        self._transform = None
        if velocity is not None:
            self.setVelocity(velocity)
        else:
            self.CreateVelocity(True)
        self.setVelocityType(velocityType)
        if passiveScalar is not None:
            self.setPassiveScalar(passiveScalar)
        elif domain.hasPassiveScalar():
            self.CreatePassiveScalar(True)
        if self.hasPassiveScalar():
            self.setPassiveScalarType(passiveScalarType if passiveScalarType is not None else velocityType)
        if transform is not None:
            self.setTransform(transform)
        self._dims = dims

    # --- size ---
    def _setSizeFromTensor(self, t, channels_first):
        dims = self._domain.getSpatialDims()
        _check(t.dim() == dims + 2, "Invalid tensor for boundary shape.")
        s = _spatial_size_xyz(t, dims, channels_first)
        self._size = [s[0], s[1], s[2], dims]

    def _setSize(self, size):
        self._size = [int(size[0]), int(size[1]), int(size[2]), int(size[3])]

    def hasSize(self):
        return self._size is not None

    def getSizes(self):
        s = self._size if self._size is not None else [1, 1, 1, self._domain.getSpatialDims()]
        return Int4(*s)

    def getStrides(self):
        s = self.getSizes()
        return Int4(1, s.x, s.x * s.y, s.x * s.y * s.z)

    # --- velocity ---
    velocity = property(lambda self: self._velocity)
    velocityType = property(lambda self: self._velocityType)
    isVelocityStatic = property(lambda self: self._velocityStatic)
    velocityGrad = property(lambda self: self._velocity_grad if self._velocity_grad is not None else torch.empty(0))

    def setVelocity(self, t):
        dims = self._domain.getSpatialDims()
        self._domain._check_tensor_device(t)
        _check(t.is_contiguous(), "velocity must be contiguous")
        _check(t.size(0) == 1, "Batches are not yet supported (velocity).")
        _check(t.size(1) == dims, "Velocity channels are invalid. Velocity must be %dD and either static (shape NC) "
               "or varying (shape NCDHW)." % dims)
        _check(t.dim() == 2 or t.dim() == 2 + dims, "Velocity spatial dimensions are invalid. Velocity must be %dD "
               "and either static (shape NC) or varying (shape NCDHW)." % dims)
        _check(t.dtype == self._domain.getDtype(), "Velocity dtype must match domain.")
        static = t.dim() == 2
        if not static and self.hasSize():
            _check(_spatial_size_xyz(t, dims) == list(self._size[:3]),
                   "New velocity spatial dimensions must match existing fields.")
        if not static and not self.hasSize():
            self._setSizeFromTensor(t, True)
        if self._velocity_grad is not None and self._velocity_grad.dim() != t.dim():
            raise RuntimeError("New velocity does not match existing gradient tensor.")
        self._velocity = t
        self._velocityStatic = static
        self._changed = True

    def setVelocityType(self, velocityType):
        _check(velocityType == BoundaryConditionType.DIRICHLET,
               "Invalid velocity boundary type: Currently only Dirichlet boundaries are supported")
        self._velocityType = BoundaryConditionType(velocityType)

    def CreateVelocity(self, createStatic):
        opts = dict(dtype=self._domain.getDtype(), device=self._domain.getDevice())
        dims = self._domain.getSpatialDims()
        if createStatic:
            v = torch.zeros(1, dims, **opts)
        else:
            _check(self.hasSize(), "FixedBoundary is missing size to create a varying velocity.")
            v = torch.zeros(_nc_shape(dims, dims, self._size), **opts)
        self.setVelocity(v)

    def setVelocityGrad(self, t):
        _check(t.shape == self._velocity.shape, "FixedBoundary.velocity_grad must match velocity.")
        self._velocity_grad = t
        self._changed = True

    def CreateVelocityGrad(self):
        self.setVelocityGrad(torch.zeros_like(self._velocity))

    def getVelocityVarying(self, size=None):
        if self._velocityStatic:
            _check(self.hasSize() or size is not None,
                   "FixedBoundary has no fields with spatial dimensions to infer varying velocity shape.")
            s = self._size if self.hasSize() else [size[0], size[1], size[2], size[3]]
            dims = self._domain.getSpatialDims()
            v = self._velocity
            for _ in range(dims):
                v = v.unsqueeze(-1)
            return v.expand(_nc_shape(dims, dims, s)).clone().contiguous()
        return self._velocity

    def makeVelocityVarying(self, size=None):
        if self._velocityStatic:
            self.setVelocity(self.getVelocityVarying(size))

    def getVelocity(self, computational):
        _check(self.hasSize(), "FixedBoundary has no size set.")
        v = self.getVelocityVarying(None)
        if computational and self.hasTransform():
            return TransformVectors(v, self._transform, True)
        return v

    def GetFluxes(self):
        """Boundary-normal fluxes (domain_structs.cpp:525-541)."""
        _check(self.hasSize(), "FixedBoundary has no size set.")
        v = self.getVelocityVarying(None)
        if self.hasTransform():
            det = self._transform[..., -1]
            fl = det.unsqueeze(1) * TransformVectors(v, self._transform, True)
        else:
            fl = v
        axis = self._axis()
        return fl[:, axis:axis + 1].clone()

    def _axis(self):
        for d in range(self._domain.getSpatialDims()):
            if self._size[d] == 1:
                return d
        return 0

    # --- passive scalar ---
    passiveScalar = property(lambda self: self._passiveScalar)
    passiveScalarTypes = property(lambda self: self._passiveScalarTypes)
    passiveScalarGrad = property(lambda self: self._passiveScalar_grad)

    def hasPassiveScalar(self):
        return self._passiveScalar is not None

    def getPassiveScalarChannels(self):
        return self._passiveScalar.size(1) if self._passiveScalar is not None else 0

    def isPassiveScalarStatic(self):
        return self._passiveScalar is None or self._passiveScalar.dim() == 2

    def isPassiveScalarBoundaryTypeStatic(self):
        return self._passiveScalarTypes is None or len(self._passiveScalarTypes) < 2

    def setPassiveScalar(self, t):
        dims = self._domain.getSpatialDims()
        self._domain._check_tensor_device(t)
        _check(t.is_contiguous(), "passive scalar must be contiguous")
        _check(t.size(0) == 1, "Batches are not yet supported (passive scalar).")
        _check(t.size(1) == self._domain.getPassiveScalarChannels(), "Passive scalar channels must match domain.")
        _check(t.dim() == 2 or t.dim() == 2 + dims, "Passive scalar must be static (shape NC) or varying (shape "
               "NCDHW).")
        _check(t.dtype == self._domain.getDtype(), "Passive scalar dtype must match domain.")
        if t.dim() > 2:
            s = _spatial_size_xyz(t, dims, True)
            if self._size is None:
                self._size = [s[0], s[1], s[2], dims]
            _check(list(s[:3]) == list(self._size[:3]), "Passive scalar shape must match the boundary's size.")
        self._passiveScalar = t
        self._changed = True

    def CreatePassiveScalar(self, createStatic):
        opts = dict(dtype=self._domain.getDtype(), device=self._domain.getDevice())
        c = self._domain.getPassiveScalarChannels()
        if createStatic:
            self.setPassiveScalar(torch.zeros(1, c, **opts))
        else:
            self.setPassiveScalar(torch.zeros(_nc_shape(self._domain.getSpatialDims(), c, self._size), **opts))

    def setPassiveScalarType(self, types):
        """FixedBoundary::setPassiveScalarType: one type for every channel, or one per channel."""
        if isinstance(types, (list, tuple)):
            ts = [BoundaryConditionType(t) for t in types]
            _check(len(ts) == 1 or len(ts) == self._domain.getPassiveScalarChannels(),
                   "Passive scalar boundary types must be static or match passive scalar channels.")
            _check(len(ts) <= 32, "libpict_hip supports per-channel passive scalar boundary types for at most 32 "
                   "channels.")
            self._passiveScalarTypes = ts
        else:
            self._passiveScalarTypes = [BoundaryConditionType(types)]
        self._changed = True

    def _scalar_bc_mask(self):
        """Bit c set: channel c is NEUMANN (a static type sets every bit alike)."""
        ts = self._passiveScalarTypes
        if not ts:
            return 0
        if len(ts) == 1:
            return -1 if ts[0] == BoundaryConditionType.NEUMANN else 0
        m = 0
        for c, t in enumerate(ts):
            if t == BoundaryConditionType.NEUMANN:
                m |= 1 << c
        return m - (1 << 32) if m >= (1 << 31) else m

    def setPassiveScalarGrad(self, t):
        self._passiveScalar_grad = t

    def CreatePassiveScalarGrad(self):
        if self._passiveScalar is not None:
            self._passiveScalar_grad = torch.zeros_like(self._passiveScalar)

    # --- transform ---
    transform = property(lambda self: self._transform)

    def hasTransform(self):
        return self._transform is not None

    def setTransform(self, t):
        dims = self._domain.getSpatialDims()
        self._domain._check_tensor_device(t)
        _check(t.dim() == dims + 2 and t.size(-1) == 2 * dims * dims + 1, "Invalid boundary transform shape.")
        if self.hasSize():
            _check(_spatial_size_xyz(t, dims, False) == list(self._size[:3]),
                   "Boundary transform size must match the boundary.")
        else:
            self._setSizeFromTensor(t, False)
        self._transform = t.contiguous()
        self._changed = True

    def clearTransform(self):
        self._transform = None
        self._changed = True

    # --- misc ---
    def Detach(self):
        self.DetachFwd()
        self.DetachGrad()

    def DetachFwd(self):
        self._velocity = self._velocity.detach()
        if self._passiveScalar is not None:
            self._passiveScalar = self._passiveScalar.detach()
        if self._transform is not None:
            self._transform = self._transform.detach()

    def DetachGrad(self):
        if self._velocity_grad is not None:
            self._velocity_grad = self._velocity_grad.detach()
        if self._passiveScalar_grad is not None:
            self._passiveScalar_grad = self._passiveScalar_grad.detach()

    def Copy(self):
        nb = FixedBoundary(self._velocity, self._velocityType, self._passiveScalar, None, self._transform, self._domain)
        if self.hasPassiveScalar():
            nb.setPassiveScalarType(self._passiveScalarTypes)
        return nb

    def Clone(self):
        c = lambda t: None if t is None else t.clone()  # noqa: E731
        nb = FixedBoundary(self._velocity.clone(), self._velocityType, c(self._passiveScalar), None,
                           c(self._transform), self._domain)
        if self.hasPassiveScalar():
            nb.setPassiveScalarType(self._passiveScalarTypes)
        return nb

    def __str__(self):
        return "FIXED(%dD, size=%s, vStatic=%s, transform=%s)" % (
            self._domain.getSpatialDims(), self.getSizes() if self.hasSize() else "?",
            "true" if self._velocityStatic else "false", "true" if self.hasTransform() else "false")


def _nc_shape(dims, channels, size_xyz):
    return [1, channels] + [size_xyz[d] for d in reversed(range(dims))]


# ----------------------------------------------------------------------------
# Block (domain_structs.cpp:1103-2160)
# ----------------------------------------------------------------------------

class Block:
    def __init__(self, domain, name, velocity=None, pressure=None, passiveScalar=None, vertexCoordinates=None,
                 size=None):
        self._domain = domain
        self.name = name
        dims = domain.getSpatialDims()
        opts = dict(dtype=domain.getDtype(), device=domain.getDevice())
        self._viscosity = None
        self._viscosity_grad = None
        self.velocitySource = None
        self.velocitySource_grad = None
        self._velocitySourceStatic = False
        self.passiveScalar = None
        self._velocity_grad = None
        self._pressure_grad = None
        self._passiveScalar_grad = None
        self._vertexCoordinates = None
        self._transform = None
        self._faceTransform = None
        self.globalOffset = 0
        self.csrOffset = 0
        if size is not None:
            for d in range(dims):
                _check(size[d] > 2, "all spatial dimensions must be at least 3.")
            self.velocity = torch.zeros(_nc_shape(dims, dims, [size[0], size[1], size[2]]), **opts)
        elif velocity is not None:
            domain._check_tensor_device(velocity)
            _check(velocity.is_contiguous(), "velocity must be contiguous")
            _check(velocity.dim() == dims + 2, "Velocity dimensionality must match domain. Layout should be "
                   "NC<spatial-dims>, e.g. NCDHW for 3D.")
            _check(velocity.size(1) == dims, "The velocity channels must match the spatial dimensions.")
            _check(velocity.size(0) == 1, "Batches are not yet supported.")
            for d in range(2, velocity.dim()):
                _check(velocity.size(d) > 2, "all spatial dimensions must be at least 3.")
            _check(velocity.dtype == domain.getDtype(), "Velocity dtype must match domain.")
            self.velocity = velocity
        else:
            ref = pressure if pressure is not None else (passiveScalar if passiveScalar is not None else None)
            if ref is not None:
                _check(ref.dim() == dims + 2, "Field dimensionality must match domain.")
                sz = _spatial_size_xyz(ref, dims)
            elif vertexCoordinates is not None:
                _check(vertexCoordinates.dim() == dims + 2, "Field dimensionality must match domain.")
                sz = [s - 1 for s in _spatial_size_xyz(vertexCoordinates, dims)]
                for d in range(dims, 3):
                    sz[d] = 1
            else:
                raise RuntimeError("At least one field (velocity, pressure, passiveScalar, vertexCoordinates) or the "
                                   "size must be given to create a block.")
            for d in range(dims):
                _check(sz[d] > 2, "all spatial dimensions must be at least 3.")
            self.velocity = torch.zeros(_nc_shape(dims, dims, sz), **opts)
        self.boundaries = [PeriodicBoundary(domain) for _ in range(2 * dims)]
        if pressure is not None:
            self.setPressure(pressure)
        else:
            self.CreatePressure()
        if passiveScalar is not None:
            self.setPassiveScalar(passiveScalar)
        elif domain.hasPassiveScalar():
            self.CreatePassiveScalar()
        if vertexCoordinates is not None:
            self.setVertexCoordinates(vertexCoordinates)
        self._changed = True

    # --- structure ---
    def getParentDomain(self):
        return self._domain

    def getSpatialDims(self):
        return self._domain.getSpatialDims()

    def getDim(self, dim):
        return self.velocity.size(dim)

    def getAxis(self, dim):
        _check(0 <= dim < self.getSpatialDims(), "Axis index must be within spatial dimensions")
        return self.velocity.size(self.velocity.dim() - 1 - dim)

    def getSizes(self):
        dims = self.getSpatialDims()
        s = _spatial_size_xyz(self.velocity, dims)
        return Int4(s[0], s[1], s[2], dims)

    def getStrides(self):
        s = self.getSizes()
        return Int4(1, s.x, s.x * s.y, s.x * s.y * s.z)

    def getDevice(self):
        return self._domain.getDevice()

    def getDtype(self):
        return self._domain.getDtype()

    def ComputeCSRSize(self):
        """Block::ComputeCSRSize (domain_structs.cpp:2151-2161)."""
        size = self.getSizes()
        stride = self.getStrides()
        dims = self.getSpatialDims()
        csr = stride.w * (2 * dims + 1)
        for d in range(dims):
            area = size[(d + 1) % 3] * size[(d + 2) % 3]
            if self.IsUnconnectedBoundary(2 * d):
                csr -= area
            if self.IsUnconnectedBoundary(2 * d + 1):
                csr -= area
        return csr

    def IsUnconnectedBoundary(self, index):
        return self.boundaries[index].type in (BoundaryType.FIXED, BoundaryType.DIRICHLET,
                                               BoundaryType.DIRICHLET_VARYING, BoundaryType.NEUMANN)

    def _check_data(self, t, channels, allow_static, name):
        self._domain._check_tensor_device(t)
        _check(t.is_contiguous(), name + " must be contiguous")
        n = self.getSpatialDims() + 2
        _check(t.dim() == n or (allow_static and t.dim() == 2),
               "Dimensions of %s must be %d%s." % (name, n, " or 2" if allow_static else ""))
        _check(t.size(0) == 1, "Batches (dim 0) are not yet supported (batch dimension must be 1).")
        if channels >= 1:
            _check(t.size(1) == channels, "Channels (dim 1) of %s must be %d." % (name, channels))
        if t.dim() != 2:
            for d in range(2, n):
                _check(self.velocity.size(d) == t.size(d), "Spatial dimension %d of %s must match (%d), but is (%d)."
                       % (d, name, self.velocity.size(d), t.size(d)))
        _check(t.dtype == self.velocity.dtype, name + "has wrong dtype.")

    def _new_field(self, channels):
        return torch.zeros(_nc_shape(self.getSpatialDims(), channels, self.getSizes()),
                           dtype=self.getDtype(), device=self.getDevice())

    # --- velocity ---
    def setVelocity(self, v):
        self._check_data(v, self.getSpatialDims(), False, "Velocity")
        self.velocity = v
        self._changed = True

    def CreateVelocity(self):
        self.setVelocity(self._new_field(self.getSpatialDims()))

    def getVelocity(self, computational):
        if computational:
            _check(self.hasTransform(), "Coordinates or Transformations are required to compute computational "
                   "velocities.")
            return TransformVectors(self.velocity, self._transform, True)
        return self.velocity

    velocityGrad = property(lambda self: self._velocity_grad if self._velocity_grad is not None else torch.empty(0))

    def setVelocityGrad(self, t):
        self._check_data(t, self.getSpatialDims(), False, "VelocityGrad")
        self._velocity_grad = t
        self._changed = True

    def CreateVelocityGrad(self):
        self.setVelocityGrad(torch.zeros_like(self.velocity))

    # --- velocity source ---
    isVelocitySourceStatic = property(lambda self: self._velocitySourceStatic)

    def hasVelocitySource(self):
        return self.velocitySource is not None

    def setVelocitySource(self, t):
        self._check_data(t, self.getSpatialDims(), True, "VelocitySource")
        if self.velocitySource_grad is not None and self.velocitySource_grad.dim() != t.dim():
            raise RuntimeError("New velocity source does not match existing gradient tensor.")
        self.velocitySource = t
        self._velocitySourceStatic = t.dim() == 2
        self._changed = True

    def CreateVelocitySource(self, createStatic):
        if createStatic:
            t = torch.zeros(1, self.getSpatialDims(), dtype=self.getDtype(), device=self.getDevice())
        else:
            t = self._new_field(self.getSpatialDims())
        self.setVelocitySource(t)

    def clearVelocitySource(self):
        if self.velocitySource is not None:
            self.velocitySource = None
            self._changed = True

    velocitySourceGrad = property(lambda self: self.velocitySource_grad)

    def hasVelocitySourceGrad(self):
        return self.velocitySource_grad is not None

    def setVelocitySourceGrad(self, t):
        _check(self.velocitySource is not None and t.shape == self.velocitySource.shape,
               "VelocitySourceGrad must match velocity source.")
        self.velocitySource_grad = t
        self._changed = True

    def CreateVelocitySourceGrad(self):
        if self.velocitySource is not None:
            self.setVelocitySourceGrad(torch.zeros_like(self.velocitySource))
        else:
            self.clearVelocitySourceGrad()

    def clearVelocitySourceGrad(self):
        if self.velocitySource_grad is not None:
            self.velocitySource_grad = None
            self._changed = True

    # --- viscosity ---
    viscosity = property(lambda self: self._viscosity)

    def setViscosity(self, v):
        self._check_data(v, 1, True, "BlockViscosity")
        if self._viscosity_grad is not None and self._viscosity_grad.dim() != v.dim():
            raise RuntimeError("New block viscosity does not match existing gradient tensor.")
        self._viscosity = v
        self._changed = True

    def hasViscosity(self):
        return self._viscosity is not None

    def isViscosityStatic(self):
        return self._viscosity.dim() == 2 if self._viscosity is not None else True

    def clearViscosity(self):
        if self._viscosity is not None:
            self._viscosity = None
            self._changed = True

    viscosityGrad = property(lambda self: self._viscosity_grad)

    def setViscosityGrad(self, t):
        _check(self._viscosity is not None and t.shape == self._viscosity.shape,
               "ViscosityGrad must match viscosity.")
        self._viscosity_grad = t
        self._changed = True

    def CreateViscosityGrad(self):
        """Block::CreateViscosityGrad (domain_structs.cpp:1510-1517): cleared when the block has no viscosity."""
        if self._viscosity is not None:
            self.setViscosityGrad(torch.zeros_like(self._viscosity))
        else:
            self.clearViscosityGrad()

    def hasViscosityGrad(self):
        return self._viscosity_grad is not None

    def clearViscosityGrad(self):
        if self._viscosity_grad is not None:
            self._viscosity_grad = None
            self._changed = True

    # --- pressure ---
    def setPressure(self, p):
        self._check_data(p, 1, False, "Pressure")
        self.pressure = p
        self._changed = True

    def CreatePressure(self):
        self.setPressure(self._new_field(1))

    pressureGrad = property(lambda self: self._pressure_grad if self._pressure_grad is not None else torch.empty(0))

    def setPressureGrad(self, t):
        self._check_data(t, 1, False, "PressureGrad")
        self._pressure_grad = t
        self._changed = True

    def CreatePressureGrad(self):
        self.setPressureGrad(torch.zeros_like(self.pressure))

    # --- passive scalar ---
    def setPassiveScalar(self, s):
        self._check_data(s, self._domain.getPassiveScalarChannels(), False, "PassiveScalar")
        self.passiveScalar = s
        self._changed = True

    def CreatePassiveScalar(self):
        self.setPassiveScalar(self._new_field(self._domain.getPassiveScalarChannels()))

    def hasPassiveScalar(self):
        return self.passiveScalar is not None

    def getPassiveScalarChannels(self):
        return self.passiveScalar.size(1) if self.passiveScalar is not None else 0

    passiveScalarGrad = property(lambda self: self._passiveScalar_grad)

    def setPassiveScalarGrad(self, t):
        self._passiveScalar_grad = t

    def CreatePassiveScalarGrad(self):
        if self.passiveScalar is not None:
            self._passiveScalar_grad = torch.zeros_like(self.passiveScalar)

    def hasPassiveScalarViscosity(self):
        return False

    passiveScalarViscosity = property(lambda self: None)

    # --- coordinates and transformations ---
    vertexCoordinates = property(lambda self: self._vertexCoordinates)
    transform = property(lambda self: self._transform)
    faceTransform = property(lambda self: self._faceTransform)

    def hasVertexCoordinates(self):
        return self._vertexCoordinates is not None

    def hasTransform(self):
        return self._transform is not None

    def hasFaceTransform(self):
        return self._faceTransform is not None

    def setVertexCoordinates(self, coords):
        """Block::setVertexCoordinates (domain_structs.cpp:2001-2019)."""
        self._domain._check_tensor_device(coords)
        _check(coords.dim() == self.velocity.dim(), "new vertex coordinates dimensions must match velocity.")
        _check(coords.size(0) == 1, "batches are not yet supported.")
        _check(coords.size(1) == self.getSpatialDims(), "vertex coordinates channel dimension must match spatial "
               "dimensions.")
        for d in range(2, self.velocity.dim()):
            _check(self.velocity.size(d) == coords.size(d) - 1, "vertex coordinates spatial dimension must match "
                   "velocity's +1.")
        _check(coords.dtype == self.velocity.dtype, "Vertex coordinates and velocity must have same dtype.")
        coords = coords.contiguous()
        transform = CoordsToTransforms(coords)
        face = CoordsToFaceTransforms(coords) if self.getSpatialDims() > 1 else None
        self.clearCoordsTransforms()
        self._vertexCoordinates = coords
        self._transform = transform
        self._faceTransform = face
        self._update_boundary_transforms()
        self._changed = True

    def setTransform(self, transform, faceTransform=None):
        """Block::setTransform (domain_structs.cpp:2020-2051)."""
        self._domain._check_tensor_device(transform)
        _check(transform.dim() == self.velocity.dim(), "new ransform dimensions must match velocity.")
        _check(transform.size(0) == 1, "batches are not yet supported.")
        dims = self.getSpatialDims()
        _check(transform.size(-1) == 2 * dims * dims + 1, "Transform channels must match spatial dimensions")
        for d in range(2, self.velocity.dim()):
            _check(self.velocity.size(d) == transform.size(d - 1), "Transform spatial dimension must match velocity's")
        _check(transform.dtype == self.velocity.dtype, "Transform and velocity must have same dtype.")
        if faceTransform is not None:
            _check(faceTransform.dim() == self.velocity.dim() + 1, "new face transform dimensions must match "
                   "velocity.")
            _check(faceTransform.size(1) == dims, "channel dimension must match spatial dimensions.")
        self.clearCoordsTransforms()
        self._transform = transform.contiguous()
        self._faceTransform = None if faceTransform is None else faceTransform.contiguous()
        self._update_boundary_transforms()
        self._changed = True

    def clearCoordsTransforms(self):
        self._vertexCoordinates = None
        self._transform = None
        self._faceTransform = None
        self._changed = True

    def getCellSizes(self):
        _check(self.hasTransform(), "Block has no transformation.")
        det = self._transform[..., -1]
        return det.unsqueeze(1)

    def getCellCoordinates(self):
        _check(self.hasVertexCoordinates(), "Block has no vertex coordinates.")
        c = self._vertexCoordinates
        dims = self.getSpatialDims()
        for d in range(dims):
            ax = 2 + d
            c = 0.5 * (c.narrow(ax, 0, c.size(ax) - 1) + c.narrow(ax, 1, c.size(ax) - 1))
        return c

    def GetCoordinateOrientation(self):
        if not self.hasTransform():
            return 1
        s = torch.sign(self._transform[..., -1])
        mn, mx = int(s.min().item()), int(s.max().item())
        return mn if mn == mx else 0

    def _face_transform_slice(self, face):
        """Block::GetFaceTransformBoundarySlice (domain_structs.cpp:1809-1835)."""
        _check(self.hasFaceTransform(), "faceTransform missing to slice for boundary.")
        axis, upper = face >> 1, face & 1
        dims = self.getSpatialDims()
        idx = [slice(None), axis]
        for d in reversed(range(dims)):
            if d == axis:
                idx.append(slice(-1, None) if upper else slice(0, 1))
            else:
                idx.append(slice(0, -1))
        idx.append(slice(None))
        return self._faceTransform[tuple(idx)].clone().contiguous()

    def _update_boundary_transforms(self):
        """Block::UpdateBoundaryTransforms (domain_structs.cpp:1837-1865)."""
        if not self.hasFaceTransform():
            return
        for face, b in enumerate(self.boundaries):
            if b.type == BoundaryType.FIXED:
                b.setTransform(self._face_transform_slice(face))

    # --- boundaries ---
    def getBoundaries(self):
        return list(self.boundaries)

    def getBoundary(self, face):
        f = _face_index(face)
        _check(0 <= f < 2 * self.getSpatialDims(), "Invalid boundary location specified.")
        return self.boundaries[f]

    def _setBoundary(self, face, boundary):
        f = _face_index(face)
        self._check_face(f)
        dims = self.getSpatialDims()
        if boundary.type == BoundaryType.FIXED:
            if boundary.hasSize():
                bs = boundary.getSizes()
                _check(bs[f >> 1] == 1, "Boundary axis size must be 1.")
                for i in range(1, dims):
                    ax = ((f >> 1) + i) % dims
                    _check(bs[ax] == self.getAxis(ax), "Boundary size must match block.")
            else:
                s = self.getSizes()
                sz = [s.x, s.y, s.z, dims]
                sz[f >> 1] = 1
                boundary._setSize(sz)
        elif boundary.type == BoundaryType.CONNECTED:
            other = boundary.getConnectedBlock()
            for i in range(1, dims):
                ax = ((f >> 1) + i) % dims
                _check(self.getAxis(ax) == other.getAxis(boundary.getConnectionAxis(i)),
                       "%s connection axis size does not match." % ("First" if i == 1 else "Second"))
        self.boundaries[f] = boundary
        self._changed = True

    setBoundary = _setBoundary

    def getFixedBoundaries(self):
        return [(i, b) for i, b in enumerate(self.boundaries) if b.type == BoundaryType.FIXED]

    def hasPrescribedBoundary(self):
        return any(self.IsUnconnectedBoundary(i) for i in range(len(self.boundaries)))

    def isAllFixedBoundariesPassiveScalarTypeStatic(self):
        return all(b.isPassiveScalarBoundaryTypeStatic() for _, b in self.getFixedBoundaries())

    def _check_face(self, f):
        _check(0 <= f < 2 * self.getSpatialDims(), "Invalid face index %d." % f)

    def _make_fixed(self, face, velocity=None, passiveScalar=None, scalarType=None):
        """Block::MakeFixedBoundary (domain_structs.cpp:1902-1922)."""
        self._check_face(face)
        dims = self.getSpatialDims()
        if velocity is not None or passiveScalar is not None or self.hasFaceTransform():
            tr = self._face_transform_slice(face) if self.hasFaceTransform() else None
            fb = FixedBoundary(velocity, BoundaryConditionType.DIRICHLET, passiveScalar, None, tr, self._domain)
        else:
            v = torch.zeros(1, dims, dtype=self.getDtype(), device=self.getDevice())
            fb = FixedBoundary(v, BoundaryConditionType.DIRICHLET, None, None, None, self._domain)
        if scalarType is not None and fb.hasPassiveScalar():
            fb.setPassiveScalarType(scalarType)
        self._setBoundary(face, fb)

    def _close_connected(self, face):
        """Block::CloseConnectedBoudary with useFixed=True (domain_structs.cpp:1773-1806)."""
        b = self.boundaries[face]
        if b.type == BoundaryType.CONNECTED:
            b.getConnectedBlock()._make_fixed(b.axes[0])
        elif b.type == BoundaryType.PERIODIC:
            self._make_fixed(face ^ 1)

    def ConnectBlock(self, faceIndex, otherBlock, otherFaceIndex, axis1Index=-1, axis2Index=-1):
        """Block::ConnectBlock / ConnectBlocks (domain_structs.cpp:1924-1935, :1064-1097)."""
        if isinstance(faceIndex, str):
            f1, f2 = _face_index(faceIndex), _face_index(otherFaceIndex)
            a1 = _face_index(axis1Index) if axis1Index not in ("", -1, None) else -1
            a2 = _face_index(axis2Index) if axis2Index not in ("", -1, None) else -1
        else:
            f1, f2, a1, a2 = int(faceIndex), int(otherFaceIndex), int(axis1Index), int(axis2Index)
        self._check_face(f1)
        self._close_connected(f1)
        otherBlock._check_face(f2)
        otherBlock._close_connected(f2)
        _check(otherBlock.getParentDomain() is self._domain, "The blocks must belong to the same domain.")
        dims = self.getSpatialDims()
        axes1, axes2 = [f2], [f1]
        if dims > 1:
            axes1.append(a1)
            f1d, f2d = f1 >> 1, f2 >> 1
            swapped = False
            if dims == 2 or (a1 >> 1) == (f2d + 1) % dims:
                axes2.append((((f1d + 1) % dims) << 1) | (a1 & 1))
            else:
                _check((a2 >> 1) == (f2d + 1) % dims, "Invalid connection.")
                axes2.append((((f1d + 2) % dims) << 1) | (a2 & 1))
                swapped = True
            if dims > 2:
                axes1.append(a2)
                if not swapped:
                    axes2.append((((f1d + 2) % dims) << 1) | (a2 & 1))
                else:
                    axes2.append((((f1d + 1) % dims) << 1) | (a1 & 1))
        self._setBoundary(f1, ConnectedBoundary(otherBlock, axes1, self._domain))
        otherBlock._setBoundary(f2, ConnectedBoundary(self, axes2, self._domain))

    def MakePeriodic(self, axisIndex):
        """Block::MakePeriodic (domain_structs.cpp:1937-1955)."""
        axis = _axis_index(axisIndex)
        _check(0 <= axis < self.getSpatialDims(), "Invalid axis index.")
        lo, up = axis << 1, (axis << 1) | 1
        bl, bu = self.boundaries[lo], self.boundaries[up]
        if not (bl.type == BoundaryType.PERIODIC and bu.type == BoundaryType.PERIODIC):
            if bl.type == BoundaryType.CONNECTED:
                self._close_connected(lo)
            if bu.type == BoundaryType.CONNECTED:
                self._close_connected(up)
            if self.boundaries[lo].type != BoundaryType.PERIODIC:
                self._setBoundary(lo, PeriodicBoundary(self._domain))
            if self.boundaries[up].type != BoundaryType.PERIODIC:
                self._setBoundary(up, PeriodicBoundary(self._domain))

    def MakeAllPeriodic(self):
        for d in range(self.getSpatialDims()):
            self.MakePeriodic(d)

    def CloseBoundary(self, faceIndex, velocity=None, passiveScalar=None):
        """Block::CloseBoundary (domain_structs.cpp:1976-1985): FixedBoundary with Dirichlet velocity."""
        f = _face_index(faceIndex)
        self._check_face(f)
        self._close_connected(f)
        self._make_fixed(f, velocity, passiveScalar)

    def CloseAllBoundaries(self):
        for f in range(2 * self.getSpatialDims()):
            self.CloseBoundary(f)

    # --- queries ---
    def getMaxVelocity(self, withBounds=True, computational=False):
        """Largest velocity component magnitude of the block and, withBounds, its fixed boundaries, as a 0-dim
        tensor (domain_structs.cpp:1564-1600)."""
        m = torch.amax(self.getVelocity(computational).abs())
        if withBounds:
            for _, b in self.getFixedBoundaries():
                bv = b.getVelocity(computational) if b.hasSize() else b.velocity
                m = torch.maximum(m, torch.amax(bv.abs()))
        return m

    def getMaxVelocityMagnitude(self, withBounds=True, computational=False):
        v = self.getVelocity(computational)
        m = torch.max(torch.linalg.vector_norm(v, dim=1))
        if withBounds:
            for _, b in self.getFixedBoundaries():
                bv = b.getVelocity(computational) if b.hasSize() else b.velocity
                m = torch.maximum(m, torch.max(torch.linalg.vector_norm(bv, dim=1)))
        return m.reshape(1)

    def IsTensorChanged(self):
        return self._changed or any(b.IsTensorChanged() for b in self.boundaries)

    def _set_changed(self, c):
        self._changed = c
        for b in self.boundaries:
            b._changed = c

    def Detach(self):
        self.DetachFwd()
        self.DetachGrad()

    def DetachFwd(self):
        self.velocity = self.velocity.detach()
        self.pressure = self.pressure.detach()
        for n in ("passiveScalar", "velocitySource", "_viscosity", "_vertexCoordinates", "_transform",
                  "_faceTransform"):
            t = getattr(self, n)
            if t is not None:
                setattr(self, n, t.detach())
        for b in self.boundaries:
            if b.type == BoundaryType.FIXED:
                b.DetachFwd()

    def DetachGrad(self):
        for n in ("_velocity_grad", "_pressure_grad", "_passiveScalar_grad", "velocitySource_grad",
                  "_viscosity_grad"):
            t = getattr(self, n)
            if t is not None:
                setattr(self, n, t.detach())
        for b in self.boundaries:
            if b.type == BoundaryType.FIXED:
                b.DetachGrad()

    def __str__(self):
        return 'Block["%s" %dD, size=%s, bounds=( %s )]' % (
            self.name, self.getSpatialDims(), self.getSizes(),
            " ".join("%s=%s" % (n, b) for n, b in zip(["-x", "+x", "-y", "+y", "-z", "+z"], self.boundaries)))

    # --- descriptor packing ---
    def _pack_signature(self, block_index_of):
        """Everything the block descriptor is built from: tensor addresses and ranks, offsets and the boundary
        objects.  An unchanged signature means an identical descriptor."""
        def t(x):
            return None if x is None else (x.data_ptr(), x.dim())
        sig = [self.globalOffset, self.csrOffset, t(self.velocity), t(self.pressure), t(self._viscosity),
               t(self.velocitySource), t(self._transform), t(self._faceTransform), t(self._velocity_grad),
               t(self._pressure_grad), t(self._viscosity_grad), t(self.velocitySource_grad), t(self.passiveScalar),
               t(self._passiveScalar_grad)]
        for bd in self.boundaries:
            if bd.type == BoundaryType.CONNECTED:
                sig.append((id(bd), block_index_of(bd.getConnectedBlock()), tuple(bd.axes)))
            elif bd.type == BoundaryType.FIXED:
                sig.append((id(bd), t(bd.velocity), t(bd._velocity_grad), t(bd.transform), int(bd.velocityType),
                            None if bd._size is None else tuple(bd._size), t(bd._passiveScalar),
                            t(bd._passiveScalar_grad), bd._scalar_bc_mask()))
            else:
                sig.append(int(bd.type))
        return tuple(sig)

    def _pack(self, block_index_of):
        sig = self._pack_signature(block_index_of)
        cached = getattr(self, "_packed", None)
        if cached is not None and cached[0] == sig:
            return cached[1]
        b = self._pack_full(block_index_of)
        self._packed = (sig, b)
        return b

    def _pack_full(self, block_index_of):
        b = _abi.BlockDesc()
        s, st = self.getSizes(), self.getStrides()
        b.global_offset = self.globalOffset
        b.csr_offset = self.csrOffset
        b.size[:] = [s.x, s.y, s.z, s.w]
        b.stride[:] = [st.x, st.y, st.z, st.w]
        b.viscosity_static = int(self.isViscosityStatic())
        b.velocity_source_static = int(self._velocitySourceStatic)
        b.has_transform = int(self._transform is not None)
        b.has_face_transform = int(self._faceTransform is not None)
        b.velocity = _ptr(self.velocity)
        b.pressure = _ptr(self.pressure)
        b.viscosity = _ptr(self._viscosity)
        b.velocity_source = _ptr(self.velocitySource)
        b.transform = _ptr(self._transform)
        b.face_transform = _ptr(self._faceTransform)
        b.velocity_grad = _ptr(self._velocity_grad)
        b.pressure_grad = _ptr(self._pressure_grad)
        b.viscosity_grad = _ptr(self._viscosity_grad)
        b.velocity_source_grad = _ptr(self.velocitySource_grad)
        b.scalar = _ptr(self.passiveScalar)
        b.scalar_grad = _ptr(self._passiveScalar_grad)
        for f, bd in enumerate(self.boundaries):
            d = b.bounds[f]
            d.type = int(bd.type)
            if bd.type == BoundaryType.CONNECTED:
                d.connected_block = block_index_of(bd.getConnectedBlock())
                ax = bd.axes + [0] * (4 - len(bd.axes))
                d.axes[:] = ax
            elif bd.type == BoundaryType.FIXED:
                d.velocity_static = int(bd.isVelocityStatic)
                d.velocity_bc = int(bd.velocityType)
                d.has_transform = int(bd.hasTransform())
                bs, bst = bd.getSizes(), bd.getStrides()
                d.size[:] = [bs.x, bs.y, bs.z, bs.w]
                d.stride[:] = [bst.x, bst.y, bst.z, bst.w]
                d.velocity = _ptr(bd.velocity)
                d.velocity_grad = _ptr(bd._velocity_grad)
                d.transform = _ptr(bd.transform)
                d.scalar_static = int(bd.isPassiveScalarStatic())
                d.scalar_bc_mask = bd._scalar_bc_mask()
                d.scalar = _ptr(bd._passiveScalar)
                d.scalar_grad = _ptr(bd._passiveScalar_grad)
        return b


# ----------------------------------------------------------------------------
# Domain (domain_structs.cpp:2207-3364)
# ----------------------------------------------------------------------------

class Domain:
    def __init__(self, spatialDims, viscosity, name, dtype, device, passiveScalarChannels=1, scalarViscosity=None,
                 _allow_cpu=False):
        _check(0 < spatialDims < 4, "spatialDims must be 1, 2, or 3.")
        _check(passiveScalarChannels >= 0, "passiveScalarChannels can't be negative.")
        self.name = name
        self._dims = int(spatialDims)
        self._channels = int(passiveScalarChannels)
        self._dtype = dtype
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", 0)
        _check(dev.type == "cuda" or _allow_cpu, "device must be a CUDA device.")
        self._device = dev
        self.setViscosity(viscosity)
        self.passiveScalarViscosity = None
        if scalarViscosity is not None:
            self.setScalarViscosity(scalarViscosity)
        self.blocks = []
        self._initialized = False
        self._changed = True
        self._atlas = None
        self._desc = None
        self._blocks_arr = None
        self.C = self.P = self.A = None
        self.C_grad = self.P_grad = None
        self.A_grad = torch.empty(0)
        self.viscosity_grad = None
        for n in ("scalarRHS", "scalarResult", "velocityRHS", "velocityResult", "pressureRHS", "pressureRHSdiv",
                  "pressureResult"):
            setattr(self, n, None)
            setattr(self, n + "_grad", torch.empty(0))
        self._totalSize = 0

    def __del__(self):
        try:
            self._destroy_atlas()
        except Exception:
            pass

    def _destroy_atlas(self):
        if self._atlas is not None:
            _abi.load_library().pict_atlas_destroy(self._atlas)
            self._atlas = None

    def _check_tensor_device(self, t):
        _check(isinstance(t, torch.Tensor), "expected a tensor")
        if self._device.type == "cuda":
            _check(t.is_cuda, "tensor must be a CUDA tensor")
        else:
            _check(t.device.type == "cpu", "tensor must be on the domain's device")

    # --- basic queries ---
    def getSpatialDims(self):
        return self._dims

    def getDtype(self):
        return self._dtype

    def getDevice(self):
        return self._device

    def getPassiveScalarChannels(self):
        return self._channels

    def hasPassiveScalar(self):
        return self._channels > 0

    def getNumBlocks(self):
        return len(self.blocks)

    def getBlock(self, i):
        return self.blocks[i]

    def getBlocks(self):
        return list(self.blocks)

    def getTotalSize(self):
        return self._totalSize

    def getMaxBlockSize(self):
        return max(b.getStrides().w for b in self.blocks)

    def IsInitialized(self):
        return self._initialized

    # --- viscosity ---
    def setViscosity(self, viscosity):
        _check(isinstance(viscosity, torch.Tensor) and viscosity.device.type == "cpu",
               "viscosity must be a host tensor")
        _check(viscosity.dim() == 1 and viscosity.size(0) == 1, "viscosity must be a scalar tensor [1].")
        _check(viscosity.dtype == self._dtype, "Data type of viscosity does not match.")
        self.viscosity = viscosity
        self._changed = True

    def hasBlockViscosity(self):
        return any(b.hasViscosity() for b in self.blocks)

    def setScalarViscosity(self, v):
        """Domain::setScalarViscosity (domain_structs.cpp:2530-2539): a device tensor [1] or [channels]."""
        _check(isinstance(v, torch.Tensor) and v.dim() == 1, "Scalar viscosity must be 1D.")
        _check(v.size(0) == 1 or v.size(0) == self._channels,
               "Scalar viscosity must be static (a scalar) or match passive scalar channels.")
        _check(v.dtype == self._dtype, "Data type of scalar viscosity does not match.")
        self._check_tensor_device(v)
        _check(v.is_contiguous(), "Scalar viscosity must be contiguous.")
        self.passiveScalarViscosity = v
        self._changed = True

    def clearScalarViscosity(self):
        self.passiveScalarViscosity = None
        self.passiveScalarViscosity_grad = None
        self._changed = True

    passiveScalarViscosityGrad = property(lambda self: getattr(self, "passiveScalarViscosity_grad", None))

    def setPassiveScalarViscosityGrad(self, t):
        """domain_structs.cpp:2884-2893."""
        _check(self.hasPassiveScalarViscosity(), "Domain has no passive scalar viscosity.")
        _check(t.dim() == 1 and t.size(0) == self.passiveScalarViscosity.size(0),
               "Scalar viscosity grad shape must match Scalar viscosity.")
        _check(t.dtype == self._dtype, "Data type of scalar viscosity does not match.")
        self._check_tensor_device(t)
        self.passiveScalarViscosity_grad = t
        self._changed = True

    def clearPassiveScalarViscosityGrad(self):
        if getattr(self, "passiveScalarViscosity_grad", None) is not None:
            self.passiveScalarViscosity_grad = None
            self._changed = True

    def CreatePassiveScalarViscosityGrad(self):
        """domain_structs.cpp:2900-2907."""
        if self.hasPassiveScalarViscosity():
            self.setPassiveScalarViscosityGrad(torch.zeros_like(self.passiveScalarViscosity))
        else:
            self.clearPassiveScalarViscosityGrad()

    def hasPassiveScalarViscosity(self):
        return self.passiveScalarViscosity is not None

    def isPassiveScalarViscosityStatic(self):
        return self.passiveScalarViscosity is None or self.passiveScalarViscosity.size(0) == 1

    def hasPassiveScalarBlockViscosity(self):
        return False

    viscosityGrad = property(lambda self: self.viscosity_grad)

    def hasViscosityGrad(self):
        return self.viscosity_grad is not None

    def setViscosityGrad(self, t):
        self.viscosity_grad = t
        self._changed = True

    def clearViscosityGrad(self):
        self.viscosity_grad = None
        self._changed = True

    def CreateViscosityGrad(self):
        """Domain::CreateViscosityGrad (domain_structs.cpp:2875-2886): the domain's [1] gradient (a device tensor,
        while the viscosity itself is a host tensor) and the blocks' viscosity gradients."""
        self.setViscosityGrad(torch.zeros(1, dtype=self._dtype, device=self._device))
        for b in self.blocks:
            b.CreateViscosityGrad()
        self.CreatePassiveScalarViscosityGrad()

    # --- blocks ---
    def CreateBlock(self, velocity=None, pressure=None, passiveScalar=None, vertexCoordinates=None, name="Block"):
        b = Block(self, name, velocity, pressure, passiveScalar, vertexCoordinates)
        self.AddBlock(b)
        return b

    def CreateBlockWithSize(self, size, name="Block"):
        b = Block(self, name, size=[size[0], size[1], size[2]])
        self.AddBlock(b)
        return b

    def AddBlock(self, block):
        _check(block.getParentDomain() is self, "Block belongs to a different domain.")
        self.blocks.append(block)
        self._initialized = False
        self._changed = True

    def _block_index(self, block):
        for i, b in enumerate(self.blocks):
            if b is block:
                return i
        return -1

    def hasPrescribedBoundary(self):
        return any(b.hasPrescribedBoundary() for b in self.blocks)

    def isAllFixedBoundariesPassiveScalarTypeStatic(self):
        return all(b.isAllFixedBoundariesPassiveScalarTypeStatic() for b in self.blocks)

    def hasVertexCoordinates(self):
        return all(b.hasVertexCoordinates() for b in self.blocks)

    def getVertexCoordinates(self):
        return [b.vertexCoordinates for b in self.blocks]

    def GetCoordinateOrientation(self):
        s = {b.GetCoordinateOrientation() for b in self.blocks}
        return s.pop() if len(s) == 1 else 0

    def CreateVelocityOnBlocks(self):
        for b in self.blocks:
            b.CreateVelocity()

    def CreatePressureOnBlocks(self):
        for b in self.blocks:
            b.CreatePressure()

    def CreatePassiveScalarOnBlocks(self):
        for b in self.blocks:
            b.CreatePassiveScalar()

    def CreateVelocityGradOnBlocks(self):
        for b in self.blocks:
            b.CreateVelocityGrad()

    def CreatePressureGradOnBlocks(self):
        for b in self.blocks:
            b.CreatePressureGrad()

    def CreateVelocitySourceGradOnBlocks(self):
        for b in self.blocks:
            b.CreateVelocitySourceGrad()

    def CreatePassiveScalarGradOnBlocks(self):
        for b in self.blocks:
            b.CreatePassiveScalarGrad()

    def CreateVelocityGradOnBoundaries(self):
        for b in self.blocks:
            for _, fb in b.getFixedBoundaries():
                fb.CreateVelocityGrad()

    def CreatePassiveScalarGradOnBoundaries(self):
        for b in self.blocks:
            for _, fb in b.getFixedBoundaries():
                fb.CreatePassiveScalarGrad()

    # --- global vectors ---
    def _vec(self, n):
        return torch.zeros(n, dtype=self._dtype, device=self._device)

    def _set_vec(self, attr, t, n):
        _check(isinstance(t, torch.Tensor) and t.dim() == 1 and t.size(0) == n,
               "%s must be a vector of size %d." % (attr, n))
        _check(t.dtype == self._dtype, "%s has wrong dtype." % attr)
        self._check_tensor_device(t)
        _check(t.is_contiguous(), "%s must be contiguous." % attr)
        setattr(self, attr, t)
        self._changed = True

    def setA(self, t):
        self._set_vec("A", t, self._totalSize)

    def CreateA(self):
        self.setA(self._vec(self._totalSize))

    def setVelocityRHS(self, t):
        self._set_vec("velocityRHS", t, self._totalSize * self._dims)

    def CreateVelocityRHS(self):
        self.setVelocityRHS(self._vec(self._totalSize * self._dims))

    def setVelocityResult(self, t):
        self._set_vec("velocityResult", t, self._totalSize * self._dims)

    def CreateVelocityResult(self):
        self.setVelocityResult(self._vec(self._totalSize * self._dims))

    def setPressureRHS(self, t):
        self._set_vec("pressureRHS", t, self._totalSize * self._dims)

    def CreatePressureRHS(self):
        self.setPressureRHS(self._vec(self._totalSize * self._dims))

    def setPressureRHSdiv(self, t):
        self._set_vec("pressureRHSdiv", t, self._totalSize)

    def CreatePressureRHSdiv(self):
        self.setPressureRHSdiv(self._vec(self._totalSize))

    def setPressureResult(self, t):
        self._set_vec("pressureResult", t, self._totalSize)

    def CreatePressureResult(self):
        self.setPressureResult(self._vec(self._totalSize))

    def setScalarRHS(self, t):
        self._set_vec("scalarRHS", t, self._totalSize * self._channels)

    def CreateScalarRHS(self):
        self.setScalarRHS(self._vec(self._totalSize * self._channels))

    def setScalarResult(self, t):
        self._set_vec("scalarResult", t, self._totalSize * self._channels)

    def CreateScalarResult(self):
        self.setScalarResult(self._vec(self._totalSize * self._channels))

    # gradients of the global vectors
    def setAGrad(self, t):
        self._set_vec("A_grad", t, self._totalSize)

    def CreateAGrad(self):
        self.setAGrad(self._vec(self._totalSize))

    AGrad = property(lambda self: self.A_grad)
    CGrad = property(lambda self: self.C_grad)
    PGrad = property(lambda self: self.P_grad)

    def setVelocityRHSGrad(self, t):
        self._set_vec("velocityRHS_grad", t, self._totalSize * self._dims)

    def CreateVelocityRHSGrad(self):
        self.setVelocityRHSGrad(self._vec(self._totalSize * self._dims))

    def setVelocityResultGrad(self, t):
        self._set_vec("velocityResult_grad", t, self._totalSize * self._dims)

    def CreateVelocityResultGrad(self):
        self.setVelocityResultGrad(self._vec(self._totalSize * self._dims))

    def setPressureRHSGrad(self, t):
        self._set_vec("pressureRHS_grad", t, self._totalSize * self._dims)

    def CreatePressureRHSGrad(self):
        self.setPressureRHSGrad(self._vec(self._totalSize * self._dims))

    def setPressureRHSdivGrad(self, t):
        self._set_vec("pressureRHSdiv_grad", t, self._totalSize)

    def CreatePressureRHSdivGrad(self):
        self.setPressureRHSdivGrad(self._vec(self._totalSize))

    def setPressureResultGrad(self, t):
        self._set_vec("pressureResult_grad", t, self._totalSize)

    def CreatePressureResultGrad(self):
        self.setPressureResultGrad(self._vec(self._totalSize))

    def setScalarRHSGrad(self, t):
        self._set_vec("scalarRHS_grad", t, self._totalSize * self._channels)

    def CreateScalarRHSGrad(self):
        self.setScalarRHSGrad(self._vec(self._totalSize * self._channels))

    def setScalarResultGrad(self, t):
        self._set_vec("scalarResult_grad", t, self._totalSize * self._channels)

    def CreateScalarResultGrad(self):
        self.setScalarResultGrad(self._vec(self._totalSize * self._channels))

    scalarRHSGrad = property(lambda self: self.scalarRHS_grad)
    scalarResultGrad = property(lambda self: self.scalarResult_grad)
    velocityRHSGrad = property(lambda self: self.velocityRHS_grad)
    velocityResultGrad = property(lambda self: self.velocityResult_grad)
    pressureRHSGrad = property(lambda self: self.pressureRHS_grad)
    pressureRHSdivGrad = property(lambda self: self.pressureRHSdiv_grad)
    pressureResultGrad = property(lambda self: self.pressureResult_grad)

    # --- setup ---
    def PrepareSolve(self):
        """Domain::PrepareSolve (domain_structs.cpp:2554-2676)."""
        _check(self.getNumBlocks() > 0, "Domain does not contain any blocks.")
        _check(self.viscosity.dtype == self._dtype, "Data type of viscosity does not match.")
        csr = 0
        total = 0
        for b in self.blocks:
            b.csrOffset = csr
            b.globalOffset = total
            csr += b.ComputeCSRSize()
            total += b.getStrides().w
            for fi, bd in enumerate(b.boundaries):
                if bd.type == BoundaryType.CONNECTED:
                    other = bd.getConnectedBlock()
                    _check(self._block_index(other) >= 0, "Connected block is not part of this domain.")
                    ob = other.getBoundary(bd.axes[0])
                    _check(ob.type == BoundaryType.CONNECTED, "Mismatch in block connection: connected block is not "
                           "connected at the target face.")
                    _check(ob.getConnectedBlock() is b, "Mismatch in block connection: connected block is connected "
                           "to a different block at the target face.")
                    _check(fi == ob.axes[0], "Mismatch in block connection: connected block is connected to a "
                           "different face of the source block.")
                    if self._dims == 2:
                        _check(bd.getConnectionAxisDirection(1) == ob.getConnectionAxisDirection(1),
                               "Direction of connection does not match.")
                elif bd.type == BoundaryType.PERIODIC:
                    _check(b.boundaries[fi ^ 1].type == BoundaryType.PERIODIC,
                           "Opposite boundary must also be periodic.")
        self._totalSize = total
        self._nnz = csr
        self.C = CSRmatrix(csr, total, self._dtype, self._device)
        # C and P share one pattern (identical by construction, PISO_multiblock_cuda_kernel.cu:3846, :4948)
        self.P = CSRmatrix(torch.zeros(csr, dtype=self._dtype, device=self._device), self.C.index, self.C.row)
        self.CreateA()
        self.C_grad = self.C.WithZeroValue()
        self.P_grad = self.P.WithZeroValue()
        if self.hasPassiveScalar():
            self.CreateScalarRHS()
            self.CreateScalarResult()
        self.CreateVelocityRHS()
        self.CreateVelocityResult()
        self.CreatePressureRHS()
        self.CreatePressureRHSdiv()
        self.CreatePressureResult()
        self._destroy_atlas()
        self._initialized = True
        self.UpdateDomainData()
        if self._device.type == "cuda":
            lib = _abi.load_library()
            _abi.check(lib.pict_build_csr_pattern(self._atlas, _stream(self._device)), "PrepareSolve")

    def _pack(self):
        """Descriptor of the domain (pict_domain_desc) and its block array; rebuilt only when a tensor address,
        an offset or a boundary object changed since the last call."""
        g = lambda t: 0 if (t is None or t.numel() == 0) else t.data_ptr()  # noqa: E731
        blocks = [b._pack(self._block_index) for b in self.blocks]
        ptrs = (self._totalSize, self._nnz, float(self.viscosity.detach()[0]), _ptr(self.C.row), _ptr(self.C.index),
                _ptr(self.C.value), _ptr(self.P.value), _ptr(self.A), _ptr(self.velocityRHS),
                _ptr(self.velocityResult), _ptr(self.pressureRHS), _ptr(self.pressureRHSdiv),
                _ptr(self.pressureResult), g(self.viscosity_grad),
                g(self.C_grad.value if self.C_grad is not None else None),
                g(self.P_grad.value if self.P_grad is not None else None), g(self.A_grad), g(self.velocityRHS_grad),
                g(self.velocityResult_grad), g(self.pressureRHS_grad), g(self.pressureRHSdiv_grad),
                g(self.pressureResult_grad))
        sv = self.passiveScalarViscosity
        sptrs = (self._channels, int(sv is not None and sv.size(0) == 1), g(sv),
                 g(getattr(self, "passiveScalarViscosity_grad", None)), g(self.scalarRHS), g(self.scalarResult),
                 g(self.scalarRHS_grad), g(self.scalarResult_grad))
        ptrs = ptrs + sptrs
        sig = (ptrs, tuple(id(b) for b in blocks))
        cached = getattr(self, "_packed", None)
        if cached is not None and cached[0] == sig:
            return cached[1], cached[2]
        n = len(self.blocks)
        arr = (_abi.BlockDesc * n)()
        for i, b in enumerate(blocks):
            arr[i] = b
        d = _abi.DomainDesc()
        d.num_dims = self._dims
        d.num_blocks = n
        d.num_cells = self._totalSize
        d.nnz = self._nnz
        d.dtype = _dtype_code(self._dtype)
        d.viscosity = ptrs[2]
        d.blocks = ctypes.cast(arr, ctypes.POINTER(_abi.BlockDesc))
        (d.csr_row, d.csr_index, d.C_value, d.P_value, d.A, d.velocity_rhs, d.velocity_result, d.pressure_rhs,
         d.pressure_rhs_div, d.pressure_result, d.viscosity_grad, d.C_grad, d.P_grad, d.A_grad, d.velocity_rhs_grad,
         d.velocity_result_grad, d.pressure_rhs_grad, d.pressure_rhs_div_grad, d.pressure_result_grad) = ptrs[3:22]
        (d.scalar_channels, d.scalar_viscosity_static, d.scalar_viscosity, d.scalar_viscosity_grad, d.scalar_rhs,
         d.scalar_result, d.scalar_rhs_grad, d.scalar_result_grad) = ptrs[22:]
        self._packed = (sig, d, arr)
        return d, arr

    def UpdateDomainData(self):
        """Domain::UpdateDomain -> UpdateDomainGPU (domain_structs.cpp:3031-3266)."""
        _check(self._initialized, "Domain is not initialized.")
        for b in self.blocks:
            for fb in b.boundaries:
                if fb.type == BoundaryType.FIXED and b.hasTransform() != fb.hasTransform():
                    pass  # the reference only warns (domain_structs.cpp:1678)
        desc, arr = self._pack()
        unchanged = desc is getattr(self, "_desc", None) and self._atlas is not None
        self._desc, self._blocks_arr = desc, arr
        if self._device.type == "cuda" and not unchanged:
            lib = _abi.load_library()
            stream = _stream(self._device)
            if self._atlas is None:
                h = ctypes.c_void_p()
                _abi.check(lib.pict_atlas_create(ctypes.byref(self._desc), ctypes.byref(h), stream), "UpdateDomainData")
                self._atlas = h
            else:
                _abi.check(lib.pict_atlas_update(self._atlas, ctypes.byref(self._desc), stream), "UpdateDomainData")
        self._set_changed(False)

    def _is_orthogonal_grid(self):
        """True if every cell metric has det * Minv_c . Minv_d == 0 for c != d: then the reference's
        non-orthogonal correction terms vanish identically (their coefficients are exactly zero,
        PISO_multiblock_cuda_kernel.cu:3753, :4902)."""
        key = tuple((b.transform.data_ptr(), b.transform._version) if b.hasTransform() else None
                    for b in self.blocks)
        if getattr(self, "_ortho_key", None) == key:
            return self._ortho
        ortho = True
        dims = self._dims
        for b in self.blocks:
            if not b.hasTransform() or dims == 1:
                continue
            T = b.transform
            Minv = T[..., dims * dims:2 * dims * dims].reshape(list(T.shape[:-1]) + [dims, dims])
            G = torch.matmul(Minv, Minv.transpose(-1, -2))
            off = G - torch.diag_embed(torch.diagonal(G, dim1=-2, dim2=-1))
            if bool((off != 0).any()):
                ortho = False
                break
        self._ortho_key = key
        self._ortho = ortho
        return ortho

    def _effective_non_ortho_flags(self, flags):
        flags = int(flags)
        if flags and self._is_orthogonal_grid():
            return 0
        return flags

    def IsTensorChanged(self):
        if self._changed or any(b.IsTensorChanged() for b in self.blocks):
            return True
        return (self.C is not None and self.C.IsTensorChanged()) or (self.P is not None and self.P.IsTensorChanged())

    def _set_changed(self, c):
        self._changed = c
        for b in self.blocks:
            b._set_changed(c)
        for m in (self.C, self.P, self.C_grad, self.P_grad):
            if m is not None:
                m._changed = c

    # --- queries ---
    def getMaxVelocity(self, withBounds=True, computational=False):
        return torch.amax(torch.stack([b.getMaxVelocity(withBounds, computational) for b in self.blocks]))

    def getMaxVelocityMagnitude(self, withBounds=True, computational=False):
        return torch.amax(torch.cat([b.getMaxVelocityMagnitude(withBounds, computational) for b in self.blocks]))\
            .reshape(1)

    def GetGlobalFluxBalance(self):
        """Net flux through all fixed boundaries (domain_structs.cpp:2722-2760)."""
        total = torch.zeros(1, dtype=self._dtype, device=self._device)
        for b in self.blocks:
            for face, fb in b.getFixedBoundaries():
                sign = 1.0 if face & 1 else -1.0
                total = total + sign * torch.sum(fb.GetFluxes())
        return total

    GetBoundaryFluxBalance = GetGlobalFluxBalance

    def CheckGlobalFluxBalance(self, eps=1e-5):
        return bool(torch.abs(self.GetGlobalFluxBalance()).item() < eps)

    CheckBoundaryFluxBalance = CheckGlobalFluxBalance

    # --- copies ---
    def Copy(self, newName=None):
        return self._copy(newName, clone=False)

    def Clone(self, newName=None):
        return self._copy(newName, clone=True)

    def _copy(self, newName, clone):
        c = (lambda t: None if t is None else t.clone()) if clone else (lambda t: t)
        d = Domain(self._dims, self.viscosity.clone() if clone else self.viscosity,
                   newName if newName is not None else self.name + ("_clone" if clone else "_copy"),
                   self._dtype, self._device, self._channels, _allow_cpu=self._device.type != "cuda")
        mapping = {}
        for b in self.blocks:
            nb = Block(d, b.name, velocity=c(b.velocity), pressure=c(b.pressure), passiveScalar=c(b.passiveScalar))
            if b._transform is not None:
                nb._transform = c(b._transform)
                nb._faceTransform = c(b._faceTransform)
                nb._vertexCoordinates = c(b._vertexCoordinates)
            if b.viscosity is not None:
                nb.setViscosity(c(b.viscosity))
            if b.velocitySource is not None:
                nb.setVelocitySource(c(b.velocitySource))
            d.AddBlock(nb)
            mapping[id(b)] = nb
        for b in self.blocks:
            nb = mapping[id(b)]
            for f, bd in enumerate(b.boundaries):
                if bd.type == BoundaryType.FIXED:
                    fb = FixedBoundary(c(bd.velocity), bd.velocityType, c(bd.passiveScalar), None, c(bd.transform), d)
                    nb.boundaries[f] = fb
                    fb._setSize(list(bd.getSizes()))
                elif bd.type == BoundaryType.CONNECTED:
                    nb.boundaries[f] = ConnectedBoundary(mapping[id(bd.getConnectedBlock())], list(bd.axes), d)
                else:
                    nb.boundaries[f] = PeriodicBoundary(d)
        return d

    def Detach(self):
        self.DetachFwd()
        self.DetachGrad()

    def DetachFwd(self):
        for b in self.blocks:
            b.DetachFwd()
        for n in ("A", "velocityRHS", "velocityResult", "pressureRHS", "pressureRHSdiv", "pressureResult",
                  "scalarRHS", "scalarResult"):
            t = getattr(self, n)
            if t is not None:
                setattr(self, n, t.detach())
        for m in (self.C, self.P):
            if m is not None:
                m.detach()

    def DetachGrad(self):
        for b in self.blocks:
            b.DetachGrad()

    def __str__(self):
        return 'Domain["%s", %dD, blocks=%d, cells=%d]' % (self.name, self._dims, len(self.blocks), self._totalSize)


# ----------------------------------------------------------------------------
# stage functions (PISOtorch.cpp:503-594)
# ----------------------------------------------------------------------------

def _check_domain(domain, timeStep=None):
    _check(domain.getNumBlocks() > 0, "Domain does not contain blocks.")
    _check(domain.IsInitialized(), "Domain is not initialized.")
    _check(not domain.IsTensorChanged(), "Domain's tensors have been changed, use UpdateDomain() to set the new "
           "pointers.")
    if domain.getDevice().type != "cuda":
        raise RuntimeError("PISOtorch stage functions run on the GPU only (domain device is %s)" % domain.getDevice())
    if timeStep is not None:
        _check(isinstance(timeStep, torch.Tensor) and timeStep.device.type == "cpu", "timeStep must be a host tensor")
        _check(timeStep.dim() == 1, "timeStep must be 1D.")
        _check(timeStep.size(0) == 1, "timeStep must be a scalar.")
        _check(timeStep.dtype == domain.getDtype(), "Data type of timeStep does not match.")
        return float(timeStep[0])
    return None


def _call(domain, name, *args):
    lib = _abi.load_library()
    _abi.check(getattr(lib, name)(domain._atlas, *args, _stream(domain.getDevice())), name)


def _check_scalar(domain, channel=None):
    """PISO_multiblock_cuda_kernel.cu:4513-4516."""
    _check(domain.hasPassiveScalar(), "domain does not have passive scalar.")
    if channel is not None:
        _check(0 <= channel < domain.getPassiveScalarChannels(), "Passive scalar channel index out of bounds.")


def SetupAdvectionMatrix(domain, timeStep, nonOrthoFlags, forPassiveScalar=False, passiveScalarChannel=0):
    dt = _check_domain(domain, timeStep)
    flags = domain._effective_non_ortho_flags(nonOrthoFlags)
    if forPassiveScalar:
        _check_scalar(domain, passiveScalarChannel)
        _call(domain, "pict_setup_advection_matrix_scalar", dt, flags, int(passiveScalarChannel))
        return
    _call(domain, "pict_setup_advection_matrix", dt, flags)


def SetupAdvectionScalar(domain, timeStep, nonOrthoFlags):
    dt = _check_domain(domain, timeStep)
    _check_scalar(domain)
    _call(domain, "pict_setup_advection_scalar", dt, domain._effective_non_ortho_flags(nonOrthoFlags))


def SetupAdvectionVelocity(domain, timeStep, nonOrthoFlags, applyPressureGradient=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_advection_velocity", dt, domain._effective_non_ortho_flags(nonOrthoFlags),
          int(bool(applyPressureGradient)))


def SetupPressureCorrection(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_pressure_correction", dt, domain._effective_non_ortho_flags(nonOrthoMode), int(bool(useFaceTransform)),
          int(bool(timeStepNorm)))


def SetupPressureMatrix(domain, timeStep, nonOrthoMode=0, useFaceTransform=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_pressure_matrix", dt, domain._effective_non_ortho_flags(nonOrthoMode), int(bool(useFaceTransform)))


def SetupPressureRHS(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_pressure_rhs", dt, domain._effective_non_ortho_flags(nonOrthoMode), int(bool(useFaceTransform)),
          int(bool(timeStepNorm)))


def SetupPressureRHSdiv(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_pressure_rhs_div", dt, domain._effective_non_ortho_flags(nonOrthoMode), int(bool(useFaceTransform)),
          int(bool(timeStepNorm)))


def CorrectVelocity(domain, timeStep, version=0, timeStepNorm=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_correct_velocity", dt, int(version), int(bool(timeStepNorm)))


def ComputeVelocityDivergence(domain):
    _check_domain(domain)
    out = torch.zeros(domain.getTotalSize(), dtype=domain.getDtype(), device=domain.getDevice())
    _call(domain, "pict_compute_velocity_divergence", ctypes.c_void_p(out.data_ptr()))
    return out


def _block_tensors(domain, g, channels):
    """Per-block tensors [1, channels, (z,) y, x] cut from a global component-major vector g [channels, N]."""
    out = []
    for b in domain.getBlocks():
        n = int(b.pressure.numel())
        out.append(g[:, b.globalOffset:b.globalOffset + n].reshape((1, channels) + tuple(b.pressure.shape[2:]))
                   .contiguous())
    return out


def ComputeSpatialVelocityGradients(domain):
    """Spatial gradients of the block velocities (PISOtorch.cpp:530; ComputeSpatialVelocityGradients
    PISO_multiblock_cuda_kernel.cu:6516-6538): for every block a list of D tensors shaped like its velocity, the
    d-th holding the gradient of velocity component d (channel i: d u_d / d x_i, in physical coordinates when the
    block has a transform).  Central differences; a fixed boundary counts at half a cell; across a connection the
    neighbour's component of the same index is used, as in the reference."""
    _check_domain(domain)
    D, N = domain.getSpatialDims(), domain.getTotalSize()
    g = torch.empty(D * D * N, dtype=domain.getDtype(), device=domain.getDevice())
    _call(domain, "pict_compute_spatial_velocity_gradients", ctypes.c_void_p(g.data_ptr()))
    g = g.view(D, D, N)
    per_comp = [_block_tensors(domain, g[d], D) for d in range(D)]
    return [[per_comp[d][bi] for d in range(D)] for bi in range(domain.getNumBlocks())]


def SGSviscosityIncompressibleSmagorinsky(domain, coefficient):
    """Additive viscosities of a Smagorinsky SGS model from the velocity field (PISOtorch.cpp:540-541;
    PISO_multiblock_cuda_kernel.cu:6898-7017): a list of one tensor per block shaped like its pressure, holding
    coefficient * Delta^2 * |S| with |S| = sqrt(2 S_ij S_ij) of the strain rate and Delta^2 the largest squared
    column norm of the block's transform M (1 without one).  `coefficient`: host tensor [1]."""
    _check_domain(domain)
    _check(isinstance(coefficient, torch.Tensor) and coefficient.device.type == "cpu",
           "coefficient must be a host tensor")
    _check(coefficient.dim() == 1 and coefficient.size(0) == 1, "coefficient tensor must have shape [1]")
    N = domain.getTotalSize()
    g = torch.empty(N, dtype=domain.getDtype(), device=domain.getDevice())
    _call(domain, "pict_sgs_viscosity_smagorinsky", ctypes.c_double(float(coefficient[0])),
          ctypes.c_void_p(g.data_ptr()))
    return _block_tensors(domain, g.view(1, N), 1)


def CopyVelocityResultToBlocks(domain):
    _check_domain(domain)
    _call(domain, "pict_copy_velocity_result_to_blocks")


def CopyVelocityResultFromBlocks(domain):
    _check_domain(domain)
    _call(domain, "pict_copy_velocity_result_from_blocks")


def CopyPressureResultToBlocks(domain):
    _check_domain(domain)
    _call(domain, "pict_copy_pressure_result_to_blocks")


def CopyPressureResultFromBlocks(domain):
    _check_domain(domain)
    _call(domain, "pict_copy_pressure_result_from_blocks")


def CopyScalarResultToBlocks(domain):
    _check_domain(domain)
    _check_scalar(domain)
    _call(domain, "pict_copy_scalar_result_to_blocks")


def CopyScalarResultFromBlocks(domain):
    _check_domain(domain)
    _check_scalar(domain)
    _call(domain, "pict_copy_scalar_result_from_blocks")


# ----------------------------------------------------------------------------
# backward stages (PISOtorch.cpp:596-621): read the *_grad tensors of a stage's outputs, write the *_grad tensors
# of its inputs; the caller creates those (zero-initialised) first, as PISOtorch_diff does
# ----------------------------------------------------------------------------

def SetupAdvectionMatrixGrad(domain, timeStep, nonOrthoFlags, forPassiveScalar=False, passiveScalarChannel=0):
    dt = _check_domain(domain, timeStep)
    flags = domain._effective_non_ortho_flags(nonOrthoFlags)
    if forPassiveScalar:
        _check_scalar(domain, passiveScalarChannel)
        _call(domain, "pict_setup_advection_matrix_scalar_grad", dt, flags, int(passiveScalarChannel))
        return
    _call(domain, "pict_setup_advection_matrix_grad", dt, flags)


def SetupAdvectionScalarGrad(domain, timeStep, nonOrthoFlags):
    dt = _check_domain(domain, timeStep)
    _check_scalar(domain)
    _call(domain, "pict_setup_advection_scalar_grad", dt, domain._effective_non_ortho_flags(nonOrthoFlags))


def SetupAdvectionVelocityGrad(domain, timeStep, nonOrthoFlags, applyPressureGradient=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_advection_velocity_grad", dt, domain._effective_non_ortho_flags(nonOrthoFlags),
          int(bool(applyPressureGradient)))


def SetupPressureCorrectionGrad(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_pressure_correction_grad", dt, domain._effective_non_ortho_flags(nonOrthoMode),
          int(bool(useFaceTransform)), int(bool(timeStepNorm)))


def SetupPressureMatrixGrad(domain, timeStep, nonOrthoMode=0, useFaceTransform=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_pressure_matrix_grad", dt, domain._effective_non_ortho_flags(nonOrthoMode),
          int(bool(useFaceTransform)))


def SetupPressureRHSGrad(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_pressure_rhs_grad", dt, domain._effective_non_ortho_flags(nonOrthoMode),
          int(bool(useFaceTransform)), int(bool(timeStepNorm)))


def SetupPressureRHSdivGrad(domain, timeStep, nonOrthoMode=0, useFaceTransform=False, timeStepNorm=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_setup_pressure_rhs_div_grad", dt, domain._effective_non_ortho_flags(nonOrthoMode),
          int(bool(useFaceTransform)), int(bool(timeStepNorm)))


def CorrectVelocityGrad(domain, timeStep, timeStepNorm=False):
    dt = _check_domain(domain, timeStep)
    _call(domain, "pict_correct_velocity_grad", dt, int(bool(timeStepNorm)))


def CopyVelocityResultGradToBlocks(domain):
    _check_domain(domain)
    _call(domain, "pict_copy_velocity_result_grad_to_blocks")


def CopyVelocityResultGradFromBlocks(domain):
    _check_domain(domain)
    _call(domain, "pict_copy_velocity_result_grad_from_blocks")


def CopyPressureResultGradFromBlocks(domain):
    _check_domain(domain)
    _call(domain, "pict_copy_pressure_result_grad_from_blocks")


def _subtract_mean(t):
    """t - mean(t) as a new tensor; the mean is accumulated in FP64 in a fixed order by the library (the pressure
    normalisation of the driver, PISOtorch_simulation.py:1243-1244)."""
    _check(t.is_cuda and t.dim() == 1, "a 1D device tensor is required")
    out = t.detach().clone().contiguous()
    if out.numel() > 0:
        _abi.check(_abi.load_library().pict_subtract_mean(ctypes.c_void_p(out.data_ptr()), out.numel(),
                                                          _dtype_code(out.dtype), _stream(out.device)), "mean")
    return out


def CopyScalarResultGradFromBlocks(domain):
    _check_domain(domain)
    _check_scalar(domain)
    _call(domain, "pict_copy_scalar_result_grad_from_blocks")


def CopyScalarResultGradToBlocks(domain):
    _check_domain(domain)
    _check_scalar(domain)
    _call(domain, "pict_copy_scalar_result_grad_to_blocks")


# ----------------------------------------------------------------------------
# linear solver (PISOtorch.cpp:559-563, PISO_multiblock_cuda_kernel.cu:7021-7103)
# ----------------------------------------------------------------------------

def _solve_opts(maxIterations, tolerance, convergenceCriterion, useBiCG, matrixRankDeficient, residualResetSteps,
                transposeA, printResidual, returnBestResult, BiCGwithPreconditioner, zero_rhs_shortcut=False):
    o = _abi.SolveOpts()
    o.max_iterations = int(maxIterations)
    o.convergence_criterion = int(convergenceCriterion)
    o.use_bicg = int(bool(useBiCG))
    o.matrix_rank_deficient = int(bool(matrixRankDeficient))
    o.residual_reset_steps = int(residualResetSteps)
    o.transpose_a = int(bool(transposeA))
    o.print_residual = int(bool(printResidual))
    o.return_best_result = int(bool(returnBestResult))
    o.bicg_with_preconditioner = int(bool(BiCGwithPreconditioner))
    o.zero_rhs_shortcut = int(bool(zero_rhs_shortcut))
    o.tolerance = float(tolerance)
    return o


# Solver diagnostics: when enabled, every solve appends
# (kind, n, nnz, nrhs, dtype bytes, [LinearSolverResultInfo], kernel ms, transposed, kernel name) -- kernel time
# from HIP events recorded by the library around its launch on the launch stream.
solve_log = None


def solver_kernel_name(A, nrhs, opts):
    """Name of the device kernel a solve of `A` with `nrhs` right-hand sides and `opts` (a SolveOpts) runs, as a
    profiler lists it."""
    buf = ctypes.create_string_buffer(128)
    desc = A._desc()
    _abi.check(_abi.load_library().pict_solver_kernel_name(ctypes.byref(desc), nrhs, ctypes.byref(opts), buf, 128),
               "solver_kernel_name")
    return buf.value.decode()


def enable_solve_log(enable=True):
    global solve_log
    solve_log = [] if enable else None
    _abi.check(_abi.load_library().pict_set_kernel_timing(1 if enable else 0))


def _solve(A, RHS, x, opts):
    n = A.getRows()
    _check(RHS.is_cuda and RHS.is_contiguous(), "RHS must be a contiguous CUDA tensor")
    _check(RHS.dim() == 1, "RHS must be a vector.")
    _check(RHS.size(0) % n == 0, "Size of RHS must be a multiple of the matrix size.")
    _check(RHS.dtype == A.getDtype(), "Data type of RHS must match A.")
    _check(x.is_cuda and x.is_contiguous(), "x must be a contiguous CUDA tensor")
    _check(x.dim() == 1, "x must be a vector.")
    _check(x.size(0) == RHS.size(0), "Size of x must match RHS.")
    _check(x.dtype == A.getDtype(), "Data type of x must match A.")
    nb = RHS.size(0) // n
    infos = (_abi.SolveInfo * nb)()
    if opts.transpose_a:
        A._ensure_transpose_perm()
    desc = A._desc()
    lib = _abi.load_library()
    _abi.check(lib.pict_solve_linear(ctypes.byref(desc), ctypes.c_void_p(RHS.data_ptr()),
                                     ctypes.c_void_p(x.data_ptr()), nb, ctypes.byref(opts), infos,
                                     _stream(RHS.device)), "SolveLinear")
    res = [LinearSolverResultInfo(i.final_residual, i.used_iterations, i.converged, i.is_finite_residual)
           for i in infos]
    if solve_log is not None:
        solve_log.append(("BiCG" if opts.use_bicg else "CG", n, A.getSize(), nb, RHS.element_size(), res,
                          lib.pict_last_solve_kernel_ms(), bool(opts.transpose_a), solver_kernel_name(A, nb, opts)))
    return res


def SolveLinear(A, RHS, x, maxIterations=1000, tolerance=1e-8, convergenceCriterion=ConvergenceCriterion.NORM2_NORMALIZED,
                useBiCG=False, matrixRankDeficient=False, residualResetSteps=0, transposeA=False, printResidual=False,
                returnBestResult=False, BiCGwithPreconditioner=True):
    """maxIterations: CPU int32 tensor [1] (or int); tolerance: CPU tensor [1] of the matrix dtype (or float)."""
    if isinstance(maxIterations, torch.Tensor):
        _check(maxIterations.device.type == "cpu" and maxIterations.dtype == torch.int32 and
               maxIterations.numel() == 1, "maxit must be a host int32 tensor [1].")
        maxIterations = int(maxIterations[0])
    if isinstance(tolerance, torch.Tensor):
        _check(tolerance.device.type == "cpu" and tolerance.numel() == 1, "tol must be a host tensor [1].")
        _check(tolerance.dtype == A.getDtype(), "Data type of tol must match A.")
        tolerance = float(tolerance[0])
    opts = _solve_opts(maxIterations, tolerance, convergenceCriterion, useBiCG, matrixRankDeficient,
                       residualResetSteps, transposeA, printResidual, returnBestResult, BiCGwithPreconditioner)
    return _solve(A, RHS, x, opts)


def SolveLinearRefined(A, RHS, x, maxIterations=1000, tolerance=1e-5, refineTolerance=1e-10, maxSweeps=8,
                       convergenceCriterion=ConvergenceCriterion.NORM2_NORMALIZED, useBiCG=False, nullSpace=None,
                       residualResetSteps=0):
    """Mixed-precision solve (no counterpart in the reference): x is an FP64 CUDA vector (initial guess in,
    solution out); A and RHS are in the matrix dtype.  Sweeps of r = b - A x in FP64 and a correction solve in the
    matrix dtype (maxIterations / tolerance / criterion / useBiCG apply to each correction solve, whose right-hand
    side is r scaled to unit RMS) until ||r||_2/||b||_2 < refineTolerance.  nullSpace: True if the constant vector
    spans A's null space (its component of b and of the residuals is projected out), False if not, None to detect
    it from the row sums.  Returns one LinearSolverResultInfo: finalResidual is the relative FP64 residual,
    usedIterations the correction iterations in total.  residualResetSteps applies to each correction solve."""
    n = A.getRows()
    _check(RHS.is_cuda and RHS.is_contiguous() and RHS.dim() == 1 and RHS.size(0) == n,
           "RHS must be a contiguous CUDA vector of the matrix size.")
    _check(RHS.dtype == A.getDtype(), "Data type of RHS must match A.")
    _check(x.is_cuda and x.is_contiguous() and x.dtype == torch.float64 and x.dim() == 1 and x.size(0) == n,
           "x must be a contiguous FP64 CUDA vector of the matrix size.")
    opts = _solve_opts(maxIterations, tolerance, convergenceCriterion, useBiCG, False, residualResetSteps, False,
                       False, False, False)
    info = _abi.SolveInfo()
    desc = A._desc()
    _abi.check(_abi.load_library().pict_solve_refined(ctypes.byref(desc), ctypes.c_void_p(RHS.data_ptr()),
                                                      ctypes.c_void_p(x.data_ptr()), ctypes.byref(opts),
                                                      float(refineTolerance), int(maxSweeps),
                                                      -1 if nullSpace is None else int(bool(nullSpace)),
                                                      ctypes.byref(info),
                                                      _stream(RHS.device)), "SolveLinearRefined")
    res = LinearSolverResultInfo(info.final_residual, info.used_iterations, info.converged, info.is_finite_residual)
    if solve_log is not None:  # one entry for all correction solves (their iterations and kernel time summed)
        lib = _abi.load_library()
        solve_log.append(("BiCG" if useBiCG else "CG", n, A.getSize(), 1, RHS.element_size(), [res],
                          lib.pict_last_solve_kernel_ms(), False, solver_kernel_name(A, 1, opts)))
    return res


def SparseOuterProduct(a, b, out_pattern):
    """out_pattern.value[k] = a[i] * b[j] on the pattern (cg_solver_kernel.cu:869-908)."""
    n = out_pattern.getRows()
    for t, nm in ((a, "a"), (b, "b")):
        _check(t.is_cuda and t.dim() == 1, "%s must be a CUDA vector." % nm)
        _check(t.size(0) == n, "Size of %s must match matrix size." % nm)
        _check(t.dtype == out_pattern.getDtype(), "Data type of %s must match C." % nm)
    desc = out_pattern._desc()
    lib = _abi.load_library()
    _abi.check(lib.pict_sparse_outer_product(ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                             ctypes.byref(desc), ctypes.c_void_p(out_pattern.value.data_ptr()),
                                             _stream(a.device)), "SparseOuterProduct")


# ----------------------------------------------------------------------------
# metrics (grid_gen.cu:284-514) and vector transforms
# ----------------------------------------------------------------------------

def _coords_args(coords):
    _check(coords.is_cuda and coords.is_contiguous(), "coords must be a contiguous CUDA tensor")
    _check(2 < coords.dim() < 6, "coords must have batch and channel dimension and be 1-3D.")
    _check(coords.size(0) == 1, "coords batch dimension must be 1.")
    dims = coords.dim() - 2
    _check(coords.size(1) == dims, "coords channel dimension must match spatial dimensionality.")
    vs = _spatial_size_xyz(coords, dims)
    return dims, (ctypes.c_int32 * 3)(*vs), vs


def CoordsToTransforms(coords):
    dims, vs_c, vs = _coords_args(coords)
    out = torch.zeros([1] + [vs[d] - 1 for d in reversed(range(dims))] + [2 * dims * dims + 1],
                      dtype=coords.dtype, device=coords.device)
    lib = _abi.load_library()
    _abi.check(lib.pict_coords_to_transforms(ctypes.c_void_p(coords.data_ptr()), dims, vs_c,
                                             _dtype_code(coords.dtype), ctypes.c_void_p(out.data_ptr()),
                                             _stream(coords.device)), "CoordsToTransforms")
    return out


def CoordsToFaceTransforms(coords):
    dims, vs_c, vs = _coords_args(coords)
    out = torch.zeros([1, dims] + [vs[d] for d in reversed(range(dims))] + [2 * dims * dims + 1],
                      dtype=coords.dtype, device=coords.device)
    lib = _abi.load_library()
    _abi.check(lib.pict_coords_to_face_transforms(ctypes.c_void_p(coords.data_ptr()), dims, vs_c,
                                                  _dtype_code(coords.dtype), ctypes.c_void_p(out.data_ptr()),
                                                  _stream(coords.device)), "CoordsToFaceTransforms")
    return out


# ----------------------------------------------------------------------------
# per-cell vector / matrix operations (matrix_vector_ops.cu, matrix_vector_ops_grads.cu; bindings
# PISOtorch.cpp:604-627): vectors or flat row-major square matrices in the channel dimension of NCDHW tensors
# ----------------------------------------------------------------------------

def _vm_check(t, name, allow_vector=True, allow_matrix=True):
    _check(isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous(), "%s must be a contiguous CUDA tensor" % name)
    _check(2 < t.dim() < 6, "%s must have batch and channel dimension and be 1-3D." % name)
    _check(t.size(0) == 1, "%s batch dimension must be 1." % name)
    _check(t.dtype in (torch.float32, torch.float64), "%s must be float32 or float64." % name)
    dims = t.dim() - 2
    ok = (allow_vector and t.size(1) == dims) or (allow_matrix and t.size(1) == dims * dims)
    return dims, ok


def _matmul_args(vectorMatrixA, vectorMatrixB):
    dims, ok = _vm_check(vectorMatrixA, "vectorMatrixA")
    _check(ok, "vectorMatrixA channel dimension must be a vector or flat square matrix that matches spatial "
               "dimensionality.")
    _check(isinstance(vectorMatrixB, torch.Tensor) and vectorMatrixB.dim() == dims + 2,
           "vectorMatrixB dimensionality must match vectorMatrixA.")
    _, okb = _vm_check(vectorMatrixB, "vectorMatrixB")
    _check(okb, "vectorMatrixB channel dimension must be a vector or flat square matrix that matches spatial "
                "dimensionality.")
    _check(vectorMatrixA.dtype == vectorMatrixB.dtype, "Data type of vectorMatrixB does not match vectorMatrixA.")
    _check(vectorMatrixA.shape[2:] == vectorMatrixB.shape[2:], "vectorMatrixB spatial size must match vectorMatrixA.")
    _check(vectorMatrixA.device == vectorMatrixB.device, "vectorMatrixB device must match vectorMatrixA.")
    # with one dimension a channel is both a vector and a 1x1 matrix: it counts as a matrix
    a_mat = vectorMatrixA.size(1) == dims * dims
    b_mat = vectorMatrixB.size(1) == dims * dims
    out_ch = dims * dims if (a_mat and b_mat) else dims if (a_mat or b_mat) else 1
    return dims, a_mat, b_mat, out_ch


def _matmul_flags(transposeA, invertA, transposeB, invertB, transposeOutput, invertOutput):
    return ((_abi.PICT_MATMUL_TRANSPOSE_A if transposeA else 0) | (_abi.PICT_MATMUL_INVERT_A if invertA else 0) |
            (_abi.PICT_MATMUL_TRANSPOSE_B if transposeB else 0) | (_abi.PICT_MATMUL_INVERT_B if invertB else 0) |
            (_abi.PICT_MATMUL_TRANSPOSE_OUT if transposeOutput else 0) |
            (_abi.PICT_MATMUL_INVERT_OUT if invertOutput else 0))


def matmul(vectorMatrixA, vectorMatrixB, transposeA=False, invertA=False, transposeB=False, invertB=False,
           transposeOutput=False, invertOutput=False):
    """Per-cell product matrix/vector x matrix/vector (PISOtorch.cpp:604-611, matrix_vector_ops.cu:113-200): a matrix
    if both are matrices, a vector if one is, a scalar if both are vectors.  transpose/invert apply to matrices only:
    an operand is loaded (transposed), then inverted; the product is inverted, then stored (transposed)."""
    dims, a_mat, b_mat, out_ch = _matmul_args(vectorMatrixA, vectorMatrixB)
    out = torch.empty([1, out_ch] + list(vectorMatrixA.shape[2:]), dtype=vectorMatrixA.dtype,
                      device=vectorMatrixA.device)
    cells = out[0, 0].numel()
    lib = _abi.load_library()
    _abi.check(lib.pict_matmul(_ptr(vectorMatrixA), int(a_mat), _ptr(vectorMatrixB), int(b_mat), _ptr(out), dims,
                               cells, _dtype_code(out.dtype),
                               _matmul_flags(transposeA, invertA, transposeB, invertB, transposeOutput, invertOutput),
                               _stream(out.device)), "matmul")
    return out


def matmulGrad(vectorMatrixA, vectorMatrixB, outputGrad, transposeA=False, invertA=False, transposeB=False,
               invertB=False, transposeOutput=False, invertOutput=False):
    """Gradients of matmul (PISOtorch.cpp:613-621, matrix_vector_ops_grads.cu:217-310): [gradA, gradB].  An inverted
    operand's gradient is zero; a matrix product with an inverted output, or with both operands inverted, has zero
    gradients."""
    dims, a_mat, b_mat, out_ch = _matmul_args(vectorMatrixA, vectorMatrixB)
    _check(isinstance(outputGrad, torch.Tensor) and outputGrad.is_cuda and outputGrad.dim() == dims + 2,
           "outputGrad dimensionality must match vectorMatrixA.")
    _check(outputGrad.size(0) == 1, "outputGrad batch dimension must be 1.")
    _check(outputGrad.size(1) == out_ch, "outputGrad channel dimension must match operation type.")
    _check(outputGrad.dtype == vectorMatrixA.dtype, "Data type of outputGrad does not match vectorMatrixA.")
    _check(outputGrad.shape[2:] == vectorMatrixA.shape[2:], "outputGrad spatial size must match vectorMatrixA.")
    g = outputGrad.contiguous()
    ga = torch.empty_like(vectorMatrixA)
    gb = torch.empty_like(vectorMatrixB)
    cells = ga[0, 0].numel()
    lib = _abi.load_library()
    _abi.check(lib.pict_matmul_grad(_ptr(vectorMatrixA), int(a_mat), _ptr(vectorMatrixB), int(b_mat), _ptr(g),
                                    _ptr(ga), _ptr(gb), dims, cells, _dtype_code(ga.dtype),
                                    _matmul_flags(transposeA, invertA, transposeB, invertB, transposeOutput,
                                                  invertOutput), _stream(ga.device)), "matmulGrad")
    return [ga, gb]


def VectorToDiagMatrix(vectors):
    """Matrices with the vectors on their diagonals (PISOtorch.cpp:622-623, matrix_vector_ops.cu:203-262)."""
    dims, ok = _vm_check(vectors, "vectors", allow_matrix=False)
    _check(ok, "vectors channel dimension must match spatial dimensionality.")
    out = torch.empty([1, dims * dims] + list(vectors.shape[2:]), dtype=vectors.dtype, device=vectors.device)
    lib = _abi.load_library()
    _abi.check(lib.pict_vector_to_diag_matrix(_ptr(vectors), _ptr(out), dims, out[0, 0].numel(),
                                              _dtype_code(out.dtype), _stream(out.device)), "VectorToDiagMatrix")
    return out


def InvertMatrix(matrices, inPlace=False):
    """Per-cell inverse of square matrices (PISOtorch.cpp:624-625, matrix_vector_ops.cu:265-303)."""
    dims, ok = _vm_check(matrices, "matrices", allow_vector=False)
    _check(ok, "matrices channel dimension must match spatial dimensionality.")
    out = matrices if inPlace else torch.empty_like(matrices)
    lib = _abi.load_library()
    _abi.check(lib.pict_invert_matrix(_ptr(matrices), _ptr(out), dims, out[0, 0].numel(), _dtype_code(out.dtype),
                                      _stream(out.device)), "InvertMatrix")
    return out


def TransformVectors(vectors, transforms, inverse):
    """Per-cell M v (or Minv v) for NCDHW vectors and N(DHW)T transforms (k_TransformVectors :7152-7176)."""
    dims = vectors.size(1)
    off = dims * dims if inverse else 0
    M = transforms[..., off:off + dims * dims]
    M = M.reshape(list(M.shape[:-1]) + [dims, dims])
    v = torch.movedim(vectors, 1, -1).unsqueeze(-1)
    r = torch.matmul(M, v).squeeze(-1)
    return torch.movedim(r, -1, 1).contiguous()
