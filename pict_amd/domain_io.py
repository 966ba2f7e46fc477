"""Saving and loading a domain as `<path>.npz` (the tensors) and `<path>.json` (the structure), in the file format of
the reference's lib/util/domain_io.py:44-236 (same JSON keys, tensors numbered in order of first use and shared
tensors stored once), so that files move between the two.

Loading builds the domain through the public construction API (CreateBlock, CloseBoundary, ConnectBlock,
MakePeriodic) and leaves PrepareSolve to the caller, as the reference does: it would allocate the matrices and the
global vectors."""
import json
import warnings

import numpy as np
import torch

from . import PISOtorch as P

_DTYPES = {torch.float32: "float32", torch.float64: "float64"}
_FACES = ["-x", "+x", "-y", "+y", "-z", "+z"]


def _bc_name(t):
    if t == P.BoundaryConditionType.DIRICHLET:
        return "DIRICHLET"
    if t == P.BoundaryConditionType.NEUMANN:
        return "NEUMANN"
    raise TypeError("Unsupported boundary condition type.")


def _bc_type(name):
    try:
        return {"DIRICHLET": P.BoundaryConditionType.DIRICHLET, "NEUMANN": P.BoundaryConditionType.NEUMANN}[name]
    except KeyError:
        raise TypeError("Unsupported boundary condition type.") from None


class _Tensors:
    """The flat tensor list of a save: each distinct tensor object once, referenced by its index (as a string)."""

    def __init__(self):
        self.items = []

    def ref(self, t):
        for i, u in enumerate(self.items):
            if u is t:
                return str(i)
        if t.dtype not in _DTYPES:
            raise TypeError("Unsupported dtype.")
        self.items.append(t)
        return str(len(self.items) - 1)


def save_domain(domain, path):
    """Writes `path`.npz and `path`.json (`path` without extension)."""
    data = _Tensors()
    blocks = domain.getBlocks()
    d = {"name": domain.name, "spatialDims": domain.getSpatialDims(), "viscosity": data.ref(domain.viscosity),
         "passiveScalarChannels": domain.getPassiveScalarChannels()}
    if domain.hasPassiveScalarViscosity():
        d["passiveScalarViscosity"] = data.ref(domain.passiveScalarViscosity)
    d["blocks"] = []
    for block in blocks:
        bd = {"name": block.name}
        if block.hasViscosity():
            bd["viscosity"] = data.ref(block.viscosity)
        if block.hasPassiveScalarViscosity():
            bd["passiveScalarViscosity"] = data.ref(block.passiveScalarViscosity)
        bd["velocity"] = data.ref(block.velocity)
        bd["pressure"] = data.ref(block.pressure)
        if block.hasPassiveScalar():
            bd["scalar"] = data.ref(block.passiveScalar)
        if block.hasVelocitySource():
            bd["velocitySource"] = data.ref(block.velocitySource)
        if block.hasVertexCoordinates():
            bd["vertexCoordinates"] = data.ref(block.vertexCoordinates)
        elif block.hasTransform():
            bd["transform"] = data.ref(block.transform)
            if block.hasFaceTransform():
                bd["faceTransform"] = data.ref(block.faceTransform)
        bd["boundaries"] = []
        for face in range(2 * block.getSpatialDims()):
            b = block.getBoundary(face)
            if b.type == P.BoundaryType.FIXED:
                e = {"type": "FIXED", "velocityType": _bc_name(b.velocityType), "velocity": data.ref(b.velocity)}
                if b.hasPassiveScalar():
                    e["passiveScalarType"] = [_bc_name(t) for t in b.passiveScalarTypes]
                    e["scalar"] = data.ref(b.passiveScalar)
                if b.hasTransform():
                    e["transform"] = data.ref(b.transform)
            elif b.type == P.BoundaryType.CONNECTED:
                e = {"type": "CONNECTED", "connectedBlock": blocks.index(b.getConnectedBlock()),
                     "axes": [int(a) for a in b.axes]}
            elif b.type == P.BoundaryType.PERIODIC:
                e = {"type": "PERIODIC"}
            else:
                raise TypeError("Unknown boundary type.")
            bd["boundaries"].append(e)
        d["blocks"].append(bd)
    d["data_info"] = {str(i): {"shape": list(t.shape), "dtype": _DTYPES[t.dtype], "device": t.device.type}
                      for i, t in enumerate(data.items)}
    np.savez_compressed(path + ".npz", **{str(i): t.detach().cpu().numpy() for i, t in enumerate(data.items)})
    with open(path + ".json", "w") as fh:
        json.dump(d, fh)


def load_domain(path, dtype=None, device=None, with_scalar=True):
    """The domain saved at `path` (without extension), in `dtype` (default: the saved one) on `device` (default: the
    device each tensor was saved from; the domain viscosity stays a host tensor).  with_scalar=False drops the
    passive scalars."""
    with open(path + ".json") as fh:
        d = json.load(fh)
    info = d["data_info"]
    with np.load(path + ".npz") as z:
        arrays = [z[str(i)] for i in range(len(z.files))]
    saved_dtype = {"float32": torch.float32, "float64": torch.float64}[info["0"]["dtype"]] if info else torch.float32
    dtype = dtype or saved_dtype
    tensors = []
    for i, a in enumerate(arrays):
        dev = device if device is not None else info[str(i)]["device"]
        tensors.append(torch.from_numpy(a).to(device=dev, dtype=dtype).contiguous())

    def get(entry, key):
        return tensors[int(entry[key])] if key in entry else None

    dev = torch.device(device) if device is not None else None
    if dev is None:
        cells = [get(b, "velocity") for b in d["blocks"]]
        dev = cells[0].device if cells else torch.device("cuda")
    viscosity = get(d, "viscosity").cpu()
    channels = d.get("passiveScalarChannels", 1) if with_scalar else 0
    domain = P.Domain(d["spatialDims"], viscosity, d["name"], dtype=dtype, device=dev, passiveScalarChannels=channels,
                      scalarViscosity=get(d, "passiveScalarViscosity") if with_scalar else None,
                      _allow_cpu=dev.type == "cpu")
    for bd in d["blocks"]:
        coords = get(bd, "vertexCoordinates")
        block = domain.CreateBlock(get(bd, "velocity"), get(bd, "pressure"),
                                   get(bd, "scalar") if with_scalar else None, vertexCoordinates=coords,
                                   name=bd["name"])
        if "viscosity" in bd:
            block.setViscosity(get(bd, "viscosity"))
        if with_scalar and "passiveScalarViscosity" in bd:
            raise NotImplementedError("per-block passive-scalar viscosity is not supported")
        if "velocitySource" in bd:
            block.setVelocitySource(get(bd, "velocitySource"))
        if coords is None and "transform" in bd:
            block.setTransform(get(bd, "transform"), get(bd, "faceTransform"))
    blocks = domain.getBlocks()
    for bd, block in zip(d["blocks"], blocks):
        for face, e in enumerate(bd["boundaries"]):
            kind = e["type"]
            if kind in ("DIRICHLET", "DIRICHLET_VARYING"):
                warnings.warn("%s boundaries are deprecated, creating FixedBoundary instead." % kind)
                block.CloseBoundary(face, get(e, "velocity"), get(e, "scalar") if with_scalar else None)
            elif kind == "FIXED":
                scalar = get(e, "scalar") if with_scalar else None
                block.CloseBoundary(face, get(e, "velocity"), scalar)
                if scalar is not None:
                    types = e["passiveScalarType"]
                    block.getBoundary(face).setPassiveScalarType(
                        [_bc_type(t) for t in types] if isinstance(types, list) else _bc_type(types))
            elif kind == "CONNECTED":
                block.ConnectBlock(_FACES[face], blocks[e["connectedBlock"]], *[_FACES[a] for a in e["axes"]])
            elif kind == "PERIODIC":
                block.MakePeriodic(face // 2)
            else:
                raise TypeError("Unknown boundary type: " + kind)
    return domain
