"""Shared helpers of the gradient tests: the tracked inputs and outputs of each PISO stage (names as in the
reference's PISOtorch_diff.flatten_domain, PISOtorch_diff.py:44-125) and their gradient tensors."""


def input_tensors(d, name):
    """(value tensor, gradient getter) pairs for a tracked input name (PISOtorch_diff.flatten_domain names)."""
    out = []
    for b in d.blocks:
        if name == "VELOCITY":
            out.append((b.velocity, lambda b=b: b.velocityGrad))
        elif name == "VELOCITY_SOURCE" and b.velocitySource is not None:
            out.append((b.velocitySource, lambda b=b: b.velocitySourceGrad))
        elif name == "VISCOSITY_BLOCK" and b.viscosity is not None:
            out.append((b.viscosity, lambda b=b: b.viscosityGrad))
        elif name == "PRESSURE":
            out.append((b.pressure, lambda b=b: b.pressureGrad))
        elif name == "BOUNDARY_VELOCITY":
            for _, fb in b.getFixedBoundaries():
                out.append((fb.velocity, lambda fb=fb: fb.velocityGrad))
        elif name == "PASSIVE_SCALAR" and b.passiveScalar is not None:
            out.append((b.passiveScalar, lambda b=b: b.passiveScalarGrad))
        elif name == "BOUNDARY_PASSIVE_SCALAR":
            for _, fb in b.getFixedBoundaries():
                if fb.passiveScalar is not None:
                    out.append((fb.passiveScalar, lambda fb=fb: fb.passiveScalarGrad))
    if name == "VISCOSITY":
        out.append((d.viscosity, lambda: d.viscosityGrad))
    elif name == "VELOCITY_RESULT":
        out.append((d.velocityResult, lambda: d.velocityResultGrad))
    elif name == "A":
        out.append((d.A, lambda: d.AGrad))
    elif name == "C":
        out.append((d.C.value, lambda: d.CGrad.value))
    elif name == "PRESSURE_RHS":
        out.append((d.pressureRHS, lambda: d.pressureRHSGrad))
    elif name == "PRESSURE_RESULT":
        out.append((d.pressureResult, lambda: d.pressureResultGrad))
    elif name == "PASSIVE_SCALAR_VISCOSITY" and d.passiveScalarViscosity is not None:
        out.append((d.passiveScalarViscosity, lambda: d.passiveScalarViscosityGrad))
    elif name == "PASSIVE_SCALAR_RESULT" and d.hasPassiveScalar():
        out.append((d.scalarResult, lambda: d.scalarResultGrad))
    return out


OUTPUTS = {
    "C": (lambda d: d.C.value, lambda d, t: d.CGrad.setValue(t)),
    "A": (lambda d: d.A, lambda d, t: d.setAGrad(t)),
    "P": (lambda d: d.P.value, lambda d, t: d.PGrad.setValue(t)),
    "VELOCITY_RHS": (lambda d: d.velocityRHS, lambda d, t: d.setVelocityRHSGrad(t)),
    "PRESSURE_RHS": (lambda d: d.pressureRHS, lambda d, t: d.setPressureRHSGrad(t)),
    "PRESSURE_RHS_DIV": (lambda d: d.pressureRHSdiv, lambda d, t: d.setPressureRHSdivGrad(t)),
    "VELOCITY_RESULT": (lambda d: d.velocityResult, lambda d, t: d.setVelocityResultGrad(t)),
    "PASSIVE_SCALAR_RHS": (lambda d: d.scalarRHS, lambda d, t: d.setScalarRHSGrad(t)),
}

def create_grads(d):
    """Zero gradient tensors for every input (PISOtorch_diff's Create*Grad calls before a backward stage)."""
    d.CreateAGrad()
    d.CGrad.CreateValue()
    d.PGrad.CreateValue()
    d.CreateViscosityGrad()
    d.CreateVelocityGradOnBlocks()
    d.CreateVelocitySourceGradOnBlocks()
    d.CreateVelocityGradOnBoundaries()
    d.CreatePressureGradOnBlocks()
    d.CreateVelocityRHSGrad()
    d.CreateVelocityResultGrad()
    d.CreatePressureRHSGrad()
    d.CreatePressureRHSdivGrad()
    d.CreatePressureResultGrad()
    if d.hasPassiveScalar():
        d.CreatePassiveScalarGradOnBlocks()
        d.CreatePassiveScalarGradOnBoundaries()
        d.CreateScalarRHSGrad()
        d.CreateScalarResultGrad()


