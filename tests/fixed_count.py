"""A simulation whose solves run given iteration counts (tolerance 0), for comparisons that must not depend on on
which side of a stopping tolerance a solve lands: near the tolerance one more CG iteration moves a lid cavity's
velocity by about 1e-3 relative at tolerance 1e-6, so two runs that differ in rounding only can differ by that much
once one of their solves stops an iteration earlier.  With the counts of a reference run fixed, a rollout is a
smooth function of its inputs (finite differences) and two precisions are compared at equal work."""
import torch

from pict_amd import PISOtorch as P
from pict_amd import PISOtorch_simulation as S


class FixedCountSim(S.Simulation):
    """Stage-path Simulation; `counts` lists the used iterations of each solve and right-hand side in the
    driver's order (advection components, then one pressure solve per corrector), as the solvers report them: CG
    the index of its last iteration (-1 for none), BiCGStab its full iterations."""

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.native_step = False
        self.counts = []

    def linear_solve(self, csrMat, rhs, x=None, use_BiCG=False, **kw):
        if x is None:
            x = torch.zeros_like(rhs)
        n = csrMat.getRows()
        for k in range(rhs.numel() // n):
            it = self.counts.pop(0) + (0 if use_BiCG else 1)
            if it > 0:
                P.SolveLinear(csrMat, rhs[k * n:(k + 1) * n], x[k * n:(k + 1) * n], maxIterations=it,
                              tolerance=0.0, useBiCG=use_BiCG, BiCGwithPreconditioner=False)
        return x, True


def counted_steps(sim, steps):
    """`steps` steps of `sim`; returns the used iterations of every solve in order."""
    sim.record_solver_infos = True
    sim.solver_infos = []
    assert sim._PISO_split_step(steps)
    sim.record_solver_infos = False
    return [i.usedIterations for _, infos in sim.solver_infos for i in infos]
