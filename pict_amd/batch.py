"""Independent simulations stepped together (pict_batch_*; no counterpart in the reference, which steps one Domain
per PISOtorch_simulation.Simulation).

    sims = [PISOtorch_simulation.Simulation(domain, time_step=dt, ...) for domain in domains]
    batch = SimulationBatch(sims)
    batch.tune()      # optional: the launch shapes that are fastest on this batch
    batch.run(100)    # each simulation advances as by `sim._PISO_split_step(1)` per step

Every simulation must be eligible for the native step (Simulation._native_step_eligible: orthogonal, not
differentiable, no hooks) and all must share their step options except the time step and the domain itself
(sizes, block structure, viscosity and boundaries may differ).  One batched step runs each stage as one launch
over the cells of every domain and each solve as one launch over every system; its result for each
simulation is the result of that simulation's own native step, up to the reduction order of the Krylov dot
products (a system may get another workgroup count).
"""
import ctypes

import numpy as np

from . import PISOtorch, _abi

_INFO_DTYPE = np.dtype([("final_residual", np.float64), ("used_iterations", np.int32), ("converged", np.int32),
                        ("is_finite_residual", np.int32), ("reserved", np.int32)])


def _opts_key(o):
    """Bytes of the step options without the time step (the batch carries one time step per domain)."""
    c = _abi.StepOpts.from_buffer_copy(o)
    c.time_step = 0.0
    return bytes(c)


class SimulationBatch:
    def __init__(self, sims, config=None):
        self.sims = list(sims)
        if not self.sims:
            raise ValueError("a batch needs at least one simulation")
        for i, s in enumerate(self.sims):
            s._check_domain()
            if not s._native_step_eligible():
                raise PISOtorch.PictUnsupportedError(
                    "simulation %d cannot run the native step (see Simulation._native_step_eligible)" % i)
        d0 = self.sims[0].domain
        self.dims = d0.getSpatialDims()
        self.dtype = d0.getDtype()
        self.device = d0.getDevice()
        self.corrector_steps = self.sims[0].corrector_steps
        # the time step in the domain dtype, as the native step takes it from Simulation._get_time_step_torch
        self._np_dtype = np.float64 if str(self.dtype) == "torch.float64" else np.float32
        key = _opts_key(self.sims[0]._native_opts(self.sims[0]._get_time_step_torch()))
        for i, s in enumerate(self.sims):
            d = s.domain
            if d.getSpatialDims() != self.dims or d.getDtype() != self.dtype or d.getDevice() != self.device:
                raise ValueError("simulation %d: dimension, dtype or device differs from simulation 0" % i)
            if _opts_key(s._native_opts(s._get_time_step_torch())) != key:
                raise ValueError("simulation %d: step options differ from simulation 0" % i)
            d.UpdateDomainData()
        self._lib = _abi.load_library()
        n = len(self.sims)
        atlases = (ctypes.c_void_p * n)(*[s.domain._atlas for s in self.sims])
        handle = ctypes.c_void_p()
        _abi.check(self._lib.pict_batch_create(atlases, n, ctypes.byref(handle)), "batch create")
        self._handle = handle
        self._atlases = atlases
        self._atlas_ids = [s.domain._atlas for s in self.sims]
        per = self.dims + self.corrector_steps
        self._infos = (_abi.SolveInfo * (n * per))()
        self._dts = (ctypes.c_double * n)()
        if config:
            self.set_config(**config)

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            self._lib.pict_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.sims)

    @property
    def total_cells(self):
        return sum(s.domain.getTotalSize() for s in self.sims)

    def config(self):
        c = _abi.BatchConfig()
        _abi.check(self._lib.pict_batch_get_config(self._handle, ctypes.byref(c)), "batch config")
        return {f: getattr(c, f) for f, _ in c._fields_}

    def set_config(self, **kw):
        c = _abi.BatchConfig()
        _abi.check(self._lib.pict_batch_get_config(self._handle, ctypes.byref(c)), "batch config")
        for k, v in kw.items():
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, int(v))
        _abi.check(self._lib.pict_batch_set_config(self._handle, ctypes.byref(c)), "batch config")

    def describe(self):
        i = _abi.BatchInfo()
        _abi.check(self._lib.pict_batch_describe(self._handle, ctypes.byref(i)), "batch describe")
        out = {f: getattr(i, f) for f, _ in i._fields_}
        out["advection_kernel"] = i.advection_kernel.decode()
        out["pressure_kernel"] = i.pressure_kernel.decode()
        return out

    def set_timing(self, enable):
        """Diagnostics: HIP events around every solve launch (steps then run without the graph)."""
        _abi.check(self._lib.pict_batch_set_timing(self._handle, int(bool(enable))), "batch timing")

    def last_solve_ms(self):
        """Kernel times of the last step's solves (advection, then one per corrector), milliseconds."""
        ms = (ctypes.c_double * (1 + self.corrector_steps))()
        n = self._lib.pict_batch_last_solve_ms(self._handle, ms, len(ms))
        return list(ms)[:n]

    def _prepare(self):
        for i, s in enumerate(self.sims):
            # the setters flag a changed domain (IsTensorChanged); an unchanged one keeps its uploaded descriptor
            if s.domain.IsTensorChanged():
                s.domain.UpdateDomainData()
            if s.domain._atlas != self._atlas_ids[i]:
                # the batch holds the library's view of each domain; a re-prepared domain has a new one
                raise RuntimeError("simulation %d: its domain was prepared again; create a new batch" % i)
            self._dts[i] = float(self._np_dtype(s._time_step))
        return self.sims[0]._native_opts(self.sims[0]._get_time_step_torch())

    def tune(self, repeats=3):
        """Time the candidate launch shapes on the current state and keep the fastest (block velocities and
        pressures are left as they are).  Returns the chosen configuration."""
        o = self._prepare()
        c = _abi.BatchConfig()
        _abi.check(self._lib.pict_batch_tune(self._handle, ctypes.byref(o), self._dts, int(repeats), ctypes.byref(c),
                                             PISOtorch._stream(self.device)), "batch tune")
        return {f: getattr(c, f) for f, _ in c._fields_}

    def step(self):
        """One PISO step of every simulation.  Returns the solver infos as a structured array
        [simulation, num_dims + corrector_steps] (advection solves, then one pressure solve per corrector)."""
        o = self._prepare()
        _abi.check(self._lib.pict_batch_step(self._handle, ctypes.byref(o), self._dts, self._infos,
                                             PISOtorch._stream(self.device)), "batch step")
        per = self.dims + self.corrector_steps
        arr = np.frombuffer(self._infos, dtype=_INFO_DTYPE).reshape(len(self.sims), per).copy()
        bad = (arr["converged"] == 0) | (arr["is_finite_residual"] == 0)
        logging = PISOtorch.solve_log is not None
        for i, s in enumerate(self.sims):
            if logging or s.record_solver_infos or bad[i].any():
                if s.record_solver_infos:
                    s.solver_infos = []
                s._native_record(self._infos[i * per:(i + 1) * per], o)
            s.end_step((self._dts[i],))
        return arr

    def run(self, steps):
        infos = None
        for _ in range(int(steps)):
            infos = self.step()
        return infos
