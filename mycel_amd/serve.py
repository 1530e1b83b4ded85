"""Job server: continuous batching of FEA solves and SIMP runs on one GPU (C ABI mycel_server_*).

Jobs are admitted as their HBM footprint fits the budget and a worker stream is free; jobs that arrive
while others run overlap with them on the device. Results do not depend on the mix.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from .fem import FeModel
from .optimization import OptimizerOptions, simp_options_c

JOB_FEA = 0
JOB_SIMP = 1
QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3


@dataclass
class JobInfo:
    state: int
    status: int
    compliance: float
    max_displacement: float
    pcg_iterations: int
    simp_iterations: int
    converged: bool
    footprint_bytes: float
    queue_ms: float
    run_ms: float
    latency_ms: float


def _info(i: L.JobInfoC) -> JobInfo:
    return JobInfo(i.state, i.status, i.compliance, i.max_displacement, i.pcg_iterations, i.simp_iterations,
                   bool(i.converged), i.footprint_bytes, i.queue_ms, i.run_ms, i.latency_ms)


def _simp_c(o: OptimizerOptions) -> L.SimpOptionsC:
    return simp_options_c(o)


class JobServer:
    def __init__(self, hbm_budget_bytes: float = 0.0, workers: int = 4):
        h = C.c_void_p()
        L.check(L.load().mycel_server_create(float(hbm_budget_bytes), int(workers), C.byref(h)))
        self._h = h
        self._sizes = {}

    def close(self):
        if getattr(self, "_h", None):
            L.load().mycel_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _submit(self, kind, model: FeModel, forces, tol=1e-8, max_iter=20000, precision=L.PRECISION_FP64,
                options: Optional[OptimizerOptions] = None):
        dom = L.u8(model.Domain)
        fx = L.i32(model.FixedDofs)
        f = L.f64(forces)
        patch = (np.concatenate([lc.Patch.Nodes for lc in model.Loads]).astype(np.int32)
                 if model.Loads else np.zeros(1, np.int32))
        den = L.f64(model.ElementDensity)
        d = L.JobDescC()
        d.kind = kind
        d.grid = model.Grid._c()
        d.domain = L.ptr(dom)
        d.fixed_dofs = L.ptr(fx) if fx.size else None
        d.n_fixed = int(fx.size)
        d.forces = L.ptr(f)
        d.patch_nodes = L.ptr(patch)
        d.n_patch_nodes = int(sum(lc.Patch.NodeCount for lc in model.Loads))
        d.e0, d.nu, d.emin_ratio, d.penal = model.Material.E0, model.Material.Nu, model.Material.EminRatio, model.Penal
        d.density_box = L.ptr(den)
        d.tol, d.max_iter, d.precision = float(tol), int(max_iter), int(precision)
        if options is not None:
            d.simp = _simp_c(options)
        jid = C.c_int64()
        L.check(L.load().mycel_server_submit(self._h, C.byref(d), C.byref(jid)))
        self._sizes[jid.value] = (model.TotalDofs, model.ActiveElementCount)
        return jid.value

    def submit_fea(self, model: FeModel, tol=1e-8, max_iter=20000, precision=L.PRECISION_FP64, forces=None):
        """StaticAnalysis.Run as a job (Fem/StaticAnalysis.cs:35)."""
        return self._submit(JOB_FEA, model, model.AssembleForces() if forces is None else forces, tol, max_iter,
                            precision)

    def submit_simp(self, model: FeModel, options: OptimizerOptions, forces=None):
        """TopologyOptimizer.Run as a job (Optimization/TopologyOptimizer.cs:16)."""
        return self._submit(JOB_SIMP, model, model.AssembleForces() if forces is None else forces,
                            options=options)

    def poll(self, jid) -> JobInfo:
        i = L.JobInfoC()
        L.check(L.load().mycel_server_poll(self._h, int(jid), C.byref(i)))
        return _info(i)

    def wait(self, jid, timeout_s: float = -1.0) -> JobInfo:
        i = L.JobInfoC()
        L.check(L.load().mycel_server_wait(self._h, int(jid), float(timeout_s), C.byref(i)))
        return _info(i)

    def fetch(self, jid):
        """(u, rho_bar) of a finished job; rho_bar is None for FEA jobs."""
        ndof, na = self._sizes[jid]
        u = np.zeros(ndof)
        rb = np.zeros(max(1, na))
        L.check(L.load().mycel_server_fetch(self._h, int(jid), L.ptr(u), L.ptr(rb)))
        return u, rb[:na]

    def release(self, jid):
        L.check(L.load().mycel_server_release(self._h, int(jid)))
        self._sizes.pop(jid, None)

    def set_packing(self, enable: bool):
        """Pack the resident jobs' multigrid solves into shared launches (default on); before the first submit."""
        L.check(L.load().mycel_server_set_packing(self._h, 1 if enable else 0))

    def pack_stats(self):
        st = L.PackStatsC()
        L.check(L.load().mycel_server_pack_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in L.PackStatsC._fields_}

    def stats(self):
        b, u = C.c_double(), C.c_double()
        q, a = C.c_int32(), C.c_int32()
        L.check(L.load().mycel_server_stats(self._h, C.byref(b), C.byref(u), C.byref(q), C.byref(a)))
        return dict(budget=b.value, in_use=u.value, queued=q.value, active=a.value)
