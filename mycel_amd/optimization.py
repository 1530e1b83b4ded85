"""Mycel.Core.Optimization on MI355X: DensityFilter, SensitivityAnalysis, OcOptimizer, TopologyOptimizer.

Host-side mirror of the reference's managed API over the C ABI of libmycel_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import _lib as L
from .fem import FemOperator, FeModel, GridSpec


class DensityFilter:
    """Linear cone filter W, W^T and S (Optimization/DensityFilter.cs:15-304), evaluated on the GPU."""

    def __init__(self, grid: GridSpec, domain, activeElements, filterRadiusInElements: float):
        if grid is None or domain is None or activeElements is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "grid, domain and activeElements are required")
        self.Grid = grid
        self.FilterRadiusInElements = float(filterRadiusInElements)
        dom = L.u8(domain)
        act = L.i32(activeElements)
        self.ActiveCount = int(act.size)
        h = C.c_void_p()
        g = grid._c()
        L.check(L.load().mycel_filter_create(C.byref(g), L.ptr(dom), L.ptr(act) if act.size else L.ptr(np.zeros(1, np.int32)),
                                             self.ActiveCount, self.FilterRadiusInElements, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            L.load().mycel_filter_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def S(self):
        s = np.zeros(max(1, self.ActiveCount))
        L.check(L.load().mycel_filter_get_s(self._h, L.ptr(s)))
        return s[: self.ActiveCount]

    @property
    def NonZeros(self):
        n = C.c_int64()
        L.check(L.load().mycel_filter_nnz(self._h, C.byref(n)))
        return n.value

    def _io(self, x, out):
        x = L.f64(x)
        if x.size < self.ActiveCount:
            raise L.ArgumentError(L.E_ARGUMENT, "Buffer lengths must be at least ActiveCount.")
        if out is None:
            out = np.zeros(self.ActiveCount)
        elif out.size < self.ActiveCount:
            raise L.ArgumentError(L.E_ARGUMENT, "Buffer lengths must be at least ActiveCount.")
        return x, out

    def ApplyFilter(self, rhoIn, rhoFilteredOut=None):
        """rho_tilde = W rho (DensityFilter.cs:205-251)."""
        x, out = self._io(rhoIn, rhoFilteredOut)
        if self.ActiveCount:
            L.check(L.load().mycel_filter_apply(self._h, L.ptr(x), L.ptr(out)))
        return out

    def ApplySensitivityFilter(self, dcIn, dcFilteredOut=None):
        """dc_out = W^T dc (DensityFilter.cs:257-303)."""
        x, out = self._io(dcIn, dcFilteredOut)
        if self.ActiveCount:
            L.check(L.load().mycel_filter_apply_adjoint(self._h, L.ptr(x), L.ptr(out)))
        return out


class SensitivityAnalysis:
    @staticmethod
    def ComputeComplianceSensitivities(model: FeModel, op: FemOperator, u, rhoPhysical, dcDrhoPhysical=None, penal=3.0):
        """dc/drho_bar_e = -p rho_bar^(p-1) (E0 - Emin) u_e^T Ke u_e (SensitivityAnalysis.cs:13-146)."""
        if model is None or op is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "model/op")
        na = model.ActiveElementCount
        rho = L.f64(rhoPhysical)
        if rho.size < na or (dcDrhoPhysical is not None and dcDrhoPhysical.size < na):
            raise L.ArgumentError(L.E_ARGUMENT, "Density spans must have length >= ActiveElementCount.")
        out = np.zeros(na) if dcDrhoPhysical is None else dcDrhoPhysical
        u = op._vec(u)
        L.check(L.load().mycel_sensitivities(op._h, L.ptr(u), L.ptr(rho), L.ptr(out), float(penal)))
        return out


class OcOptimizer:
    @staticmethod
    def Update(rho, dcDrho, s, isPassive=None, rhoNew=None, targetVolumeFraction=0.4, moveLimit=0.2, rhoMin=1e-3,
               lambdaGuess=1.0, maxBisectionSteps=60, volumeTolerance=1e-8):
        """OC update with log-bisection on lambda, on the GPU (OcOptimizer.cs:12-132). Returns
        (lambda, rhoNew)."""
        rho = L.f64(rho)
        n = rho.size
        dc, sv = L.f64(dcDrho), L.f64(s)
        if dc.size < n or sv.size < n or (rhoNew is not None and rhoNew.size < n):
            raise L.ArgumentError(L.E_ARGUMENT, "All spans must have length >= rho.Length.")
        out = np.zeros(n) if rhoNew is None else rhoNew
        pas = None
        if isPassive is not None and len(isPassive) > 0:
            p = np.zeros(n, np.uint8)
            m = min(n, len(isPassive))
            p[:m] = np.asarray(isPassive[:m], bool)
            pas = p
        lam = C.c_double()
        L.check(L.load().mycel_oc_update(L.ptr(rho), L.ptr(dc), L.ptr(sv), L.ptr(pas), L.ptr(out), n,
                                         float(targetVolumeFraction), float(moveLimit), float(rhoMin),
                                         float(lambdaGuess), int(maxBisectionSteps), float(volumeTolerance),
                                         C.byref(lam)))
        return lam.value, out


@dataclass
class OptimizerOptions:
    """Optimization/OptimizerOptions.cs:5-26 (Heaviside and RngSeed are not on the GPU path)."""

    VolumeFraction: float = 0.4
    FilterRadiusMm: float = 2.0
    Penal: float = 3.0
    MoveLimit: float = 0.2
    MaxIterations: int = 200
    SolverToleranceInitial: float = 1e-4
    SolverToleranceFinal: float = 1e-6
    SolverMaxIterations: int = 5000
    EminRatio: float = 1e-6  # not read by the SIMP loop, as in the reference (Emin comes from FeMaterial)
    Continuation: bool = False
    EnableHeaviside: bool = False
    RhoMin: float = 1e-3
    PassiveMask: Optional[np.ndarray] = None
    ProtectBoundaryHinges: bool = True
    InitialDensity: Optional[np.ndarray] = None
    StartIteration: int = 1
    WarmStartDisplacements: Optional[np.ndarray] = None
    EnforceMinimumFilterRadius: bool = True
    Precision: int = L.PRECISION_FP64
    InitialLambda: float = 1.0  # the OC multiplier's first guess (TopologyOptimizer.cs:154); a resume passes the saved one
    InitialDesign64: Optional[np.ndarray] = None  # FP64 initial design (box- or active-sized): an exact resume
    # The loop state of an exact resume (no reference counterpart): the previous iteration's max change (None:
    # restart the tolerance schedule as the reference's StartIteration does), the Heaviside beta in effect and
    # the last <= 6 compliances, oldest first.
    ResumeMaxChange: Optional[float] = None
    ResumeBeta: float = 1.0
    ResumeComplianceHistory: Optional[Sequence[float]] = None
    IslandDeflation: float = 0.0  # multigrid island deflation's solid ratio for this run (0: the operator's setting)


@dataclass
class IterationSnapshot:
    """Optimization/IterationSnapshot.cs:12-31 plus lambda; PhysicalDensity (rho_bar) and DesignDensity (rho,
    after this iteration's OC update) are float32 box-sized, as TopologyOptimizer.cs:340-347 builds them."""

    Iteration: int = 0
    Compliance: float = 0.0
    VolumeFraction: float = 0.0
    MaxChange: float = 0.0
    GreyLevel: float = 0.0
    SolverIterations: int = 0
    SolverResidual: float = 0.0
    MaxDisplacement: float = 0.0
    Lambda: float = 0.0
    SolveMs: float = 0.0
    SensitivityMs: float = 0.0
    FilterMs: float = 0.0
    UpdateMs: float = 0.0
    TotalMs: float = 0.0
    PhysicalDensity: Optional[np.ndarray] = None
    DesignDensity: Optional[np.ndarray] = None
    BetaNext: float = 1.0  # the Heaviside beta of the next iteration (resume state)


@dataclass
class TopologyOptimizationResult:
    """Optimization/TopologyOptimizationResult.cs:6-17 plus the FP64 design arrays."""

    Converged: bool = False
    TotalIterations: int = 0
    Snapshots: List[IterationSnapshot] = field(default_factory=list)
    FinalCompliance: float = 0.0
    FinalVolumeFraction: float = 0.0
    FinalPhysicalDensity: Optional[np.ndarray] = None
    FinalDisplacements: Optional[np.ndarray] = None
    FinalDesign: Optional[np.ndarray] = None
    FinalPhysical64: Optional[np.ndarray] = None
    Canceled: bool = False


def _snap(s: L.SnapshotC) -> IterationSnapshot:
    return IterationSnapshot(s.iteration, s.compliance, s.volume_fraction, s.max_change, s.grey_level,
                             s.solver_iterations, s.solver_residual, s.max_displacement, s.lambda_, s.solve_ms,
                             s.sensitivity_ms, s.filter_ms, s.update_ms, s.total_ms, BetaNext=s.beta_next)


def simp_options_c(o: "OptimizerOptions") -> L.SimpOptionsC:
    """The C options struct of an OptimizerOptions (initial_design is set by the caller that owns the array)."""
    co = L.SimpOptionsC(o.VolumeFraction, o.FilterRadiusMm, o.Penal, o.MoveLimit, o.MaxIterations,
                        o.SolverToleranceInitial, o.SolverToleranceFinal, o.SolverMaxIterations, o.RhoMin,
                        int(o.ProtectBoundaryHinges), o.StartIteration, int(o.EnforceMinimumFilterRadius),
                        int(o.Continuation), int(o.Precision), int(o.EnableHeaviside), float(o.InitialLambda))
    co.island_deflation = float(o.IslandDeflation)
    if o.ResumeMaxChange is not None:
        hist = list(o.ResumeComplianceHistory or [])[-6:]
        co.resume_state = 1
        co.resume_max_change = float(o.ResumeMaxChange)
        co.resume_beta = float(o.ResumeBeta)
        co.resume_history_len = len(hist)
        for k, c in enumerate(hist):
            co.resume_history[k] = float(c)
    return co


class TopologyOptimizer:
    @staticmethod
    def Run(model: FeModel, options: Optional[OptimizerOptions] = None, progress: Optional[Callable] = None,
            cancel: Optional[Callable[[], bool]] = None, op: Optional[FemOperator] = None):
        """SIMP with filter, sensitivities and OC (TopologyOptimizer.cs:16-407), all on the GPU.
        progress(snapshot) receives each IterationSnapshot (with float32 box-sized PhysicalDensity);
        cancel() returning True stops after the current iteration."""
        if model is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "model")
        return TopologyOptimizer._run(model, model.AssembleForces(), 1, options, progress, cancel, op)

    @staticmethod
    def RunCases(model: FeModel, forces, options: Optional[OptimizerOptions] = None,
                 progress: Optional[Callable] = None, cancel: Optional[Callable[[], bool]] = None,
                 op: Optional[FemOperator] = None):
        """SIMP on the summed compliance of independent load cases (mycel_simp_run_cases): forces is an
        (n_cases, TotalDofs) array of assembled RHS vectors; FinalDisplacements comes back with the same
        shape. The reference instead sums load cases into one RHS (Fem/FeModel.cs:267-318)."""
        if model is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "model")
        f = np.ascontiguousarray(forces, np.float64)
        if f.ndim != 2 or f.shape[1] != model.TotalDofs or f.shape[0] < 1:
            raise L.ArgumentError(L.E_ARGUMENT, "forces must be (n_cases, TotalDofs)")
        return TopologyOptimizer._run(model, f, f.shape[0], options, progress, cancel, op)

    @staticmethod
    def _run(model, f, n_cases, options, progress, cancel, op):
        o = options or OptimizerOptions()
        co = simp_options_c(o)
        design64 = None
        if o.InitialDesign64 is not None:
            design64 = np.ascontiguousarray(o.InitialDesign64, np.float64).ravel()
            co.initial_design = design64.ctypes.data
            co.initial_design_len = int(design64.size)
        patch = np.concatenate([lc.Patch.Nodes for lc in model.Loads]).astype(np.int32) if model.Loads else np.zeros(0, np.int32)
        pm = L.u8(o.PassiveMask) if o.PassiveMask is not None else None
        idn = np.ascontiguousarray(o.InitialDensity, np.float32) if o.InitialDensity is not None else None
        wu = None
        if o.WarmStartDisplacements is not None and np.size(o.WarmStartDisplacements) == model.TotalDofs * n_cases:
            wu = L.f64(np.ravel(o.WarmStartDisplacements))
        n_it = max(1, o.MaxIterations - o.StartIteration + 1)
        snaps = (L.SnapshotC * n_it)()
        res = L.SimpResultC()
        na = model.ActiveElementCount
        rho, rb = np.zeros(max(1, na)), np.zeros(max(1, na))
        u = np.zeros((n_cases, model.TotalDofs)) if f.ndim == 2 else np.zeros(model.TotalDofs)
        collected: List[IterationSnapshot] = []
        act = model.ActiveElements

        def _cb(sp, rbp, rp, _u):
            s = _snap(sp.contents)
            box = np.zeros(model.TotalElements, np.float32)
            box[act] = np.ctypeslib.as_array(rbp, shape=(na,)).astype(np.float32)
            s.PhysicalDensity = box
            des = np.zeros(model.TotalElements, np.float32)
            des[act] = np.ctypeslib.as_array(rp, shape=(na,)).astype(np.float32)
            s.DesignDensity = des
            collected.append(s)
            if progress is not None:
                progress(s)
            return 1 if (cancel is not None and cancel()) else 0

        cb = L.SNAPSHOT_CB(_cb)
        op_own = op is None
        op = op or FemOperator(model)
        try:
            L.check(L.load().mycel_simp_run_cases(op._h, L.ptr(f), int(n_cases), C.byref(co), L.ptr(pm),
                                                  L.ptr(patch) if patch.size else None, int(patch.size), L.ptr(idn),
                                                  0 if idn is None else int(idn.size), L.ptr(wu), cb, None,
                                                  C.byref(res), snaps, n_it, L.ptr(rho), L.ptr(rb), L.ptr(u)))
        finally:
            if op_own:
                op.close()
        final_box = np.zeros(model.TotalElements, np.float32)
        final_box[act] = rb[:na].astype(np.float32)
        model.ElementDensity[act] = rb[:na]
        return TopologyOptimizationResult(bool(res.converged), res.total_iterations, collected, res.final_compliance,
                                          res.final_volume_fraction, final_box, u, rho[:na], rb[:na],
                                          bool(res.canceled))
