"""Mycel.Core.Fem on MI355X: FeModel, FemOperator, PcgSolver, StaticAnalysis.

Host-side mirror of the reference's managed API (same class and member names, argument meaning and
exception behaviour) over the C ABI of libmycel_hip.so. FeModel is the input description and its
index bookkeeping (Fem/FeModel.cs); every operator application, solve and reduction runs in the HIP
library on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Callable, Iterable, List, Optional, Sequence

import numpy as np

from . import _lib as L


# --------------------------------------------------------------------------------------------
# Input types (Voxelization/GridSpec.cs, Voxelization/SurfacePatch.cs, Fem/LoadCase.cs, FeMaterial.cs)
# --------------------------------------------------------------------------------------------
class GridSpec:
    """Structured Hex8 grid (Voxelization/GridSpec.cs:11-204). H is float32 as in the reference."""

    def __init__(self, nx: int, ny: int, nz: int, h: float = 1.0, origin=(0.0, 0.0, 0.0)):
        if not (np.float32(h) > 0):
            raise L.ArgumentOutOfRangeError(L.E_OUT_OF_RANGE, "Voxel pitch must be positive.")
        if nx <= 0 or ny <= 0 or nz <= 0:
            raise L.ArgumentOutOfRangeError(L.E_OUT_OF_RANGE, "Grid dimensions must be strictly positive.")
        self.Nx, self.Ny, self.Nz = int(nx), int(ny), int(nz)
        self.H = float(np.float32(h))
        self.Origin = tuple(float(np.float32(o)) for o in origin)

    @property
    def TotalElements(self):
        return self.Nx * self.Ny * self.Nz

    @property
    def TotalNodes(self):
        return (self.Nx + 1) * (self.Ny + 1) * (self.Nz + 1)

    def ElementIndex(self, ix, iy, iz):
        if not (0 <= ix < self.Nx and 0 <= iy < self.Ny and 0 <= iz < self.Nz):
            raise L.ArgumentOutOfRangeError(L.E_OUT_OF_RANGE, "Element coordinates out of bounds.")
        return ix + self.Nx * (iy + self.Ny * iz)

    def NodeIndex(self, i, j, k):
        if not (0 <= i <= self.Nx and 0 <= j <= self.Ny and 0 <= k <= self.Nz):
            raise L.ArgumentOutOfRangeError(L.E_OUT_OF_RANGE, "Node coordinates out of bounds.")
        return i + (self.Nx + 1) * (j + (self.Ny + 1) * k)

    def InverseElementIndex(self, e):
        iz, rem = divmod(e, self.Nx * self.Ny)
        iy, ix = divmod(rem, self.Nx)
        return ix, iy, iz

    def NodeCoords(self, n):
        k, rem = divmod(n, (self.Nx + 1) * (self.Ny + 1))
        j, i = divmod(rem, self.Nx + 1)
        return i, j, k

    def NodePosition(self, i, j, k):
        h = np.float32(self.H)
        o = np.asarray(self.Origin, np.float32)
        return o + np.array([i * h, j * h, k * h], np.float32)

    def ElementCenter(self, ix, iy, iz):
        h = np.float32(self.H)
        o = np.asarray(self.Origin, np.float32)
        return o + np.array([(ix + np.float32(0.5)) * h, (iy + np.float32(0.5)) * h, (iz + np.float32(0.5)) * h],
                            np.float32)

    def _c(self):
        return L.Grid(self.Nx, self.Ny, self.Nz, self.H)


@dataclass
class FeMaterial:
    """Isotropic material (Fem/FeMaterial.cs:11-17)."""

    E0: float
    Nu: float
    EminRatio: float = 1e-6

    @property
    def Emin(self):
        return self.E0 * self.EminRatio


class SurfacePatch:
    """Boundary nodes with tributary areas (Voxelization/SurfacePatch.cs:25-49)."""

    def __init__(self, nodes: Sequence[int], node_area: Sequence[float]):
        self.Nodes = np.asarray(nodes, np.int32)
        self.NodeArea = np.asarray(node_area, np.float64)

    @property
    def NodeCount(self):
        return int(self.Nodes.size)


class LoadCase:
    """Total force spread over a surface patch (Fem/LoadCase.cs:10-20); the force is float32."""

    def __init__(self, patch: SurfacePatch, total_force):
        if patch is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "patch")
        self.Patch = patch
        self.TotalForce = np.asarray(total_force, np.float32).reshape(3)


class FeModel:
    """Structured Hex8 finite element model (Fem/FeModel.cs:11-322)."""

    def __init__(self, grid: GridSpec, domain, material: FeMaterial, fixed_dofs, loads: Optional[Iterable[LoadCase]] = None,
                 element_density=None, penal: float = 3.0):
        if grid is None or domain is None or material is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "grid, domain and material are required")
        self.Grid = grid
        self.Domain = L.u8(domain)
        self.Material = material
        self.FixedDofs = L.i32(fixed_dofs if fixed_dofs is not None else [])
        self.Penal = float(penal)
        if self.Domain.size != self.TotalElements:
            raise L.ArgumentError(L.E_ARGUMENT, f"Domain array length ({self.Domain.size}) does not match Grid "
                                                f"TotalElements ({self.TotalElements}).")
        self.Loads: List[LoadCase] = list(loads) if loads is not None else []
        if element_density is not None:
            ed = L.f64(element_density).copy()
            if ed.size != self.TotalElements:
                raise L.ArgumentError(L.E_ARGUMENT, "ElementDensity length must match TotalElements.")
            self.ElementDensity = ed
        else:
            self.ElementDensity = np.ones(self.TotalElements)
        # Active elements in ascending order and their nodes (FeModel.cs:80-147).
        g = grid
        dom = self.Domain.reshape(g.Nz, g.Ny, g.Nx) != 0
        self.ActiveElements = np.flatnonzero(self.Domain).astype(np.int32)
        an = np.zeros((g.Nz + 1, g.Ny + 1, g.Nx + 1), bool)
        for oz in (0, 1):
            for oy in (0, 1):
                for ox in (0, 1):
                    an[oz:oz + g.Nz, oy:oy + g.Ny, ox:ox + g.Nx] |= dom
        self.ActiveNodes = an.reshape(-1)
        cons = np.repeat(~self.ActiveNodes, 3)
        if self.FixedDofs.size:
            bad = (self.FixedDofs < 0) | (self.FixedDofs >= self.TotalDofs)
            if bad.any():
                dof = int(self.FixedDofs[np.argmax(bad)])
                raise L.ArgumentOutOfRangeError(
                    L.E_OUT_OF_RANGE, f"Fixed DOF {dof} is out of bounds [0, {self.TotalDofs - 1}].")
            cons[self.FixedDofs] = True
            uniq = np.unique(self.FixedDofs)
            self.ActiveFixedDofCount = int(self.ActiveNodes[uniq // 3].sum())
        else:
            self.ActiveFixedDofCount = 0
        self.ConstrainedDofs = cons

    @property
    def TotalElements(self):
        return self.Grid.TotalElements

    @property
    def TotalNodes(self):
        return self.Grid.TotalNodes

    @property
    def TotalDofs(self):
        return 3 * self.TotalNodes

    @property
    def ActiveElementCount(self):
        return int(self.ActiveElements.size)

    @property
    def ActiveNodeCount(self):
        return int(self.ActiveNodes.sum())

    @staticmethod
    def FixDofs(nodes, fixX=True, fixY=True, fixZ=True):
        """FeModel.FixDofs (Fem/FeModel.cs:189-200)."""
        out = []
        for n in nodes:
            if fixX:
                out.append(3 * n)
            if fixY:
                out.append(3 * n + 1)
            if fixZ:
                out.append(3 * n + 2)
        return np.asarray(out, np.int32)

    def AssembleForces(self):
        """FeModel.AssembleForces (Fem/FeModel.cs:263-321)."""
        lib = L.load()
        offs = [0]
        nodes, areas, forces = [], [], []
        for lc in self.Loads:
            nodes.append(lc.Patch.Nodes)
            areas.append(lc.Patch.NodeArea)
            offs.append(offs[-1] + lc.Patch.NodeCount)
            forces.append(lc.TotalForce)
        f = np.zeros(self.TotalDofs)
        o = L.i32(offs)
        nd = L.i32(np.concatenate(nodes)) if nodes else np.zeros(1, np.int32)
        ar = L.f64(np.concatenate(areas)) if areas else np.zeros(1)
        fv = np.ascontiguousarray(np.concatenate(forces), np.float32) if forces else np.zeros(3, np.float32)
        L.check(lib.mycel_assemble_forces(self.TotalNodes, len(self.Loads), L.ptr(o), L.ptr(nd), L.ptr(ar),
                                          L.ptr(fv), L.ptr(f)))
        return f


# --------------------------------------------------------------------------------------------
# FemOperator (Fem/FemOperator.cs)
# --------------------------------------------------------------------------------------------
class FemOperator:
    """Matrix-free K(rho) and A = PKP + (I - P) on the GPU (Fem/FemOperator.cs:14-440)."""

    def __init__(self, model: FeModel):
        if model is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "model")
        self.Model = model
        lib = L.load()
        g = model.Grid._c()
        h = C.c_void_p()
        L.check(lib.mycel_op_create(C.byref(g), L.ptr(model.Domain), L.ptr(model.FixedDofs) if model.FixedDofs.size else None,
                                    int(model.FixedDofs.size), float(model.Material.E0), float(model.Material.Nu),
                                    float(model.Material.EminRatio), float(model.Penal), L.ptr(model.ElementDensity),
                                    C.byref(h)))
        self._h = h
        self._n = model.TotalDofs

    def close(self):
        if getattr(self, "_h", None):
            L.load().mycel_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def Ke(self):
        ke = np.zeros(576)
        L.check(L.load().mycel_op_get_ke(self._h, L.ptr(ke)))
        return ke

    @property
    def DiagKe(self):
        return self.Ke.reshape(24, 24).diagonal().copy()

    @property
    def DiagA(self):
        d = np.zeros(self._n)
        L.check(L.load().mycel_op_get_diag_a(self._h, L.ptr(d)))
        return d

    def info(self):
        i = L.OpInfo()
        L.check(L.load().mycel_op_info_get(self._h, C.byref(i)))
        return i

    def UpdateDensities(self, densities, penal):
        """FemOperator.UpdateDensities (Fem/FemOperator.cs:114-138); syncs Model.ElementDensity/Penal."""
        d = L.f64(densities)
        if d.size != self.Model.TotalElements:
            raise L.ArgumentError(L.E_ARGUMENT, "densities must be box-sized")
        if d is not self.Model.ElementDensity:
            self.Model.ElementDensity[:] = d
        self.Model.Penal = float(penal)
        L.check(L.load().mycel_op_update_densities(self._h, L.ptr(d), float(penal)))

    def _vec(self, x):
        x = L.f64(x)
        if x.size != self._n:
            raise L.ArgumentError(L.E_ARGUMENT, f"vector length {x.size} != TotalDofs {self._n}")
        return x

    def ApplyK(self, x, y=None):
        """y = K(rho) x (Fem/FemOperator.cs:196-243). Returns y."""
        x = self._vec(x)
        out = np.empty(self._n) if y is None else y
        L.check(L.load().mycel_op_apply_k(self._h, L.ptr(x), L.ptr(out)))
        return out

    def SetMatvecForm(self, form: int, columns: int = 0) -> None:
        """K*u kernel form of ApplyK / ApplyA and the FP32 operator (no counterpart in the reference; the
        solvers' fused-dot launches always use the automatic form). MATVEC_NODE_GATHER: columns 0 = one node
        per thread, 2/4/8 = each thread marches that many nodes along z. MATVEC_ELEMENT: each element's product
        once in the monomial basis, columns 2/3/4/8 node layers per tile. MATVEC_AUTO: by grid size. Forms of one
        family give bit-identical results; the two families agree to rounding."""
        L.check(L.load().mycel_op_set_matvec_form(self._h, int(form), int(columns)))

    def SetIslandDeflation(self, solid_ratio: float) -> None:
        """Multigrid island deflation (mycel_op_set_island_deflation): 0 turns it off."""
        L.check(L.load().mycel_op_set_island_deflation(self._h, float(solid_ratio)))

    def SetMatvecColumns(self, columns: int) -> None:
        """The gather form with columns 0/2/4/8, or -1 for the automatic form."""
        L.check(L.load().mycel_op_set_matvec_columns(self._h, int(columns)))

    def ApplyA(self, x, y=None, xIsAlreadyProjected=False):
        """y = [PKP + (I-P)] x (Fem/FemOperator.cs:250-294). Returns y."""
        x = self._vec(x)
        out = np.empty(self._n) if y is None else y
        L.check(L.load().mycel_op_apply_a(self._h, L.ptr(x), L.ptr(out), int(bool(xIsAlreadyProjected))))
        return out

    # Device-resident forms (torch tensors on the GPU; used by the benchmark and serving paths).
    def update_densities_device(self, d_t, penal, stream=None):
        """UpdateDensities from a box-sized float64 device tensor (enqueued on the operator's stream)."""
        L.check(L.load().mycel_op_update_densities_d(self._h, L.device_ptr(d_t), float(penal), stream))

    def apply_k_device(self, x_t, y_t, stream=None):
        L.check(L.load().mycel_op_apply_k_d(self._h, L.device_ptr(x_t), L.device_ptr(y_t), stream))

    def apply_k_device_f32(self, x_t, y_t, stream=None):
        """The FP32 operator of the mixed-precision solver on float32 device tensors."""
        L.check(L.load().mycel_op_apply_k_f32_d(self._h, L.device_ptr(x_t), L.device_ptr(y_t), stream))

    def stream(self):
        return L.load().mycel_op_stream(self._h)


class JacobiPreconditioner:
    """Point Jacobi z = r / diag(A) (Fem/JacobiPreconditioner.cs:9-36). The GPU solver applies it fused
    into its vector kernels, as the reference special-cases Jacobi (Fem/PcgSolver.cs:213-226)."""


class MultigridPreconditioner:
    """A second IPreconditioner (Fem/IPreconditioner.cs:9-20): one FP32 geometric-multigrid V-cycle per
    iteration (2x coarsening, Chebyshev-Jacobi smoothing, dense coarsest solve) inside an FP64 flexible PCG.
    Passing it to PcgSolver.Solve selects PRECISION_MIXED_GMG."""


def precond_apply_device(op: "FemOperator", r_t, z_t, precision=L.PRECISION_MIXED_GMG, stream=None):
    """IPreconditioner.Apply on device tensors (mycel_precond_apply_d): z = M^-1 r, Jacobi or one V-cycle."""
    L.check(L.load().mycel_precond_apply_d(op._h, int(precision), L.device_ptr(r_t), L.device_ptr(z_t), stream))


@dataclass
class PcgResult:
    """Fem/PcgResult.cs:7-14."""

    Iterations: int = 0
    FinalResidual: float = 0.0
    Converged: bool = False
    ElapsedMs: float = 0.0
    Message: Optional[str] = None
    MessageCode: int = 0
    OuterCorrections: int = 0
    FallbackIterations: int = 0


def _pcg_result(r: L.PcgResultC) -> PcgResult:
    buf = C.create_string_buffer(512)
    L.check(L.load().mycel_pcg_message(C.byref(r), buf, len(buf)))
    text = buf.value.decode("utf-8") or None  # PcgResult.Message is null for a converged solve
    return PcgResult(r.iterations, r.final_residual, bool(r.converged), r.elapsed_ms, text,
                     r.message, r.outer_corrections, r.fallback_iterations)


def _precision_for(preconditioner, precision):
    """The solver mode a preconditioner object selects (None and Jacobi keep `precision`)."""
    if preconditioner is None:
        return precision
    if isinstance(preconditioner, JacobiPreconditioner):
        if precision == L.PRECISION_MIXED_GMG:
            raise L.ArgumentError(L.E_ARGUMENT, "PRECISION_MIXED_GMG is the multigrid preconditioner, not Jacobi")
        return precision
    if isinstance(preconditioner, MultigridPreconditioner):
        return L.PRECISION_MIXED_GMG
    raise L.ArgumentError(L.E_ARGUMENT, "preconditioner must be JacobiPreconditioner or MultigridPreconditioner")


class PcgSolver:
    """Preconditioned CG on A (Fem/PcgSolver.cs:16-554), device-resident."""

    def Solve(self, op: FemOperator, b, x, preconditioner=None, tol=1e-6, maxIterations=5000, progress=None,
              cancel=None, precision=L.PRECISION_FP64):
        """Solves A x = b in place in x (a float64 numpy array). progress(iteration, relres) is called
        between device batches; cancel is an optional ctypes c_int32 flag."""
        if op is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "op")
        precision = _precision_for(preconditioner, precision)
        b = op._vec(b)
        if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags["C_CONTIGUOUS"] and x.size == op._n):
            raise L.ArgumentError(L.E_ARGUMENT, "x must be a contiguous float64 array of length TotalDofs")
        res = L.PcgResultC()
        cb = L.PROGRESS_CB(lambda it, rel, _u: progress(it, rel)) if progress else L.PROGRESS_CB()
        cptr = C.cast(C.pointer(cancel), C.c_void_p) if cancel is not None else None
        L.check(L.load().mycel_pcg_solve(op._h, L.ptr(b), L.ptr(x), float(tol), int(maxIterations), int(precision),
                                         cptr, cb, None, C.byref(res)))
        return _pcg_result(res)

    def SolveBatch(self, op: FemOperator, B, X, tol=1e-6, maxIterations=5000,
                   precision=L.PRECISION_MIXED_GMG) -> List["PcgResult"]:
        """R systems in lockstep (mycel_pcg_solve_batch): B, X are C-contiguous (R, TotalDofs) float64 arrays;
        X holds the warm starts and receives the solutions."""
        if not (isinstance(X, np.ndarray) and X.dtype == np.float64 and X.flags["C_CONTIGUOUS"] and X.ndim == 2
                and X.shape[1] == op._n):
            raise L.ArgumentError(L.E_ARGUMENT, "X must be a contiguous float64 array of shape (R, TotalDofs)")
        B = np.ascontiguousarray(B, np.float64)
        if B.shape != X.shape:
            raise L.ArgumentError(L.E_ARGUMENT, "B and X must have the same shape")
        R = int(X.shape[0])
        res = (L.PcgResultC * R)()
        L.check(L.load().mycel_pcg_solve_batch(op._h, L.ptr(B), L.ptr(X), R, float(tol), int(maxIterations),
                                               int(precision), res))
        return [_pcg_result(r) for r in res]

    def solve_batch_device(self, op: FemOperator, B_t, X_t, tol=1e-6, maxIterations=5000,
                           precision=L.PRECISION_MIXED_GMG) -> List["PcgResult"]:
        """R systems in lockstep (mycel_pcg_solve_batch_d): B_t, X_t are (R, TotalDofs) float64 device tensors."""
        R = int(B_t.shape[0])
        res = (L.PcgResultC * R)()
        L.check(L.load().mycel_pcg_solve_batch_d(op._h, L.device_ptr(B_t), L.device_ptr(X_t), R, float(tol),
                                                 int(maxIterations), int(precision), res))
        return [_pcg_result(r) for r in res]

    def solve_packed_device(self, ops, bs, xs, tol=1e-6, maxIterations=5000) -> List["PcgResult"]:
        """Independent multigrid systems on distinct operators (different grids) solved together in shared
        launches (mycel_pcg_solve_packed_d): bs[i], xs[i] are float64 device tensors of ops[i]'s box DOFs (xs:
        warm starts in, solutions out). Each result is bit-identical to solve_device of that system alone."""
        n = len(ops)
        if not (len(bs) == n and len(xs) == n):
            raise L.ArgumentError(L.E_ARGUMENT, "ops, bs and xs must have the same length")
        res = (L.PcgResultC * n)()
        o = (C.c_void_p * n)(*[op._h.value if isinstance(op._h, C.c_void_p) else op._h for op in ops])
        b = (C.c_void_p * n)(*[t.data_ptr() for t in bs])
        x = (C.c_void_p * n)(*[t.data_ptr() for t in xs])
        L.check(L.load().mycel_pcg_solve_packed_d(o, b, x, n, float(tol), int(maxIterations), res))
        return [_pcg_result(r) for r in res]

    def solve_device(self, op: FemOperator, b_t, x_t, tol=1e-6, maxIterations=5000, precision=L.PRECISION_FP64):
        res = L.PcgResultC()
        L.check(L.load().mycel_pcg_solve_d(op._h, L.device_ptr(b_t), L.device_ptr(x_t), float(tol), int(maxIterations),
                                           int(precision), C.byref(res)))
        return _pcg_result(res)


# --------------------------------------------------------------------------------------------
# StaticAnalysis (Fem/StaticAnalysis.cs)
# --------------------------------------------------------------------------------------------
@dataclass
class StaticAnalysisOptions:
    """Fem/StaticAnalysis.cs:8-17."""

    Tolerance: float = 1e-6
    MaxIterations: int = 5000
    WarmStartDisplacements: Optional[np.ndarray] = None
    Preconditioner: Optional[object] = None  # JacobiPreconditioner or MultigridPreconditioner
    ThrowOnNonConvergence: bool = True
    Precision: int = L.PRECISION_FP64


@dataclass
class StaticAnalysisResult:
    """Fem/StaticAnalysis.cs:19-27."""

    U: np.ndarray = field(default_factory=lambda: np.zeros(0))
    Compliance: float = 0.0
    Reactions: np.ndarray = field(default_factory=lambda: np.zeros(0))
    VonMisesPerElement: np.ndarray = field(default_factory=lambda: np.zeros(0))
    MaxDisplacement: float = 0.0
    SolverInfo: PcgResult = field(default_factory=PcgResult)


class StaticAnalysis:
    @staticmethod
    def Run(model: FeModel, options: Optional[StaticAnalysisOptions] = None, op: Optional[FemOperator] = None):
        """StaticAnalysis.Run (Fem/StaticAnalysis.cs:35-142)."""
        if model is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "model")
        options = options or StaticAnalysisOptions()
        if model.ActiveFixedDofCount < 6:
            raise L.InvalidOperationError(
                L.E_INVALID_OPERATION,
                f"Model has fewer than 6 unique fixed DOFs on active nodes (found {model.ActiveFixedDofCount}). "
                "At least 6 fixed DOFs are required to prevent rigid body modes in 3D.")
        f = model.AssembleForces()
        n = model.TotalDofs
        u = np.zeros(n)
        if options.WarmStartDisplacements is not None:
            w = np.asarray(options.WarmStartDisplacements, np.float64)
            if w.size != n:
                raise L.ArgumentError(L.E_ARGUMENT, f"WarmStartDisplacements length ({w.size}) must match "
                                                    f"TotalDofs ({n}).")
            u[:] = w
        own = op is None
        op = op or FemOperator(model)
        try:
            comp, mx = C.c_double(), C.c_double()
            react = np.zeros(n)
            vm = np.zeros(model.TotalElements)
            res = L.PcgResultC()
            L.check(L.load().mycel_static_analysis(op._h, L.ptr(f), L.ptr(u), float(options.Tolerance),
                                                   int(options.MaxIterations), int(bool(options.ThrowOnNonConvergence)),
                                                   int(_precision_for(options.Preconditioner, options.Precision)), C.byref(comp), C.byref(mx), L.ptr(react),
                                                   L.ptr(vm), C.byref(res)))
        finally:
            if own:
                op.close()
        return StaticAnalysisResult(u, comp.value, react, vm, mx.value, _pcg_result(res))

    @staticmethod
    def AverageElementToNode(op: FemOperator, element_scalar):
        """StaticAnalysis.AverageElementToNode (Fem/StaticAnalysis.cs:222-272)."""
        e = L.f64(element_scalar)
        if e.size != op.Model.TotalElements:
            raise L.ArgumentError(L.E_ARGUMENT, "elementScalar length must equal TotalElements.")
        out = np.zeros(op.Model.TotalNodes)
        L.check(L.load().mycel_average_element_to_node(op._h, L.ptr(e), L.ptr(out)))
        return out

    @staticmethod
    def SampleNodalScalars(op: FemOperator, nodal_scalar, points):
        """StaticAnalysis.SampleNodalScalar (Fem/StaticAnalysis.cs:278-366) at many points: points is (n, 3), taken
        as float32 (System.Numerics.Vector3); returns n FP64 values."""
        nd = L.f64(nodal_scalar)
        if nd.size != op.Model.TotalNodes:
            raise L.ArgumentError(L.E_ARGUMENT, "nodalScalar length must equal TotalNodes.")
        pts = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 3))
        org = np.ascontiguousarray(np.asarray(op.Model.Grid.Origin, np.float32))
        out = np.zeros(pts.shape[0])
        L.check(L.load().mycel_sample_nodal_scalar(op._h, L.ptr(nd), int(pts.shape[0]), L.ptr(pts), L.ptr(org),
                                                   L.ptr(out)))
        return out

    @staticmethod
    def SampleNodalScalar(op: FemOperator, nodal_scalar, point):
        """StaticAnalysis.SampleNodalScalar at one point (x, y, z)."""
        return float(StaticAnalysis.SampleNodalScalars(op, nodal_scalar, [point])[0])

    @staticmethod
    def ComputeElementVonMises(op: FemOperator, u):
        """StaticAnalysis.ComputeElementVonMises (Fem/StaticAnalysis.cs:144-216)."""
        u = op._vec(u)
        out = np.zeros(op.Model.TotalElements)
        L.check(L.load().mycel_von_mises(op._h, L.ptr(u), L.ptr(out)))
        return out


class Hex8Element:
    """Fem/Hex8Element.cs:48-252 (computed by the library)."""

    @staticmethod
    def CreateKe(h, nu):
        ke = np.zeros(576)
        L.check(L.load().mycel_hex8_ke(float(h), float(nu), L.ptr(ke)))
        return ke

    @staticmethod
    def CreateBc(h):
        bc = np.zeros(144)
        L.check(L.load().mycel_hex8_bc(float(h), L.ptr(bc)))
        return bc
