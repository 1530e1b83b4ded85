"""FP64 CPU oracle for the Hex8 elasticity path -- TEST INFRASTRUCTURE ONLY.

This package wraps ``oracle/liboracle.so`` (a plain-C restatement of Mycel.Core's FemOperator,
PcgSolver, StaticAnalysis, DensityFilter, SensitivityAnalysis, OcOptimizer and TopologyOptimizer;
see ``mycel_oracle.c`` for the reference file:line of each function).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it,
and only as the checker or the timed CPU baseline. The product package ``mycel_amd`` never imports
it and has no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

PCG_MESSAGES = {
    0: "Converged",
    1: "RHS is zero. Solution is zero.",
    2: "Initial guess already converged.",
    3: "PCG curvature breakdown",
    4: "PCG step breakdown",
    5: "PCG stagnated",
    6: "PCG search direction breakdown",
    7: "PCG reached maximum iterations without converging",
}


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (incremental: make rebuilds it when a source or
    header is newer, so a stale library is never loaded)."""
    subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class PcgResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("final_residual", C.c_double), ("converged", C.c_int),
                ("message", C.c_int), ("elapsed_ms", C.c_double)]


class SimpOptions(C.Structure):
    _fields_ = [("volume_fraction", C.c_double), ("filter_radius_mm", C.c_double), ("penal", C.c_double),
                ("move_limit", C.c_double), ("max_iterations", C.c_int), ("solver_tol_initial", C.c_double),
                ("solver_tol_final", C.c_double), ("solver_max_iterations", C.c_int), ("rho_min", C.c_double),
                ("protect_boundary_hinges", C.c_int), ("start_iteration", C.c_int),
                ("enforce_min_filter_radius", C.c_int), ("continuation", C.c_int), ("enable_heaviside", C.c_int)]


class Snapshot(C.Structure):
    _fields_ = [("iteration", C.c_int), ("compliance", C.c_double), ("volume_fraction", C.c_double),
                ("max_change", C.c_double), ("grey_level", C.c_double), ("solver_iterations", C.c_int),
                ("solver_residual", C.c_double), ("max_displacement", C.c_double), ("lambda_", C.c_double),
                ("solver_tolerance", C.c_double)]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def _declare(L):
    L.orc_pow_penal_n.argtypes = [_dp, C.c_int, C.c_double, _dp]
    L.orc_pow_penal_n.restype = None
    L.orc_hex8_ke.argtypes = [C.c_double, C.c_double, _dp]
    L.orc_hex8_bc.argtypes = [C.c_double, _dp]
    L.orc_constitutive.argtypes = [C.c_double, C.c_double, _dp]
    L.orc_model_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, _up, _vp, C.c_int, _vp, C.POINTER(C.c_void_p)]
    L.orc_model_destroy.argtypes = [_vp]
    L.orc_model_active_count.argtypes = [_vp]
    L.orc_model_active_fixed_count.argtypes = [_vp]
    L.orc_model_active_elements.argtypes = [_vp, _ip]
    L.orc_model_constrained.argtypes = [_vp, _up]
    L.orc_model_active_nodes.argtypes = [_vp, _up]
    L.orc_model_diag_a.argtypes = [_vp, _dp]
    L.orc_model_ke.argtypes = [_vp]
    L.orc_model_ke.restype = C.POINTER(C.c_double)
    L.orc_assemble_forces.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _fp, _dp]
    L.orc_update_densities.argtypes = [_vp, _dp, C.c_double]
    L.orc_apply_k.argtypes = [_vp, _dp, _dp, C.c_int]
    L.orc_apply_a.argtypes = [_vp, _dp, _dp, C.c_int, C.c_int]
    L.orc_apply_k_scalar.argtypes = [_vp, _dp, _dp]
    L.orc_pcg_solve.argtypes = [_vp, _dp, _dp, C.c_double, C.c_int, C.c_int, C.POINTER(PcgResult)]
    L.orc_static_analysis.argtypes = [_vp, _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.c_double), C.POINTER(C.c_double), _vp, _vp,
                                      C.POINTER(PcgResult)]
    L.orc_von_mises.argtypes = [_vp, _dp, _dp]
    L.orc_average_element_to_node.argtypes = [_vp, _dp, _dp]
    L.orc_sample_nodal_scalar.restype = C.c_double
    L.orc_sample_nodal_scalar.argtypes = [_vp, _dp] + [C.c_float] * 6
    L.orc_filter_create.argtypes = [C.c_int, C.c_int, C.c_int, _up, _ip, C.c_int, C.c_double,
                                    C.POINTER(C.c_void_p)]
    L.orc_filter_destroy.argtypes = [_vp]
    L.orc_filter_nnz.argtypes = [_vp]
    L.orc_filter_s.argtypes = [_vp, _dp]
    L.orc_filter_apply.argtypes = [_vp, _dp, _dp]
    L.orc_filter_apply_adjoint.argtypes = [_vp, _dp, _dp]
    L.orc_sensitivities.argtypes = [_vp, _dp, _dp, _dp, C.c_double]
    L.orc_oc_update.argtypes = [_dp, _dp, _dp, _vp, _dp, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_int, C.c_double, C.POINTER(C.c_int)]
    L.orc_oc_update.restype = C.c_double
    L.orc_heaviside_project.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double]
    L.orc_heaviside_derivative.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double]
    L.orc_passive_region.argtypes = [C.c_int, C.c_int, C.c_int, _ip, C.c_int, C.c_int, _up]
    L.orc_simp_run.argtypes = [_vp, _dp, C.POINTER(SimpOptions), _vp, _vp, C.c_int, _vp, C.c_int, _vp,
                               C.c_int, C.POINTER(Snapshot), C.POINTER(C.c_int), C.POINTER(C.c_int),
                               _dp, _dp, _dp]
    L.orc_simp_run_cases.argtypes = [_vp, _dp, C.c_int, C.POINTER(SimpOptions), _vp, _vp, C.c_int, _vp, C.c_int,
                                     _vp, C.c_int, C.POINTER(Snapshot), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     _dp, _dp, _dp]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"oracle {what} failed with status {code}")
        self.code = code


def pow_penal(r, p):
    """rho^p as the oracle's stiffness and sensitivities form it (orc_pow_penal_n)."""
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty_like(r)
    lib().orc_pow_penal_n(r, r.size, float(p), out)
    return out


def hex8_ke(h=1.0, nu=0.3):
    ke = np.zeros(576)
    lib().orc_hex8_ke(float(h), float(nu), ke)
    return ke.reshape(24, 24)


def hex8_bc(h=1.0):
    bc = np.zeros(144)
    lib().orc_hex8_bc(float(h), bc)
    return bc.reshape(6, 24)


def assemble_forces(n_nodes_total, cases):
    """cases: list of (nodes int array, areas float64 array, (fx, fy, fz) float32)."""
    offs = [0]
    nodes, areas, forces = [], [], []
    for nd, ar, fv in cases:
        nodes.append(np.asarray(nd, np.int32))
        areas.append(np.asarray(ar, np.float64))
        offs.append(offs[-1] + len(nd))
        forces.extend(np.asarray(fv, np.float32).tolist())
    f = np.zeros(3 * n_nodes_total)
    st = lib().orc_assemble_forces(n_nodes_total, len(cases), np.asarray(offs, np.int32),
                                   np.concatenate(nodes).astype(np.int32) if nodes else np.zeros(1, np.int32),
                                   np.concatenate(areas) if areas else np.zeros(1),
                                   np.asarray(forces, np.float32) if forces else np.zeros(3, np.float32), f)
    if st:
        raise OracleError(st, "assemble_forces")
    return f


class Model:
    """FeModel + FemOperator state (Fem/FeModel.cs:43-184, Fem/FemOperator.cs:33-108)."""

    def __init__(self, nx, ny, nz, domain=None, fixed_dofs=(), h=1.0, e0=1.0, nu=0.3, emin_ratio=1e-9,
                 penal=3.0, density=None):
        self.nx, self.ny, self.nz, self.h = nx, ny, nz, h
        self.e0, self.nu, self.emin_ratio, self.penal = e0, nu, emin_ratio, penal
        self.n_elem = nx * ny * nz
        self.n_nodes = (nx + 1) * (ny + 1) * (nz + 1)
        self.n_dofs = 3 * self.n_nodes
        dom = np.ones(self.n_elem, np.uint8) if domain is None else np.ascontiguousarray(domain, np.uint8)
        self.domain = dom
        fx = np.ascontiguousarray(np.asarray(fixed_dofs, np.int32))
        self.fixed_dofs = fx
        den = None if density is None else np.ascontiguousarray(density, np.float64)
        h_ = C.c_void_p()
        st = lib().orc_model_create(nx, ny, nz, float(h), float(e0), float(nu), float(emin_ratio), float(penal),
                                    dom, _ptr(fx) if fx.size else None, int(fx.size), _ptr(den), C.byref(h_))
        if st:
            raise OracleError(st, "model_create")
        self._h = h_

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_model_destroy(self._h)
            self._h = None

    @property
    def active_count(self):
        return lib().orc_model_active_count(self._h)

    @property
    def active_fixed_count(self):
        return lib().orc_model_active_fixed_count(self._h)

    @property
    def active_elements(self):
        out = np.zeros(self.active_count, np.int32)
        lib().orc_model_active_elements(self._h, out)
        return out

    @property
    def constrained(self):
        out = np.zeros(self.n_dofs, np.uint8)
        lib().orc_model_constrained(self._h, out)
        return out.astype(bool)

    @property
    def active_nodes(self):
        out = np.zeros(self.n_nodes, np.uint8)
        lib().orc_model_active_nodes(self._h, out)
        return out.astype(bool)

    @property
    def diag_a(self):
        out = np.zeros(self.n_dofs)
        lib().orc_model_diag_a(self._h, out)
        return out

    @property
    def ke(self):
        p = lib().orc_model_ke(self._h)
        return np.ctypeslib.as_array(p, shape=(576,)).copy().reshape(24, 24)

    def update_densities(self, density_box, penal):
        lib().orc_update_densities(self._h, np.ascontiguousarray(density_box, np.float64), float(penal))
        self.penal = penal

    def apply_k(self, x, threads=0):
        y = np.zeros(self.n_dofs)
        lib().orc_apply_k(self._h, np.ascontiguousarray(x, np.float64), y, int(threads))
        return y

    def apply_a(self, x, projected=False, threads=0):
        y = np.zeros(self.n_dofs)
        lib().orc_apply_a(self._h, np.ascontiguousarray(x, np.float64), y, int(bool(projected)), int(threads))
        return y

    def apply_k_scalar(self, x):
        y = np.zeros(self.n_dofs)
        lib().orc_apply_k_scalar(self._h, np.ascontiguousarray(x, np.float64), y)
        return y

    def pcg(self, b, x0=None, tol=1e-6, max_iter=5000, threads=0):
        x = np.zeros(self.n_dofs) if x0 is None else np.array(x0, np.float64)
        res = PcgResult()
        st = lib().orc_pcg_solve(self._h, np.ascontiguousarray(b, np.float64), x, float(tol), int(max_iter),
                                 int(threads), C.byref(res))
        if st:
            raise OracleError(st, "pcg_solve")
        return x, res

    def static_analysis(self, f, u0=None, tol=1e-6, max_iter=5000, throw_on_nonconv=True, threads=0):
        u = np.zeros(self.n_dofs) if u0 is None else np.array(u0, np.float64)
        react = np.zeros(self.n_dofs)
        vm = np.zeros(self.n_elem)
        comp, mx = C.c_double(), C.c_double()
        res = PcgResult()
        st = lib().orc_static_analysis(self._h, np.ascontiguousarray(f, np.float64), u, float(tol), int(max_iter),
                                       int(bool(throw_on_nonconv)), int(threads), C.byref(comp), C.byref(mx),
                                       _ptr(react), _ptr(vm), C.byref(res))
        if st:
            raise OracleError(st, "static_analysis")
        return dict(u=u, compliance=comp.value, max_displacement=mx.value, reactions=react, von_mises=vm,
                    solver=res)

    def sensitivities(self, u, rho_phys, penal):
        dc = np.zeros(self.active_count)
        lib().orc_sensitivities(self._h, np.ascontiguousarray(u, np.float64),
                                np.ascontiguousarray(rho_phys, np.float64), dc, float(penal))
        return dc

    def average_element_to_node(self, elem_box):
        out = np.zeros(self.n_nodes)
        lib().orc_average_element_to_node(self._h, np.ascontiguousarray(elem_box, np.float64), out)
        return out

    def von_mises(self, u):
        vm = np.zeros(self.n_elem)
        lib().orc_von_mises(self._h, np.ascontiguousarray(u, np.float64), vm)
        return vm

    def sample_nodal_scalar(self, nodal, point, origin=(0.0, 0.0, 0.0)):
        return lib().orc_sample_nodal_scalar(self._h, np.ascontiguousarray(nodal, np.float64), *[float(v) for v in point],
                                             *[float(v) for v in origin])

    def simp(self, f, *, volume_fraction=0.4, filter_radius_mm=2.0, penal=3.0, move_limit=0.2,
             max_iterations=200, solver_tol_initial=1e-4, solver_tol_final=1e-6, solver_max_iterations=5000,
             rho_min=1e-3, protect_boundary_hinges=True, start_iteration=1, enforce_min_filter_radius=True,
             continuation=False, passive_mask=None, patch_nodes=None, initial_density=None, warm_u=None,
             threads=0, enable_heaviside=False):
        o = SimpOptions(volume_fraction, filter_radius_mm, penal, move_limit, max_iterations, solver_tol_initial,
                        solver_tol_final, solver_max_iterations, rho_min, int(protect_boundary_hinges),
                        start_iteration, int(enforce_min_filter_radius), int(continuation), int(enable_heaviside))
        snaps = (Snapshot * max(1, max_iterations - start_iteration + 1))()
        nsn, conv = C.c_int(), C.c_int()
        na = self.active_count
        f = np.ascontiguousarray(f, np.float64)
        n_cases = 1 if f.ndim == 1 else f.shape[0]  # a 2-D f holds one load case per row (summed compliance)
        rho, rb, u = np.zeros(na), np.zeros(na), np.zeros((n_cases, self.n_dofs) if f.ndim == 2 else self.n_dofs)
        pm = None if passive_mask is None else np.ascontiguousarray(passive_mask, np.uint8)
        pn = np.zeros(0, np.int32) if patch_nodes is None else np.ascontiguousarray(patch_nodes, np.int32)
        idn = None if initial_density is None else np.ascontiguousarray(initial_density, np.float32)
        wu = None if warm_u is None else np.ascontiguousarray(warm_u, np.float64)
        st = lib().orc_simp_run_cases(self._h, f, int(n_cases), C.byref(o), _ptr(pm), _ptr(pn),
                                      int(pn.size), _ptr(idn), 0 if idn is None else int(idn.size), _ptr(wu),
                                      int(threads), snaps, C.byref(nsn), C.byref(conv), rho, rb, u)
        hist = [dict(iteration=s.iteration, compliance=s.compliance, volume_fraction=s.volume_fraction,
                     max_change=s.max_change, grey_level=s.grey_level, solver_iterations=s.solver_iterations,
                     solver_residual=s.solver_residual, max_displacement=s.max_displacement, lam=s.lambda_,
                     solver_tolerance=s.solver_tolerance)
                for s in snaps[:nsn.value]]
        return dict(status=st, snapshots=hist, converged=bool(conv.value), rho=rho, rho_bar=rb, u=u)


class Filter:
    """DensityFilter (Optimization/DensityFilter.cs:34-303)."""

    def __init__(self, nx, ny, nz, domain, active, rmin):
        self.n = len(active)
        h_ = C.c_void_p()
        st = lib().orc_filter_create(nx, ny, nz, np.ascontiguousarray(domain, np.uint8),
                                     np.ascontiguousarray(active, np.int32), self.n, float(rmin), C.byref(h_))
        if st:
            raise OracleError(st, "filter_create")
        self._h = h_

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_filter_destroy(self._h)
            self._h = None

    @property
    def nnz(self):
        return lib().orc_filter_nnz(self._h)

    @property
    def s(self):
        out = np.zeros(self.n)
        lib().orc_filter_s(self._h, out)
        return out

    def apply(self, x):
        y = np.zeros(self.n)
        lib().orc_filter_apply(self._h, np.ascontiguousarray(x, np.float64), y)
        return y

    def apply_adjoint(self, x):
        y = np.zeros(self.n)
        lib().orc_filter_apply_adjoint(self._h, np.ascontiguousarray(x, np.float64), y)
        return y


def heaviside_project(rho_tilde, beta, eta=0.5):
    rt = np.ascontiguousarray(rho_tilde, np.float64)
    out = np.zeros_like(rt)
    lib().orc_heaviside_project(rt, out, rt.size, float(beta), float(eta))
    return out


def heaviside_derivative(rho_tilde, beta, eta=0.5):
    rt = np.ascontiguousarray(rho_tilde, np.float64)
    out = np.zeros_like(rt)
    lib().orc_heaviside_derivative(rt, out, rt.size, float(beta), float(eta))
    return out


def oc_update(rho, dc, s, passive=None, vf=0.3, move=0.2, rho_min=1e-3, lambda_guess=1.0, max_bisect=60,
              vol_tol=1e-8):
    rho = np.ascontiguousarray(rho, np.float64)
    out = np.zeros_like(rho)
    ne = C.c_int()
    pm = None if passive is None else np.ascontiguousarray(passive, np.uint8)
    lam = lib().orc_oc_update(rho, np.ascontiguousarray(dc, np.float64), np.ascontiguousarray(s, np.float64),
                              _ptr(pm), out, rho.size, float(vf), float(move), float(rho_min), float(lambda_guess),
                              int(max_bisect), float(vol_tol), C.byref(ne))
    return out, lam, ne.value


def passive_region(nx, ny, nz, nodes, radius=2):
    mask = np.zeros(nx * ny * nz, np.uint8)
    nd = np.ascontiguousarray(nodes, np.int32)
    lib().orc_passive_region(nx, ny, nz, nd, nd.size, int(radius), mask)
    return mask
