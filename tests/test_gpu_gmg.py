"""The multigrid-preconditioned solver (MYCEL_PRECISION_MIXED_GMG) through the C ABI.

The FP64 outer iteration must deliver what the reference's FP64 Jacobi-PCG delivers: a solution whose true
FP64 residual, checked with the oracle's operator, is at the requested tolerance, whose compliance matches
the oracle solved to 1e-12 within 1e-8 (relative) and whose displacements match within 1e-7. The FP32
V-cycle only changes how many iterations that takes.
"""
import numpy as np
import pytest

import mycel_amd as M
from helpers import oracle_forces, oracle_model, random_field, rel
from mycel_amd import (ArgumentError, FemOperator, MultigridPreconditioner, OptimizerOptions, PcgSolver,
                       TopologyOptimizer)
from mycel_amd import workloads as W

pytestmark = pytest.mark.gpu


def _true_relres(m, b, u):
    Ku = m.apply_a(u, projected=True)
    return np.linalg.norm(np.where(m.constrained, 0.0, b - Ku)) / np.linalg.norm(b)


def _check_against_oracle(p, tol=1e-8, max_gmg_iters=200):
    m = oracle_model(p)
    f = oracle_forces(p)
    b = np.where(m.constrained, 0.0, f)
    u_ref, r_ref = m.pcg(b, tol=1e-12, max_iter=400000)
    assert r_ref.converged
    op = FemOperator(p.fe_model())
    try:
        u = np.zeros(p.n_dofs)
        r = PcgSolver().Solve(op, b, u, MultigridPreconditioner(), tol=tol, maxIterations=100000)
    finally:
        op.close()
    c_ref = f @ u_ref
    e_c = abs(f @ u - c_ref) / abs(c_ref)
    tr = _true_relres(m, b, u)
    print(f"[gmg] {p.name or 'problem'} dofs={p.n_dofs} it={r.Iterations} fallback={r.FallbackIterations} "
          f"relres={r.FinalResidual:.2e} true={tr:.2e} compliance err={e_c:.2e} u err={rel(u, u_ref):.2e} "
          f"(oracle Jacobi it at 1e-12: {r_ref.iterations})", flush=True)
    assert r.Converged and r.MessageCode == 0
    assert r.FinalResidual <= tol
    assert tr <= 2 * tol
    assert e_c <= 1e-8
    assert rel(u, u_ref) <= 1e-7
    assert r.Iterations <= max_gmg_iters
    return r


def test_gmg_cantilever_random_density(gpu):
    p = W.cantilever_simp()
    p.density = random_field(4, p.n_elem, 0.01, 1.0)
    _check_against_oracle(p, max_gmg_iters=120)


def test_gmg_holed_box_workstation(gpu):
    """The ~350k-DOF holed box: a few dozen V-cycles instead of ~1,900 Jacobi iterations."""
    _check_against_oracle(W.matvec_pcg_workstation(), max_gmg_iters=150)


def test_gmg_bimodal_contrast(gpu):
    """The 1e-9 stiffness-contrast field: rediscretised coarse levels are weak here, but the FP64 outer
    iteration still has to reach the same answer."""
    _check_against_oracle(W.precision_variant(), max_gmg_iters=60000)


def test_gmg_refresh_after_density_update_and_warm_start(gpu):
    """The hierarchy follows UpdateDensities (the same answer as a fresh operator), and a warm start from the
    converged solution stops at once."""
    p = W.cantilever_simp(nx=30, ny=10, nz=10)
    p.density = random_field(5, p.n_elem, 0.05, 1.0)
    rho2 = random_field(6, p.n_elem, 0.05, 1.0)
    m = oracle_model(p)
    b = np.where(m.constrained, 0.0, oracle_forces(p))
    solver = PcgSolver()
    op = FemOperator(p.fe_model())
    try:
        u1 = np.zeros(p.n_dofs)
        assert solver.Solve(op, b, u1, MultigridPreconditioner(), tol=1e-10, maxIterations=10000).Converged
        op.UpdateDensities(rho2, p.penal)
        u2 = np.zeros(p.n_dofs)
        r2 = solver.Solve(op, b, u2, MultigridPreconditioner(), tol=1e-10, maxIterations=10000)
        warm = u2.copy()
        rw = solver.Solve(op, b, warm, MultigridPreconditioner(), tol=1e-10, maxIterations=10000)
    finally:
        op.close()
    p2 = W.cantilever_simp(nx=30, ny=10, nz=10)
    p2.density = rho2
    op2 = FemOperator(p2.fe_model())
    try:
        u3 = np.zeros(p.n_dofs)
        r3 = solver.Solve(op2, b, u3, MultigridPreconditioner(), tol=1e-10, maxIterations=10000)
    finally:
        op2.close()
    print(f"[gmg] refresh it={r2.Iterations} fresh it={r3.Iterations} warm it={rw.Iterations} "
          f"msg={rw.Message!r}", flush=True)
    assert r2.Converged and r3.Converged
    assert r2.Iterations == r3.Iterations
    assert np.array_equal(u2, u3)  # deterministic: same hierarchy, same launches, same order
    assert rw.Converged and rw.Iterations <= 1


def test_gmg_rejects_grid_that_cannot_coarsen(gpu):
    p = W.Problem(30, 30, 1)
    p.fixed_dofs = W.fix_all(W.face_nodes_x(p, 0))
    b = np.zeros(p.n_dofs)
    b[3 * p.node(p.nx, p.ny, p.nz) + 2] = -1.0
    op = FemOperator(p.fe_model())
    try:
        for _ in range(2):  # a failed build leaves no half-built hierarchy behind
            with pytest.raises(ArgumentError):
                PcgSolver().Solve(op, b, np.zeros(p.n_dofs), MultigridPreconditioner(), tol=1e-8)
        assert PcgSolver().Solve(op, b, np.zeros(p.n_dofs), tol=1e-8, maxIterations=100000).Converged
    finally:
        op.close()


def test_gmg_simp_matches_fp64_history(gpu):
    """Five SIMP iterations with the multigrid solver follow the FP64 Jacobi run (both solve to 1e-10)."""
    p = W.cantilever_simp(nx=30, ny=10, nz=10)
    runs = {}
    for name, prec in (("fp64", M.PRECISION_FP64), ("gmg", M.PRECISION_MIXED_GMG)):
        o = OptimizerOptions(VolumeFraction=0.3, FilterRadiusMm=1.5, MaxIterations=5, SolverToleranceInitial=1e-10,
                             SolverToleranceFinal=1e-10, SolverMaxIterations=100000, ProtectBoundaryHinges=False,
                             Precision=prec)
        runs[name] = TopologyOptimizer.Run(p.fe_model(), o)
    c64 = np.array([s.Compliance for s in runs["fp64"].Snapshots])
    cg = np.array([s.Compliance for s in runs["gmg"].Snapshots])
    it64 = [s.SolverIterations for s in runs["fp64"].Snapshots]
    itg = [s.SolverIterations for s in runs["gmg"].Snapshots]
    print(f"[gmg] simp compliance rel diff {np.max(np.abs(cg - c64) / c64):.2e} it fp64={it64} gmg={itg}", flush=True)
    assert np.max(np.abs(cg - c64) / c64) <= 1e-7
    assert sum(itg) < sum(it64)
