"""The reference-order FP64 path on the GPU (MYCEL_PRECISION_FP64_ORDERED, MYCEL_MATVEC_REFERENCE_ORDER) against
the oracle, bit for bit.

These kernels form every sum in the order the reference forms it (colour-ordered K*u with 4-lane Ke rows,
8,192-DOF blocked dot products with four lane chains, sequential ||Pb||, OC volumes and compliance) and never
contract a*b + c, so the oracle's restatement and the GPU must agree exactly: K*u, DiagA, every PCG iterate
(same iteration count, same solution bits, same final residual) and the SIMP trajectory.
"""
import numpy as np
import pytest

import mycel_amd as M
from mycel_amd import FemOperator, PcgSolver, StaticAnalysis, StaticAnalysisOptions
from mycel_amd import workloads as W

from helpers import oracle_forces, oracle_model, random_field

pytestmark = pytest.mark.gpu


def _problems():
    tiny = W.Problem(5, 4, 3, domain=W.l_shape_domain(5, 4, 3), emin_ratio=1e-3)
    tiny.fixed_dofs = W.fix_all(W.face_nodes_x(tiny, 0))
    tiny.density = random_field(3, tiny.n_elem, 0.01, 1.0)
    c1 = W.cantilever_simp()
    c1.density = random_field(4, c1.n_elem, 0.01, 1.0)
    c2 = W.matvec_pcg_workstation()
    return [tiny, c1, c2]


IDS = ["lshape-5x4x3", "cantilever-60x20x20", "holed-80x40x40"]


@pytest.mark.parametrize("idx", [0, 1, 2], ids=IDS)
def test_reference_order_apply_k_bitwise(gpu, idx):
    p = _problems()[idx]
    m = oracle_model(p)
    op = FemOperator(p.fe_model())
    try:
        op.SetMatvecForm(M.MATVEC_REFERENCE_ORDER)
        x = random_field(30 + idx, p.n_dofs, -1, 1)
        xp = np.where(m.constrained, 0.0, x)
        assert np.array_equal(op.ApplyK(x), m.apply_k(x))
        assert np.array_equal(op.ApplyA(x), m.apply_a(x, projected=False))
        assert np.array_equal(op.ApplyA(xp, xIsAlreadyProjected=True), m.apply_a(xp, projected=True))
        assert np.array_equal(op.DiagA, m.diag_a)
    finally:
        op.close()


@pytest.mark.parametrize("idx", [0, 1, 2], ids=IDS)
def test_reference_order_pcg_bitwise(gpu, idx):
    """Cold and warm-started Jacobi-PCG: the iteration count, every bit of the solution and the final residual
    equal the oracle's; a solve stopped by the iteration limit as well."""
    p = _problems()[idx]
    if not p.loads:
        p.loads = [(np.arange(p.n_nodes - 4, p.n_nodes, dtype=np.int32), np.ones(4), np.array([0, 0, -1], np.float32))]
    m = oracle_model(p)
    f = oracle_forces(p)
    b = np.where(m.constrained, 0.0, f)
    op = FemOperator(p.fe_model())
    try:
        for x0, tol, max_it in ((None, 1e-8, 20000), (random_field(40 + idx, p.n_dofs, -1e-3, 1e-3), 1e-6, 20000),
                                (None, 1e-12, 37)):
            u = np.zeros(p.n_dofs) if x0 is None else x0.copy()
            r = PcgSolver().Solve(op, b, u, tol=tol, maxIterations=max_it, precision=M.PRECISION_FP64_ORDERED)
            ur, rr = m.pcg(b, x0=x0, tol=tol, max_iter=max_it)
            print(f"[ordered] {p.name or 'tiny'} tol={tol:g}: {r.Iterations} vs {rr.iterations} iterations", flush=True)
            assert r.Iterations == rr.iterations and r.MessageCode == rr.message
            assert r.FinalResidual == rr.final_residual
            assert np.array_equal(u, ur)
    finally:
        op.close()


def test_reference_order_static_analysis(gpu):
    """StaticAnalysis in the reference order: the solution, the compliance F.u summed in index order and the
    reactions from the colour-ordered K*u equal the oracle's."""
    p = W.cantilever_simp()
    p.density = random_field(5, p.n_elem, 0.01, 1.0)
    fm = p.fe_model()
    m = oracle_model(p)
    f = oracle_forces(p)
    res = StaticAnalysis.Run(fm, StaticAnalysisOptions(Tolerance=1e-8, MaxIterations=20000,
                                                       Precision=M.PRECISION_FP64_ORDERED))
    sa = m.static_analysis(f, tol=1e-8, max_iter=20000)
    assert res.SolverInfo.Iterations == sa["solver"].iterations
    assert np.array_equal(res.U, sa["u"])
    assert res.Compliance == sa["compliance"] and res.MaxDisplacement == sa["max_displacement"]
    assert np.array_equal(res.Reactions[p.fixed_dofs], sa["reactions"][p.fixed_dofs])
