"""Packed multigrid solves: several systems on different grids advance together, one launch per kernel for all
of them, and each system's solution, iteration count and final residual are bit-identical to its solve alone
(mycel_pcg_solve_packed_d; the job server packs its jobs' solves through the same machinery)."""
import numpy as np
import pytest
import torch

import mycel_amd as M
from helpers import random_field
from mycel_amd import FemOperator, OptimizerOptions, PcgSolver
from mycel_amd import workloads as W
from mycel_amd.serve import DONE, JobServer

pytestmark = pytest.mark.gpu


def _problems():
    """Grids of different hierarchy depth and tile depth, an L-shaped domain, a grid too thin for the hierarchy
    (its system is solved by the FP64 Jacobi iteration, outside the pack) and a zero right-hand side."""
    out = []
    for k, (nx, ny, nz) in enumerate([(24, 8, 8), (37, 19, 11), (33, 21, 9), (60, 20, 20), (30, 10, 4)]):
        p = W.cantilever_simp(nx, ny, nz)
        if (nx, ny, nz) == (33, 21, 9):
            p.domain = W.l_shape_domain(nx, ny, nz)
        p.density = random_field(100 + k, p.n_elem, 0.01, 1.0)
        out.append((p, False))
    out.append((W.cantilever_simp(20, 10, 10), True))  # zero load: ends in the set-up
    return out


def _rhs(p, zero):
    fm = p.fe_model()
    f = np.zeros(p.n_dofs) if zero else fm.AssembleForces()
    return np.where(fm.ConstrainedDofs, 0.0, f)


@pytest.mark.parametrize("warm", [False, True], ids=["cold", "warm"])
def test_packed_solves_equal_solves_alone(gpu, warm):
    dev = torch.device("cuda:0")
    probs = _problems()
    ops = [FemOperator(p.fe_model()) for p, _ in probs]
    try:
        bs, x0s = [], []
        for k, ((p, zero), op) in enumerate(zip(probs, ops)):
            bs.append(torch.tensor(_rhs(p, zero), device=dev))
            x0 = 1e-3 * (random_field(7 + k, p.n_dofs) - 0.5) if warm else np.zeros(p.n_dofs)
            x0s.append(x0)
        solver = PcgSolver()
        alone = []
        for op, b, x0 in zip(ops, bs, x0s):
            x = torch.tensor(x0, device=dev)
            r = solver.solve_device(op, b, x, tol=1e-8, maxIterations=20000, precision=M.PRECISION_MIXED_GMG)
            alone.append((x.cpu().numpy(), r))
        xs = [torch.tensor(x0, device=dev) for x0 in x0s]
        packed = solver.solve_packed_device(ops, bs, xs, tol=1e-8, maxIterations=20000)
        torch.cuda.synchronize()
        for (p, _), (xa, ra), x, rp in zip(probs, alone, xs, packed):
            xp = x.cpu().numpy()
            print(f"[packed] {p.nx}x{p.ny}x{p.nz} it alone {ra.Iterations} packed {rp.Iterations} "
                  f"fallback {rp.FallbackIterations} msg {rp.MessageCode}", flush=True)
            assert np.array_equal(xp, xa), f"{p.nx}x{p.ny}x{p.nz}"
            assert (rp.Iterations, rp.MessageCode, rp.Converged, rp.FallbackIterations) == \
                (ra.Iterations, ra.MessageCode, ra.Converged, ra.FallbackIterations)
            assert rp.FinalResidual == ra.FinalResidual
        assert all(r.Converged for r in packed)
        assert packed[4].FallbackIterations == packed[4].Iterations > 0  # the thin grid: FP64 Jacobi
        assert packed[5].Iterations == 0  # zero right-hand side
    finally:
        for op in ops:
            op.close()


def test_packed_solve_rejects_a_handle_twice(gpu):
    p = W.cantilever_simp(24, 8, 8)
    op = FemOperator(p.fe_model())
    try:
        dev = torch.device("cuda:0")
        b = torch.tensor(_rhs(p, False), device=dev)
        x1, x2 = torch.zeros_like(b), torch.zeros_like(b)
        with pytest.raises(M.ArgumentError):
            PcgSolver().solve_packed_device([op, op], [b, b], [x1, x2], tol=1e-8)
    finally:
        op.close()


def _options(j):
    return OptimizerOptions(VolumeFraction=j.volume_fraction, FilterRadiusMm=j.rmin, MaxIterations=j.simp_iterations,
                            SolverToleranceInitial=1e-8, SolverToleranceFinal=1e-8, SolverMaxIterations=50000,
                            ProtectBoundaryHinges=False, Precision=M.PRECISION_MIXED_GMG)


def _run_server(jobs, packing):
    with JobServer(workers=6) as sv:
        sv.set_packing(packing)
        ids = []
        for j in jobs:
            fm = j.problem.fe_model()
            ids.append(sv.submit_fea(fm, tol=1e-8, max_iter=50000, precision=M.PRECISION_MIXED_GMG)
                       if j.kind == "fea" else sv.submit_simp(fm, _options(j)))
        out = []
        for jid in ids:
            info = sv.wait(jid)
            assert info.state == DONE, info
            u, rb = sv.fetch(jid)
            out.append((u, rb if info.simp_iterations else None, info.compliance, info.pcg_iterations))
            sv.release(jid)
        return out, sv.pack_stats()


def test_server_packing_is_bit_identical(gpu):
    """The same stream with the jobs' solves packed (the default) and each on its worker's stream."""
    jobs = W.serving_stream(10, seed=21, min_el=1500, max_el=20000, simp_iterations=4)
    on, st = _run_server(jobs, True)
    off, st_off = _run_server(jobs, False)
    for (u1, r1, c1, i1), (u0, r0, c0, i0), j in zip(on, off, jobs):
        assert np.array_equal(u1, u0), j.problem.name
        assert c1 == c0 and i1 == i0
        if r0 is not None:
            assert np.array_equal(r1, r0)
    print(f"[packed] server stats {st}", flush=True)
    assert st_off["batches"] == 0
    assert st["solves"] > 0 and st["batches"] > 0
    assert st["member_batches"] > st["batches"]  # solves of several jobs shared batches
