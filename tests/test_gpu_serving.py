"""Job server: jobs submitted while others run are admitted and overlap on the device, and every
job's result is bit-identical to running it alone."""
import numpy as np
import pytest

import mycel_amd as M
from mycel_amd import OptimizerOptions, StaticAnalysis, StaticAnalysisOptions, TopologyOptimizer
from mycel_amd import workloads as W
from mycel_amd.serve import DONE, FAILED, JobServer

pytestmark = pytest.mark.gpu


def _stream(n, seed=11):
    return W.serving_stream(n, seed=seed, min_el=1500, max_el=12000, simp_iterations=4)


def _options(j, precision=M.PRECISION_FP64):
    return OptimizerOptions(VolumeFraction=j.volume_fraction, FilterRadiusMm=j.rmin, MaxIterations=j.simp_iterations,
                            SolverToleranceInitial=1e-8, SolverToleranceFinal=1e-8, SolverMaxIterations=50000,
                            ProtectBoundaryHinges=False, Precision=precision)


@pytest.mark.parametrize("precision", [M.PRECISION_FP64, M.PRECISION_MIXED_GMG], ids=["fp64", "gmg"])
def test_jobs_alone_equal_jobs_in_the_mix(gpu, precision):
    jobs = _stream(8)
    alone = []
    for j in jobs:
        fm = j.problem.fe_model()
        if j.kind == "fea":
            r = StaticAnalysis.Run(fm, StaticAnalysisOptions(Tolerance=1e-8, MaxIterations=50000, Precision=precision))
            alone.append((r.U, None, r.Compliance))
        else:
            r = TopologyOptimizer.Run(fm, _options(j, precision))
            alone.append((r.FinalDisplacements, r.FinalPhysical64, r.FinalCompliance))
    with JobServer(workers=4) as sv:
        ids = []
        for j in jobs:  # submitted back to back: the later ones queue while the first ones run
            fm = j.problem.fe_model()
            ids.append(sv.submit_fea(fm, tol=1e-8, max_iter=50000, precision=precision) if j.kind == "fea" else
                       sv.submit_simp(fm, _options(j, precision)))
        peak_active = 0
        for jid in ids:
            peak_active = max(peak_active, sv.stats()["active"])
            info = sv.wait(jid)
            assert info.state == DONE, info
        for (u0, rb0, c0), jid, j in zip(alone, ids, jobs):
            u, rb = sv.fetch(jid)
            info = sv.poll(jid)
            assert np.array_equal(u, u0), j.problem.name
            assert info.compliance == c0
            if rb0 is not None:
                assert np.array_equal(rb, rb0)
            sv.release(jid)
    print(f"[serving] peak concurrently active jobs: {peak_active}", flush=True)


def test_admission_by_budget_and_failures(gpu):
    jobs = _stream(4, seed=3)
    with JobServer(hbm_budget_bytes=1.0, workers=4) as sv:  # every job exceeds the budget: one at a time
        ids = [sv.submit_fea(j.problem.fe_model(), tol=1e-6) for j in jobs]
        for jid in ids:
            assert sv.stats()["active"] <= 1
            assert sv.wait(jid).state == DONE
        bad = W.Problem(3, 2, 2)
        bad.fixed_dofs = np.array([0, 1, 2], np.int32)  # fewer than 6 fixed DOFs: InvalidOperation
        jid = sv.submit_fea(bad.fe_model(), tol=1e-6)
        info = sv.wait(jid)
        assert info.state == FAILED and info.status == 4
