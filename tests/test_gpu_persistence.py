"""A SIMP run on the GPU written to a run folder (Persistence/RunWriter.cs formats) and resumed from it.

The run folder must read back as written: every iteration frame is the 8-bit quantization of that
iteration's physical density with a matching checksum, metrics.csv carries the snapshots' values, and the
checkpoint holds the final float32 design and physical densities. A resume from the checkpoint's FP64
extension (design, displacements, OC multiplier) continues the run exactly: its iterations are bit-identical
to those of the uninterrupted run. A resume from the float32 design alone, which is what the reference's
checkpoint holds, stays within float32 rounding of it.
"""
import dataclasses

import numpy as np
import pytest

import mycel_amd.persistence as P
from mycel_amd import OptimizerOptions, TopologyOptimizer
from mycel_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_simp_run_folder_and_exact_resume(gpu, tmp_path):
    p = W.cantilever_simp(nx=24, ny=8, nz=8)
    fm = p.fe_model()
    o = OptimizerOptions(VolumeFraction=0.3, FilterRadiusMm=1.5, MaxIterations=8, SolverToleranceInitial=1e-10,
                         SolverToleranceFinal=1e-10, SolverMaxIterations=100000, ProtectBoundaryHinges=False)
    full = TopologyOptimizer.Run(fm, o)
    assert full.TotalIterations == 8

    o5 = dataclasses.replace(o, MaxIterations=5)
    w = P.RunWriter(str(tmp_path / "run"), P.RunMetadata.for_grid(p.nx, p.ny, p.nz, options=o5), p.domain)
    part = TopologyOptimizer.Run(fm, o5, progress=w)
    w.WriteSnapshot(part.Snapshots[-1], isFinal=True, displacements=part.FinalDisplacements,
                    design64=part.FinalDesign)
    w.FinalizeRun(P.RunStatus.Completed)
    w.Dispose()

    r = P.RunReader.Open(str(tmp_path / "run"))
    assert r.Metadata.Status == P.RunStatus.Completed and r.GetAvailableIterations() == [1, 2, 3, 4, 5]
    rows = r.ReadMetrics()
    assert [x.Iteration for x in rows] == [1, 2, 3, 4, 5, 5]
    for s in part.Snapshots:
        hdr, dens = r.ReadIteration(s.Iteration)
        q, h = P.quantize(s.PhysicalDensity)
        assert hdr.Checksum == h and np.array_equal(dens, q.astype(np.float32) * np.float32(1 / 255))
        assert hdr.Compliance == s.Compliance and rows[s.Iteration - 1].Compliance == s.Compliance
        assert s.DesignDensity is not None and s.DesignDensity.shape == s.PhysicalDensity.shape
    ck = r.ReadLatestCheckpoint()
    last = part.Snapshots[-1]
    assert ck.Iteration == 5 and np.array_equal(ck.DesignDensity, last.DesignDensity)
    assert np.array_equal(ck.PhysicalDensity, last.PhysicalDensity)
    assert np.array_equal(ck.DesignDensity64, part.FinalDesign) and ck.Lambda == last.Lambda
    # The snapshot's float32 design is the final FP64 design rounded (TopologyOptimizer.cs:343-347).
    act = fm.ActiveElements
    assert np.array_equal(last.DesignDensity[act], part.FinalDesign.astype(np.float32))

    exact = TopologyOptimizer.Run(fm, P.resume_options(ck, o))
    c_full = [s.Compliance for s in full.Snapshots[5:]]
    c_exact = [s.Compliance for s in exact.Snapshots]
    print(f"[persist] full {c_full} resumed {c_exact}", flush=True)
    assert [s.Iteration for s in exact.Snapshots] == [6, 7, 8]
    assert c_exact == c_full
    assert np.array_equal(exact.FinalDesign, full.FinalDesign)

    plain = dataclasses.replace(ck, DesignDensity64=None, Displacements=None, Lambda=None)
    approx = TopologyOptimizer.Run(fm, P.resume_options(plain, o))
    c_approx = np.array([s.Compliance for s in approx.Snapshots])
    rel = np.max(np.abs(c_approx - np.array(c_full)) / np.array(c_full))
    print(f"[persist] float32 resume rel diff {rel:.2e}", flush=True)
    assert rel < 1e-5


@pytest.mark.parametrize("case", ["default-schedule", "heaviside-past-25"])
def test_exact_resume_continues_schedule_beta_and_stop_test(gpu, tmp_path, case):
    """A resume from the checkpoint's FP64 extension and loop state (max change, Heaviside beta, the last six
    compliances) continues the uninterrupted run bit for bit: under the default 1e-4 -> 1e-6 tolerance schedule,
    whose first resumed solve takes its tolerance from the saved max change, and with Heaviside projection
    across the beta doubling at iteration 25."""
    p = W.cantilever_simp(nx=24, ny=8, nz=8)
    fm = p.fe_model()
    heav = case == "heaviside-past-25"
    total = 30 if heav else 40
    o = OptimizerOptions(VolumeFraction=0.3, FilterRadiusMm=1.5, MaxIterations=total, ProtectBoundaryHinges=False,
                         EnableHeaviside=heav)
    full = TopologyOptimizer.Run(fm, o)
    # Cut where the schedule has left its 1e-4 end (max change below 0.1), so the first resumed solve's tolerance
    # depends on the saved max change; with Heaviside, before the beta doubling at 25.
    cut = 20 if heav else next(k for k in range(8, total - 5) if full.Snapshots[k - 1].MaxChange < 0.08)
    o_cut = dataclasses.replace(o, MaxIterations=cut)
    w = P.RunWriter(str(tmp_path / "run"), P.RunMetadata.for_grid(p.nx, p.ny, p.nz, options=o_cut), p.domain)
    part = TopologyOptimizer.Run(fm, o_cut, progress=w)
    w.WriteSnapshot(part.Snapshots[-1], isFinal=True, displacements=part.FinalDisplacements,
                    design64=part.FinalDesign)
    w.Dispose()
    ck = P.RunReader.Open(str(tmp_path / "run")).ReadLatestCheckpoint()
    assert ck.MaxChange == part.Snapshots[-1].MaxChange
    assert ck.ComplianceHistory == [s.Compliance for s in part.Snapshots[-6:]]
    resumed = TopologyOptimizer.Run(fm, P.resume_options(ck, o))
    c_full = [s.Compliance for s in full.Snapshots[cut:]]
    c_res = [s.Compliance for s in resumed.Snapshots]
    it_full = [s.SolverIterations for s in full.Snapshots[cut:]]
    it_res = [s.SolverIterations for s in resumed.Snapshots]
    print(f"[resume {case}] full {len(full.Snapshots)} it (converged {full.Converged}), resumed from {cut}: "
          f"solver iterations {it_res} vs {it_full}", flush=True)
    assert [s.Iteration for s in resumed.Snapshots] == [s.Iteration for s in full.Snapshots[cut:]]
    assert c_res == c_full and it_res == it_full
    assert resumed.Converged == full.Converged
    assert np.array_equal(resumed.FinalDesign, full.FinalDesign)
    # Without the loop state (the reference's own resume semantics) the first resumed solve restarts the
    # schedule at the initial tolerance, so the run differs.
    if not heav:
        plain = dataclasses.replace(ck, MaxChange=None, BetaNext=None, ComplianceHistory=None)
        other = TopologyOptimizer.Run(fm, P.resume_options(plain, o))
        assert other.Snapshots[0].SolverIterations < resumed.Snapshots[0].SolverIterations  # solved to 1e-4
