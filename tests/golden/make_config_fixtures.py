"""Writes the oracle's results for the full-size configuration parity tests (tests/test_gpu_configs.py).

Run from the repository root: python tests/golden/make_config_fixtures.py [name ...]. Each fixture is the FP64
oracle (oracle/mycel_oracle.c, the restatement of the reference's TopologyOptimizer.Run and its callees) on a
seeded workload from mycel_amd.workloads; the GPU tests rebuild the same inputs from the same seeds. These runs
take minutes of CPU time, which is why their outputs are committed instead of recomputed on every GPU run.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle as O  # noqa: E402
from helpers import oracle_forces, oracle_model  # noqa: E402
from mycel_amd import workloads as W  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# The SIMP cantilever configuration: 60x20x20, vf 0.3, rmin 1.5, p 3, move 0.2, solves to 1e-8 (both ends of
# the tolerance schedule), 20,000 solver iterations, hinges off, 30 iterations.
C1 = dict(volume_fraction=0.3, filter_radius_mm=1.5, penal=3.0, move_limit=0.2, max_iterations=30,
          solver_tol_initial=1e-8, solver_tol_final=1e-8, solver_max_iterations=20000, protect_boundary_hinges=False)
# The same mesh under the reference's defaults for hinge protection and the tolerance schedule (1e-4 -> 1e-6,
# OptimizerOptions.cs:5-26), 30 iterations.
C1_HINGES = dict(C1, solver_tol_initial=1e-4, solver_tol_final=1e-6, solver_max_iterations=5000,
                 protect_boundary_hinges=True)
# p-continuation (p = 1 below iteration 20, 2 below 40, then 3: TopologyOptimizer.cs:177-183) for 42 iterations.
C1_CONTINUATION = dict(C1, max_iterations=42, continuation=True)
# The multi-load-case configuration's first iteration: 100x50x50, 16 cases, vf 0.25, rmin 2.0.
C4_FIRST = dict(volume_fraction=0.25, filter_radius_mm=2.0, penal=3.0, move_limit=0.2, max_iterations=1,
                solver_tol_initial=1e-8, solver_tol_final=1e-8, solver_max_iterations=200000,
                protect_boundary_hinges=False)
# Job 21 of the serving stream (seed 5): a 54x27x27 SIMP job, 40 iterations at 1e-8 (the server's options).
C5_JOB = 21


def c5_job():
    j = W.serving_stream(512, seed=5)[C5_JOB]
    kw = dict(volume_fraction=j.volume_fraction, filter_radius_mm=j.rmin, penal=3.0, move_limit=0.2,
              max_iterations=j.simp_iterations, solver_tol_initial=1e-8, solver_tol_final=1e-8,
              solver_max_iterations=500_000, protect_boundary_hinges=False)
    return j.problem, kw


def _simp(p, f, kw, patch_nodes=None):
    m = oracle_model(p)
    t = time.time()
    ref = m.simp(f, patch_nodes=patch_nodes, **kw)
    assert ref["status"] == 0, ref["status"]
    sn = ref["snapshots"]
    out = dict(compliance=np.array([s["compliance"] for s in sn]),
               solver_iterations=np.array([s["solver_iterations"] for s in sn], np.int64),
               max_change=np.array([s["max_change"] for s in sn]),
               volume_fraction=np.array([s["volume_fraction"] for s in sn]),
               lam=np.array([s["lam"] for s in sn]), rho=ref["rho"], rho_bar=ref["rho_bar"])
    return out, ref, time.time() - t


C4_SAMPLE = 8  # the 250,000-element density fields are kept as every 8th element plus their sums
C3_SAMPLE = 8  # the bimodal solve's displacements are kept as every 8th DOF


def _sample(out):
    for k in ("rho", "rho_bar"):
        out[k + "_sum"] = np.array(out[k].sum())
        out[k] = out[k][::C4_SAMPLE]


def make(name):
    if name == "c3_solve":  # the bimodal field solved to 1e-12 (compliance; every 8th DOF of u)
        p = W.precision_variant()
        m = oracle_model(p)
        f = oracle_forces(p)
        t = time.time()
        u, r = m.pcg(np.where(m.constrained, 0.0, f), tol=1e-12, max_iter=500000)
        assert r.converged
        out = dict(compliance=np.array([f @ u]), u_sample=u[::C3_SAMPLE], iterations=np.array([r.iterations]))
        dt = time.time() - t
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {r.iterations} iterations, {dt:.1f} s, {os.path.getsize(path) / 1024:.0f} KiB", flush=True)
        return
    if name == "c1_simp":
        p = W.cantilever_simp()
        out, _, dt = _simp(p, oracle_forces(p), C1)
    elif name == "c1_simp_hinges":
        p = W.cantilever_simp()
        out, _, dt = _simp(p, oracle_forces(p), C1_HINGES, patch_nodes=p.patch_nodes())
    elif name == "c1_simp_continuation":
        p = W.cantilever_simp()
        out, _, dt = _simp(p, oracle_forces(p), C1_CONTINUATION)
    elif name == "c4_first_iteration":
        p, cases = W.multi_rhs()
        f = W.case_forces(p, cases)
        out, ref, dt = _simp(p, f, C4_FIRST)
        out["case_compliance"] = np.einsum("kd,kd->k", f, ref["u"])
        _sample(out)
    elif name == "c5_job":
        p, kw = c5_job()
        out, _, dt = _simp(p, oracle_forces(p), kw)
    else:
        raise SystemExit(f"unknown fixture {name}")
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: {len(out['compliance'])} iterations, {dt:.1f} s, {os.path.getsize(path) / 1024:.0f} KiB",
          flush=True)


if __name__ == "__main__":
    O.build()
    for n in sys.argv[1:] or ["c1_simp", "c1_simp_hinges", "c1_simp_continuation", "c4_first_iteration",
                             "c5_job", "c3_solve"]:
        make(n)
