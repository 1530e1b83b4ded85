"""Benchmark: Jacobi-PCG solves/s on the ~350k-DOF holed box, on one MI355X.

Workload (the workstation matvec/PCG configuration): 80x40x40 Hex8 grid with a radius-12 through-hole
(110,080 active elements, 359,160 active / 408,483 box DOFs), i.i.d. densities U[0.01, 1], x = 0
clamped, 8 seeded 5x5-node patch loads; E0 = 1, nu = 0.3, Emin/E0 = 1e-9, p = 3.

A step is one cold PCG solve (u = 0) to relative residual 1e-8 with all inputs resident in HBM.
Prints one JSON line (metric, value = solves/s, roofline of the dominant K*u kernel measured live with
HIP events on the operator's stream, and the oracle CPU baseline timed on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--matvecs", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    return ap.parse_args()


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 PMC counters (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data.get(f"{kernel}@{workload}")
    except (OSError, ValueError):
        return None


def cpu_baseline(p, gpu_iters: int, budget_s: float):
    """The FP64 oracle (reference-algorithm CPU restatement: 8-colour K*u in 256-element chunks,
    8,192-DOF PCG reduction blocks) on this host's cores: matvecs and a bounded run of PCG iterations,
    scaled to solves/s with the iteration count of the same solve."""
    import numpy as np

    import oracle as O

    O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    m = O.Model(p.nx, p.ny, p.nz, p.domain, p.fixed_dofs, p.h, p.e0, p.nu, p.emin_ratio, p.penal, p.density)
    x = np.random.default_rng(0).standard_normal(p.n_dofs)
    m.apply_k(x, threads=threads)
    t0 = time.perf_counter()
    n_mv = 0
    while n_mv < 100 and time.perf_counter() - t0 < budget_s * 0.3:
        m.apply_k(x, threads=threads)
        n_mv += 1
    mv_ms = (time.perf_counter() - t0) / n_mv * 1e3
    f = O.assemble_forces(p.n_nodes, p.loads)
    b = np.where(m.constrained, 0.0, f)
    it = max(10, int(budget_s * 0.6 / (mv_ms * 1.3e-3)))
    t0 = time.perf_counter()
    _, res = m.pcg(b, tol=1e-30, max_iter=it, threads=threads)
    per_it = (time.perf_counter() - t0) / max(1, res.iterations)
    return {
        "value": 1.0 / (per_it * gpu_iters),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n_mv} K*u + {res.iterations} Jacobi-PCG iterations of the same problem on the FP64 oracle "
                  f"({threads} threads); solves/s = 1 / (s per iteration x {gpu_iters} iterations)",
        "ms_per_matvec": mv_ms,
        "ms_per_pcg_iteration": per_it * 1e3,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Independent jobs would shard across GPUs with no communication; see DESIGN.md.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from mycel_amd import FemOperator, PcgSolver
    from mycel_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    dev = torch.device("cuda:0")
    p = W.matvec_pcg_workstation(seed=args.seed)
    fm = p.fe_model()
    op = FemOperator(fm)
    info = op.info()
    f = fm.AssembleForces()
    b_t = torch.tensor(np.where(fm.ConstrainedDofs, 0.0, f), dtype=torch.float64, device=dev)
    x_t = torch.zeros(p.n_dofs, dtype=torch.float64, device=dev)
    y_t = torch.empty_like(x_t)
    xin = torch.tensor(np.random.default_rng(1).standard_normal(p.n_dofs), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    # ---- K*u: timed with HIP events on the operator's own stream --------------------------------
    stream = torch.cuda.ExternalStream(op.stream())
    for _ in range(3):
        op.apply_k_device(xin, y_t)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(args.matvecs):
        op.apply_k_device(xin, y_t)
    ev1.record(stream)
    ev1.synchronize()
    mv_ms = ev0.elapsed_time(ev1) / args.matvecs

    # ---- PCG solves --------------------------------------------------------------------------
    solver = PcgSolver()
    res = None
    for _ in range(args.warmup):
        x_t.zero_()
        torch.cuda.synchronize()
        res = solver.solve_device(op, b_t, x_t, tol=args.tol, maxIterations=200_000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        x_t.zero_()
        res = solver.solve_device(op, b_t, x_t, tol=args.tol, maxIterations=200_000)
        iters.append(res.Iterations)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    if not res.Converged:
        raise SystemExit(f"PCG did not converge: {res.Message}")
    ms_per_step = elapsed / args.steps * 1e3
    solves_per_s = args.steps / elapsed
    n_it = iters[-1]

    # ---- roofline of the dominant kernel (K*u) ---------------------------------------------------
    na, nn = info.active_elements, info.active_nodes
    bytes_mv = 2 * 3 * 8 * nn + 8 * na  # read x, write y over active nodes + one FP64 E per element
    flops_mv = 1152 * na
    achieved = bytes_mv / (mv_ms * 1e-3) / 1e9
    workload = f"holed-box-80x40x40"
    out = {
        "metric": "PCG solves/s @350k DOF, mixed vs FP64 at compliance <=1e-8 rel; p99 job latency",
        "value": solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "grid": [p.nx, p.ny, p.nz],
            "active_elements": na,
            "active_dofs": 3 * nn,
            "box_dofs": p.n_dofs,
            "density": "iid U[0.01,1] (SplitMix64 seed %d)" % args.seed,
            "emin_ratio": p.emin_ratio,
            "tol": args.tol,
            "precision": "fp64",
            "parallelism": "single-gpu",
        },
        "pcg_iterations": n_it,
        "ms_per_pcg_iteration": ms_per_step / n_it,
        "ms_per_matvec": mv_ms,
        "gdof_per_s_matvec": {"active": 3 * nn / (mv_ms * 1e-3) / 1e9, "box": p.n_dofs / (mv_ms * 1e-3) / 1e9},
        "roofline": {
            "kernel": "k_apply_k_f64<kApplyK>",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles("k_apply_k_f64", workload),
            "algorithmic_bytes_per_launch": bytes_mv,
            "avg_launch_ms": mv_ms,
            "fp64": {"achieved_tflops": flops_mv / (mv_ms * 1e-3) / 1e12, "peak_tflops": FP64_PEAK_TFLOPS,
                     "frac": flops_mv / (mv_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                     "flops_per_launch": flops_mv},
        },
        "cpu_baseline": None,
    }
    op.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, n_it, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
