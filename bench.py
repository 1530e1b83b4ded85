"""Benchmark: Jacobi-PCG solves/s on the ~350k-DOF holed box, on one MI355X, FP64 and mixed precision.

Workload (the workstation matvec/PCG configuration): 80x40x40 Hex8 grid with a radius-12 through-hole
(110,080 active elements, 359,160 active / 408,483 box DOFs), i.i.d. densities U[0.01, 1], x = 0
clamped, 8 seeded 5x5-node patch loads; E0 = 1, nu = 0.3, Emin/E0 = 1e-9, p = 3.

--mode solve (default): a step is one cold PCG solve (u = 0) to relative residual 1e-8 with all inputs
  resident in HBM, in the mixed-precision solver: FP64 flexible PCG (FP64 operator, residual, dot products
  and convergence test) preconditioned by one FP32 geometric-multigrid V-cycle per iteration. The plain
  FP64 Jacobi-PCG solver (the reference's algorithm) and the FP32-iteration solver with FP64 reliable
  updates are timed the same way beside it; all three are checked against an FP64 solve to 1e-12. The
  metric's job-latency part follows (unless --no-serve): the serving stream of --mode serve through the job
  server, its capacity at saturation and its p50/p99 latency at 80% of that capacity ("serving").
  Prints one JSON line.
--mode kernels: a short fixed run of each kernel (for rocprofv3 kernel traces and PMC counters),
  including a device copy of known size that calibrates the FETCH_SIZE / WRITE_SIZE counters.
--mode serve: a stream of --jobs (default 96; the configuration's 512 with --jobs 512) fresh full-box jobs
  (20k-250k elements; FEA solves and 40-iteration SIMP runs): every job alone in plain FP64 and in the multigrid
  solver (skipped with --no-alone), the job server's capacity with every job submitted at once (value, jobs/s;
  --workers resident jobs whose multigrid solves run packed, --packing off runs each on its worker's stream),
  and its p50/p99 latency under Poisson arrivals at 80% of that capacity, beside a replay of the same arrivals
  through one-at-a-time service.
--mode cases: the batched multi-RHS SIMP configuration: 100x50x50 grid (788,103 DOFs), x = 0 clamped, 16
  independent load cases (1-4 seeded 4x4-node patches each), 20 SIMP iterations on the summed compliance
  (vf 0.25, rmin 2.0, p = 3, solves to 1e-8 warm-started per case) with the multigrid solver; one SIMP
  iteration with the plain FP64 Jacobi solver is timed beside it. Reports SIMP iterations/s.
--mode precision: the 60x20x20 cantilever with a bimodal (~1e-9 contrast) density field, one cold solve to 1e-8
  per step in plain FP64 Jacobi-PCG, FP32 Jacobi-PCG with FP64 reliable updates, the same with bfloat16 search
  directions, the multigrid solver, and the multigrid solver with island deflation; errors against an FP64
  solve to 1e-12.
--mode sweep: K*u (FP64 and FP32) on the holed box and on solid boxes up to 256x128x128 (12.7M DOFs), 100
  launches each timed with HIP events on the operator's stream; reports time per launch and the fraction of
  the HBM and FP64 rooflines at each size (the fixed launch and latency cost of one K*u is amortised as the
  grid grows). Prints one JSON line.
--mode packed: --pack (16) independent systems of the serving mix, one cold multigrid solve each, one at a time and
  then together in packed launches (the job server's machinery); solutions checked bit for bit; solves/s both ways.
--mode summarize: no GPU; reads rocprofv3 CSV output (--stats-dir, --fetch-dir, --write-dir) of a
  `--mode kernels` run and prints per-kernel durations and HBM bytes per launch. FETCH_SIZE and WRITE_SIZE
  (KiB per dispatch) are each scaled by the factor that makes the largest device copy of the run read and
  write exactly its 256 MiB; it also prints the multigrid outer iteration kernel by kernel (launches, algorithmic
  bytes, durations, PMC bytes); --traffic-json writes the K*u and the multigrid iteration entries that --mode
  solve reports as ku_roofline.traffic and roofline.traffic.
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
METRIC = "PCG solves/s @350k DOF, mixed vs FP64 at compliance <=1e-8 rel; p99 job latency"
WORKLOAD = "holed-box-80x40x40"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=("solve", "kernels", "serve", "cases", "sweep", "precision", "packed",
                                       "summarize"), default="solve")
    ap.add_argument("--pack", type=int, default=16, help="packed mode: systems solved together (<= 16)")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--matvecs", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true", help="solve mode: skip the serving-stream latency leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--jobs", type=int, default=96, help="serve mode: jobs in the stream")
    ap.add_argument("--workers", type=int, default=16, help="serve mode: resident jobs (and packed solves)")
    ap.add_argument("--simp-iterations", type=int, default=40, help="serve mode: SIMP iterations per job")
    ap.add_argument("--max-elements", type=int, default=250_000, help="serve mode: largest job")
    ap.add_argument("--cases-iterations", type=int, default=20, help="cases mode: SIMP iterations")
    ap.add_argument("--serve-precision", choices=("gmg", "fp64"), default="gmg",
                    help="serve mode: solver of the job server (the one-at-a-time baseline is always plain FP64)")
    ap.add_argument("--packing", choices=("on", "off"), default="on",
                    help="serve: the server packs its jobs' multigrid solves into shared launches (on) or runs each "
                         "job's solves on its worker's stream (off)")
    ap.add_argument("--no-alone", action="store_true", help="serve mode: skip the one-at-a-time legs")
    ap.add_argument("--stats-dir", help="summarize: rocprofv3 --kernel-trace --stats output directory")
    ap.add_argument("--fetch-dir", help="summarize: rocprofv3 --pmc FETCH_SIZE output directory")
    ap.add_argument("--write-dir", help="summarize: rocprofv3 --pmc WRITE_SIZE output directory")
    ap.add_argument("--traffic-json", help="summarize: write K*u HBM bytes per launch to this file")
    return ap.parse_args()


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 PMC counters (profiles/traffic.json), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as fh:
            return json.load(fh).get(f"{kernel}@{workload}")
    except (OSError, ValueError):
        return None


def cpu_model() -> str:
    """The host CPU's model name (lscpu's "Model name", read from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, gpu_iters: int, budget_s: float):
    """The FP64 oracle (reference-algorithm CPU restatement: 8-colour K*u in 256-element chunks, 8,192-DOF
    PCG reduction blocks) on this host's cores: K*u and a bounded run of PCG iterations of the same
    problem, scaled to solves/s with the iteration count of the same solve."""
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
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": f"{n_mv} K*u + {res.iterations} Jacobi-PCG iterations of the same problem on the FP64 oracle "
                  f"({threads} threads); solves/s = 1 / (s per iteration x {gpu_iters} iterations of the FP64 "
                  f"GPU solve)",
        "ms_per_matvec": mv_ms,
        "ms_per_pcg_iteration": per_it * 1e3,
    }


def setup(seed):
    import numpy as np
    import torch

    from mycel_amd import FemOperator
    from mycel_amd import workloads as W

    dev = torch.device("cuda:0")
    p = W.matvec_pcg_workstation(seed=seed)
    fm = p.fe_model()
    op = FemOperator(fm)
    f = fm.AssembleForces()
    b = torch.tensor(np.where(fm.ConstrainedDofs, 0.0, f), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    return p, fm, op, f, b


def time_matvecs(op, n, fp32=False):
    """Average K*u launch time from HIP events recorded on the operator's stream."""
    import torch

    dofs = op.Model.TotalDofs
    dt = torch.float32 if fp32 else torch.float64
    x = torch.randn(dofs, dtype=dt, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    run = op.apply_k_device_f32 if fp32 else op.apply_k_device
    stream = torch.cuda.ExternalStream(op.stream())
    for _ in range(3):
        run(x, y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(n):
        run(x, y)
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / n


def time_copy(nbytes, n=100):
    """Average time of a plain device-to-device copy moving nbytes in total (nbytes / 2 read, nbytes / 2
    written), HIP events over n back-to-back launches: what any kernel moving these bytes in one launch can
    expect at this size."""
    import torch

    x = torch.randn(max(1, int(nbytes) // 16), dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(n):
        y.copy_(x)
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / n


def timed_solves(op, b, steps, warmup, tol, precision, rho=None, penal=3.0):
    """A step is UpdateDensities(rho) (when given: the moduli, the Jacobi diagonal and, for the multigrid
    solver, the whole hierarchy are rebuilt) followed by one cold solve."""
    import torch

    from mycel_amd import PcgSolver

    x = torch.zeros_like(b)
    solver = PcgSolver()
    res = None
    for _ in range(warmup):
        x.zero_()
        torch.cuda.synchronize()
        if rho is not None:
            op.update_densities_device(rho, penal)
        res = solver.solve_device(op, b, x, tol=tol, maxIterations=500_000, precision=precision)
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(op.stream())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    iters = 0
    for _ in range(steps):
        x.zero_()
        if rho is not None:
            op.update_densities_device(rho, penal)
        res = solver.solve_device(op, b, x, tol=tol, maxIterations=500_000, precision=precision)
        iters += res.Iterations
    e1.record(stream)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if not res.Converged:
        raise SystemExit(f"PCG did not converge: {res.Message}")
    timed_solves.last_event_ms = e0.elapsed_time(e1)
    timed_solves.last_iterations = iters
    return el, res, x


def gmg_levels(nx, ny, nz, coarse_max_dofs=140):
    """(nodes, elements) of each multigrid level: 2x coarsening per axis (ceil) down to <= 140 DOFs."""
    levels = []
    while True:
        nodes = (nx + 1) * (ny + 1) * (nz + 1)
        levels.append((nodes, nx * ny * nz))
        if 3 * nodes <= coarse_max_dofs or min(nx, ny, nz) < 2:
            return levels
        nx, ny, nz = (nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2


def mg_iteration_model(nx, ny, nz, active_nodes, active_elements, tail_nodes=512):
    """Launches and algorithmic HBM bytes of one multigrid outer iteration, per kernel: [(kernel name prefix,
    launches per iteration, bytes per iteration)]. Every stream a launch reads or writes counts once (FP64 8 B,
    FP32 4 B per DOF; moduli 8 B (FP64) or 4 B (FP32) per element; the nodal block inverse B^-1, 6 FP32 values
    per node, 8 B per DOF; the constraint mask D^-1 4 B per DOF). The finest level counts its active DOFs D0 and
    elements E0, the coarse levels their box grids. Per launch:
      k_fpcg_update_xr_gmg   read x, p, r, Ap, D^-1, B^-1; write x, r (FP64) and the V-cycle's b, x0, d0 (FP32):
                             72 D0
      k_gmg_cheb_ec (x4)     read x, b, d, B^-1; write x', d; the moduli: 28 D0 + 4 E0
      k_gmg_residual_ec      read x, b, D^-1; write r: 16 D0 + 4 E0
      k_gmg_restrict_dof     read r; read the coarse D^-1 and B^-1 and write its b, x, d: 4 D0 + 24 D1
      k_gmg_cheb_s           levels 1 .. t0-1, the second pre-smoothing step, as the step above: 28 D + 4 E
      k_gmg_cheb2_s          levels 1 .. t0-1, both post-smoothing steps: read x, b, d, B^-1, write x'; the
                             moduli: 24 D + 4 E; below level t0-1 also the prolongation's D^-1 and coarse x:
                             + 4 D_l + 4 D_l+1
      k_gmg_residual_restrict read x, b, D^-1 of level l and write level l+1's entry: 12 D_l + 4 E_l + 24 D_l+1
      k_gmg_tail             stage level t0's b, x, d and every tail level's moduli, D^-1 and B^-1, the coarsest
                             inverse (4 n^2 B), prolong into level t0-1 (read D^-1, read and write x): 12 D_t0 +
                             sum(12 D + 4 E) + 4 n^2 + 12 D_t0-1
      k_gmg_prolong          the finest level (when t0 > 1): read the coarse x, read D^-1 and x, write x:
                             12 D_0 + 4 D_1
      k_gmg_cheb_final_ec    read x, b, d, B^-1 (FP32), r and Ap (FP64); write z (FP64); moduli: 44 D0 + 4 E0
      k_pcg_step4_ap<double, true  read z, p; write p', Ap' (FP64); the FP64 moduli: 32 D0 + 8 E0
    Levels below t0 (the first level from which every level has <= 512 nodes) run inside the tail."""
    lv = gmg_levels(nx, ny, nz)
    D = [3 * n for n, _ in lv]
    E = [e for _, e in lv]
    D[0], E[0] = 3 * active_nodes, active_elements
    top = len(lv) - 1
    t0 = top
    while t0 > 1 and lv[t0 - 1][0] <= tail_nodes:
        t0 -= 1
    nc = D[top]
    rows = [("k_fpcg_update_xr_gmg", 1, 72 * D[0]),
            ("k_gmg_cheb_ec<", 4, 4 * (28 * D[0] + 4 * E[0])),
            ("k_gmg_residual_ec<", 1, 16 * D[0] + 4 * E[0]),
            ("k_gmg_restrict_dof", 1, 4 * D[0] + 24 * D[1])]
    mid = range(1, t0)
    rows.append(("k_gmg_cheb_s", len(mid), sum(28 * D[l] + 4 * E[l] for l in mid)))
    rows.append(("k_gmg_cheb2_s", len(mid), sum(24 * D[l] + 4 * E[l] + (4 * D[l] + 4 * D[l + 1] if l < t0 - 1 else 0)
                                                for l in mid)))
    rows.append(("k_gmg_residual_restrict", len(mid), sum(12 * D[l] + 4 * E[l] + 24 * D[l + 1] for l in mid)))
    rows.append(("k_gmg_tail", 1, 12 * D[t0] + sum(12 * D[l] + 4 * E[l] for l in range(t0, top + 1)) + 4 * nc * nc
                 + 12 * D[t0 - 1]))
    rows.append(("k_gmg_prolong", 1 if t0 > 1 else 0, 12 * D[0] + 4 * D[1] if t0 > 1 else 0))
    rows.append(("k_gmg_cheb_final_ec<", 1, 44 * D[0] + 4 * E[0]))
    rows.append(("k_pcg_step4_ap<double, true", 1, 32 * D[0] + 8 * E[0]))
    return rows


def mg_iteration_bytes(nx, ny, nz, active_nodes, active_elements):
    return sum(b for _, _, b in mg_iteration_model(nx, ny, nz, active_nodes, active_elements))


def mode_solve(args):
    import numpy as np
    import torch

    import mycel_amd as M

    p, fm, op, f, b = setup(args.seed)
    info = op.info()
    mv_ms = time_matvecs(op, args.matvecs)
    mv32_ms = time_matvecs(op, args.matvecs, fp32=True)
    copy_ms = time_copy(2 * 3 * 8 * op.info().active_nodes + 8 * op.info().active_elements, args.matvecs)
    rho = torch.tensor(p.density_box(), dtype=torch.float64, device="cuda:0")
    el64, r64, x64 = timed_solves(op, b, args.steps, args.warmup, args.tol, M.PRECISION_FP64, rho, p.penal)
    elmx, rmx, xmx = timed_solves(op, b, args.steps, args.warmup, args.tol, M.PRECISION_MIXED_FP32, rho, p.penal)
    elbf, rbf, xbf = timed_solves(op, b, args.steps, args.warmup, args.tol, M.PRECISION_MIXED_BF16, rho, p.penal)
    elor, ror, xor_ = timed_solves(op, b, 1, 0, args.tol, M.PRECISION_FP64_ORDERED, rho, p.penal)
    elmg, rmg, xmg = timed_solves(op, b, args.steps, args.warmup, args.tol, M.PRECISION_MIXED_GMG, rho, p.penal)
    mg_event_ms, mg_iters = timed_solves.last_event_ms, timed_solves.last_iterations
    # The same step (UpdateDensities + multigrid solve) through the host-pointer entry points (densities, b and
    # x in host memory: PCIe copies in and out included); reported beside the device-resident value, never
    # as it.
    bh = b.cpu().numpy()
    xh = np.zeros_like(bh)
    M.PcgSolver().Solve(op, bh, xh, tol=args.tol, maxIterations=500_000, precision=M.PRECISION_MIXED_GMG)
    t0 = time.perf_counter()
    rho_h = p.density_box()
    for _ in range(args.steps):
        xh[:] = 0.0
        op.UpdateDensities(rho_h, p.penal)
        M.PcgSolver().Solve(op, bh, xh, tol=args.tol, maxIterations=500_000, precision=M.PRECISION_MIXED_GMG)
    host_ms = (time.perf_counter() - t0) / args.steps * 1e3
    # Accuracy against a tight FP64 reference of the same problem (untimed).
    xr = torch.zeros_like(b)
    M.PcgSolver().solve_device(op, b, xr, tol=1e-12, maxIterations=1_000_000)
    ft = torch.tensor(f, dtype=torch.float64, device="cuda:0")
    c_ref = float(ft @ xr)
    c64, cmx, cmg = float(ft @ x64), float(ft @ xmx), float(ft @ xmg)
    u_err = float(torch.linalg.norm(xmx - xr) / torch.linalg.norm(xr))
    u_err_mg = float(torch.linalg.norm(xmg - xr) / torch.linalg.norm(xr))
    solves64, solvesmx, solvesmg = args.steps / el64, args.steps / elmx, args.steps / elmg
    na, nn = info.active_elements, info.active_nodes
    bytes_mv = 2 * 3 * 8 * nn + 8 * na
    flops_mv = 1152 * na
    achieved = bytes_mv / (mv_ms * 1e-3) / 1e9
    # The step's unit of work: one multigrid outer iteration (HIP events on the operator's stream around the timed
    # solves; the step's UpdateDensities and hierarchy refresh are included in the time).
    it_ms = mg_event_ms / max(1, mg_iters)
    it_bytes = mg_iteration_bytes(p.nx, p.ny, p.nz, nn, na)
    it_gbs = it_bytes / (it_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": solvesmg,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elmg / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD,
            "grid": [p.nx, p.ny, p.nz],
            "active_elements": na,
            "active_dofs": 3 * nn,
            "box_dofs": p.n_dofs,
            "density": f"iid U[0.01,1] (SplitMix64 seed {args.seed})",
            "emin_ratio": p.emin_ratio,
            "tol": args.tol,
            "precision": "mixed: FP64 flexible PCG, FP32 multigrid V-cycle preconditioner",
            "parallelism": "single-gpu",
        },
        "mixed": {"solves_per_s": solvesmg, "pcg_iterations": rmg.Iterations,
                  "fp64_fallback_iterations": rmg.FallbackIterations, "ms_per_solve": elmg / args.steps * 1e3,
                  "ms_per_pcg_iteration": elmg / args.steps * 1e3 / max(1, rmg.Iterations),
                  "compliance_rel_err_vs_fp64_1e-12": abs(cmg - c_ref) / abs(c_ref), "u_rel_err": u_err_mg},
        "mixed_fp32_reliable_updates": {
            "solves_per_s": solvesmx, "pcg_iterations": rmx.Iterations, "reliable_updates": rmx.OuterCorrections,
            "ms_per_solve": elmx / args.steps * 1e3,
            "compliance_rel_err_vs_fp64_1e-12": abs(cmx - c_ref) / abs(c_ref), "u_rel_err": u_err},
        "fp64_reference_order": {"solves_per_s": 1.0 / elor, "pcg_iterations": ror.Iterations, "ms_per_solve": elor * 1e3,
                                 "compliance_rel_err_vs_fp64_1e-12": abs(float(ft @ xor_) - c_ref) / abs(c_ref)},
        "mixed_bf16_search_directions": {
            "solves_per_s": args.steps / elbf, "pcg_iterations": rbf.Iterations, "reliable_updates": rbf.OuterCorrections,
            "fp64_fallback_iterations": rbf.FallbackIterations, "ms_per_solve": elbf / args.steps * 1e3,
            "compliance_rel_err_vs_fp64_1e-12": abs(float(ft @ xbf) - c_ref) / abs(c_ref),
            "u_rel_err": float(torch.linalg.norm(xbf - xr) / torch.linalg.norm(xr))},
        "fp64": {"solves_per_s": solves64, "pcg_iterations": r64.Iterations, "ms_per_solve": el64 / args.steps * 1e3,
                 "ms_per_pcg_iteration": el64 / args.steps * 1e3 / r64.Iterations,
                 "compliance_rel_err_vs_fp64_1e-12": abs(c64 - c_ref) / abs(c_ref),
                 # One Jacobi-PCG iteration: the matvec plus 64 B per active DOF (read x, r, p, ap, 1/diag;
                 # write x, r, p).
                 "iteration_roofline": {
                     "algorithmic_bytes": bytes_mv + 64 * 3 * nn,
                     "achieved_gbs": (bytes_mv + 64 * 3 * nn) / (el64 / args.steps / r64.Iterations) / 1e9,
                     "hbm_frac": (bytes_mv + 64 * 3 * nn) / (el64 / args.steps / r64.Iterations) / 1e9
                     / HBM_PEAK_GBS}},
        "mixed_speedup_vs_fp64": solvesmg / solves64,
        "mixed_host_buffers_ms_per_solve": host_ms,
        "ms_per_matvec": {"fp64": mv_ms, "fp32": mv32_ms},
        "gdof_per_s_matvec": {"active": 3 * nn / (mv_ms * 1e-3) / 1e9, "box": p.n_dofs / (mv_ms * 1e-3) / 1e9},
        "roofline": {
            "kernel": "multigrid outer iteration: 20 launches (fused x/r update, FP32 V-cycle, fused p update + K p), bytes per mg_iteration_model",
            "bound": "hbm",
            "achieved": it_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": it_gbs / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles("mg_iteration", WORKLOAD),
            "algorithmic_bytes_per_launch": it_bytes,
            "avg_launch_ms": it_ms,
        },
        "ku_roofline": {
            "kernel": KU_KERNEL,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles("k_apply_k_ec<double>", WORKLOAD),
            "algorithmic_bytes_per_launch": bytes_mv,
            "avg_launch_ms": mv_ms,
            "fp64": {"achieved_tflops": flops_mv / (mv_ms * 1e-3) / 1e12, "peak_tflops": FP64_PEAK_TFLOPS,
                     "frac": flops_mv / (mv_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "flops_per_launch": flops_mv},
            # A plain device copy of the same bytes (half read, half written), one launch, HIP events: the
            # practical floor for one launch of this size.
            "copy_same_bytes": {"avg_launch_ms": copy_ms, "gbs": bytes_mv / (copy_ms * 1e-3) / 1e9,
                                "frac": bytes_mv / (copy_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        },
        "cpu_baseline": None,
    }
    op.close()
    if not args.no_serve:
        # The metric's job-latency part: the serving stream through the job server (capacity at saturation,
        # p50/p99 at 80% of it); --mode serve adds the one-at-a-time legs.
        out["serving"] = serving(args, one_at_a_time=False)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, r64.Iterations, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def mode_kernels(args):
    """Fixed short run for profilers: 20 FP64 K*u, 20 FP32 K*u, one cold solve in each solver mode (FP64,
    FP32 with reliable updates, multigrid), and a 256 MiB device-to-device copy (calibrates the byte
    counters)."""
    import torch

    import mycel_amd as M

    p, fm, op, f, b = setup(args.seed)
    print(f"[kernels] fp64 K*u {time_matvecs(op, 20):.4f} ms", flush=True)
    print(f"[kernels] fp32 K*u {time_matvecs(op, 20, fp32=True):.4f} ms", flush=True)
    src = torch.ones(32 << 20, dtype=torch.float64, device="cuda:0")  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("[kernels] copy 256 MiB x3 done", flush=True)
    for prec in (M.PRECISION_FP64, M.PRECISION_MIXED_FP32, M.PRECISION_MIXED_GMG):
        el, r, _ = timed_solves(op, b, 1, 0, args.tol, prec)
        print(f"[kernels] solve precision={prec} {el * 1e3:.2f} ms, {r.Iterations} it", flush=True)
    op.close()


def mode_precision(args):
    """The precision configuration: the 60x20x20 cantilever with a spatially correlated bimodal density field
    (30% rho = 1, 70% rho = 1e-3: ~1e-9 stiffness contrast), one cold solve to 1e-8 per step in each solver: plain
    FP64 Jacobi-PCG, FP32 Jacobi-PCG with FP64 reliable updates, FP64 flexible PCG with the FP32 multigrid V-cycle,
    and the same with island deflation; errors against an FP64 solve to 1e-12 (untimed)."""
    import numpy as np
    import torch

    import mycel_amd as M
    from mycel_amd import FemOperator
    from mycel_amd import workloads as W

    p = W.precision_variant()
    fm = p.fe_model()
    op = FemOperator(fm)
    f = fm.AssembleForces()
    dev = torch.device("cuda:0")
    b = torch.tensor(np.where(fm.ConstrainedDofs, 0.0, f), dtype=torch.float64, device=dev)
    ft = torch.tensor(f, dtype=torch.float64, device=dev)
    xr = torch.zeros_like(b)
    M.PcgSolver().solve_device(op, b, xr, tol=1e-12, maxIterations=2_000_000)
    c_ref = float(ft @ xr)
    legs = {}
    for name, prec, islands in (("fp64", M.PRECISION_FP64, 0.0), ("mixed_fp32_reliable_updates", M.PRECISION_MIXED_FP32, 0.0),
                                ("mixed_bf16_search_directions", M.PRECISION_MIXED_BF16, 0.0),
                                ("gmg", M.PRECISION_MIXED_GMG, 0.0), ("gmg_islands", M.PRECISION_MIXED_GMG, 1e-3)):
        op.SetIslandDeflation(islands)
        el, r, x = timed_solves(op, b, args.steps, args.warmup, args.tol, prec)
        legs[name] = {"solves_per_s": args.steps / el, "ms_per_solve": el / args.steps * 1e3,
                      "pcg_iterations": r.Iterations, "fp64_fallback_iterations": r.FallbackIterations,
                      "reliable_updates": r.OuterCorrections,
                      "compliance_rel_err_vs_fp64_1e-12": abs(float(ft @ x) - c_ref) / abs(c_ref),
                      "u_rel_err": float(torch.linalg.norm(x - xr) / torch.linalg.norm(xr))}
        print(f"[precision] {name}: {legs[name]}", file=sys.stderr, flush=True)
    op.close()
    best = legs["gmg_islands"]
    out = {
        "metric": METRIC, "value": best["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_solve"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "bimodal-cantilever-60x20x20", "grid": [p.nx, p.ny, p.nz], "box_dofs": p.n_dofs,
                   "density": "seeded N(0,1) noise, Gaussian blur sigma 2, top 30% rho = 1, else 1e-3",
                   "emin_ratio": p.emin_ratio, "tol": args.tol,
                   "precision": "mixed: FP64 flexible PCG, FP32 V-cycle + FP64 island deflation"},
        "legs": legs,
        "mixed_speedup_vs_fp64": best["solves_per_s"] / legs["fp64"]["solves_per_s"],
    }
    print(json.dumps(out), flush=True)


SWEEP_GRIDS = ((80, 40, 40, True), (100, 50, 50, False), (160, 80, 80, False), (200, 100, 100, False),
               (256, 128, 128, False))


def mode_sweep(args):
    """K*u roofline fractions against grid size. Algorithmic bytes per launch: 2 * 3 * s per active node
    (x read, y written) + s per active element (the modulus), s = 8 (FP64) or 4 (FP32); flops 1,152 per
    active element."""
    import numpy as np

    import mycel_amd as M
    from mycel_amd import FemOperator
    from mycel_amd import workloads as W

    rows = []
    for nx, ny, nz, holed in SWEEP_GRIDS:
        if holed:
            p = W.matvec_pcg_workstation(seed=args.seed)
        else:
            p = W.Problem(nx, ny, nz, name=f"box {nx}x{ny}x{nz}")
            W._clamped_x0(p)
            rng = np.random.default_rng(args.seed)
            p.density = 0.01 + 0.99 * rng.random(p.n_elem)
        act = p.active
        nodes = np.zeros((nz + 1, ny + 1, nx + 1), bool)
        ex, ey, ez = act % nx, (act // nx) % ny, act // (nx * ny)
        for oz in (0, 1):
            for oy in (0, 1):
                for ox in (0, 1):
                    nodes[ez + oz, ey + oy, ex + ox] = True
        na, nn = act.size, int(nodes.sum())
        op = FemOperator(p.fe_model())
        row = {"grid": [nx, ny, nz], "holed": holed, "active_elements": int(na), "active_nodes": nn,
               "box_dofs": 3 * (nx + 1) * (ny + 1) * (nz + 1)}
        cols = {}
        # K*u forms: the node gather (one node per thread, or kz-node z-marching columns) and the
        # element-centric product with kz node layers per tile.
        for form, tag, kzs in ((M.MATVEC_NODE_GATHER, "gather", (0, 2, 4, 8)), (M.MATVEC_ELEMENT, "element", (2, 3, 4, 8))):
            for kz in kzs:
                op.SetMatvecForm(form, kz)
                cols[f"{tag}{kz}"] = {"fp64_us": time_matvecs(op, args.matvecs) * 1e3,
                                      "fp32_us": time_matvecs(op, args.matvecs, fp32=True) * 1e3}
        row["forms"] = cols
        op.SetMatvecForm(M.MATVEC_AUTO)
        for name, fp32, sz in (("fp64", False, 8), ("fp32", True, 4)):
            ms = time_matvecs(op, args.matvecs, fp32=fp32)
            byt = 2 * 3 * sz * nn + sz * na
            gbs = byt / (ms * 1e-3) / 1e9
            cms = time_copy(byt, args.matvecs)
            row[name] = {"us_per_launch": ms * 1e3, "algorithmic_bytes": byt, "achieved_gbs": gbs,
                         "hbm_frac": gbs / HBM_PEAK_GBS,
                         "tflops": 1152.0 * na / (ms * 1e-3) / 1e12,
                         "copy_same_bytes_us": cms * 1e3, "copy_hbm_frac": byt / (cms * 1e-3) / 1e9 / HBM_PEAK_GBS}
            if not fp32:
                row[name]["fp64_frac"] = row[name]["tflops"] / FP64_PEAK_TFLOPS
        op.close()
        print(f"[sweep] {nx}x{ny}x{nz}{' holed' if holed else ''}: fp64 {row['fp64']['us_per_launch']:.2f} us "
              f"(hbm {row['fp64']['hbm_frac']:.3f}, fp64 {row['fp64']['fp64_frac']:.3f}; copy of the same bytes "
              f"{row['fp64']['copy_same_bytes_us']:.2f} us, hbm {row['fp64']['copy_hbm_frac']:.3f}), fp32 "
              f"{row['fp32']['us_per_launch']:.2f} us (hbm {row['fp32']['hbm_frac']:.3f}); by form "
              + " ".join(f"{k}:{v['fp64_us']:.1f}/{v['fp32_us']:.1f}" for k, v in cols.items()), file=sys.stderr,
              flush=True)
        rows.append(row)
    print(json.dumps({"metric": "K*u roofline fraction vs grid size", "unit": "fraction", "dtype": "f64/f32",
                      "data": "synthetic", "launches": args.matvecs, "hbm_peak_gbs": HBM_PEAK_GBS,
                      "fp64_peak_tflops": FP64_PEAK_TFLOPS, "sizes": rows}), flush=True)


def serving(args, one_at_a_time=True):
    """The serving configuration: a stream of fresh full-box jobs (20k-250k elements, half FEA solves, half SIMP
    runs of --simp-iterations iterations). Legs: (1) every job alone, one at a time, in plain FP64 and in the
    multigrid solver (service times; skipped without one_at_a_time); (2) the job server with every job submitted
    at once (its capacity at saturation, jobs/s); (3) the server under Poisson arrivals at 80% of that measured
    capacity (p50/p99 latency), beside a replay of the same arrivals through one-at-a-time service with the
    service times of (1)."""
    import numpy as np

    from mycel_amd import OptimizerOptions, StaticAnalysis, StaticAnalysisOptions, TopologyOptimizer
    from mycel_amd import workloads as W
    from mycel_amd.serve import DONE, JobServer

    import mycel_amd as M

    served = {"fp64": M.PRECISION_FP64, "gmg": M.PRECISION_MIXED_GMG}[args.serve_precision]

    def options(j, precision=M.PRECISION_FP64):
        return OptimizerOptions(VolumeFraction=j.volume_fraction, FilterRadiusMm=j.rmin,
                                MaxIterations=j.simp_iterations, SolverToleranceInitial=1e-8,
                                SolverToleranceFinal=1e-8, SolverMaxIterations=500_000, ProtectBoundaryHinges=False,
                                Precision=precision)

    stream = W.serving_stream(args.jobs, seed=args.seed, max_el=args.max_elements,
                              simp_iterations=args.simp_iterations)
    models = [j.problem.fe_model() for j in stream]
    forces = [m.AssembleForces() for m in models]

    def alone(precision):
        service = []
        for j, m in zip(stream, models):
            t0 = time.perf_counter()
            if j.kind == "fea":
                StaticAnalysis.Run(m, StaticAnalysisOptions(Tolerance=1e-8, MaxIterations=500_000,
                                                            Precision=precision))
            else:
                TopologyOptimizer.Run(m, options(j, precision))
            service.append(time.perf_counter() - t0)
        return service

    def replay(arrivals, service):  # one at a time in arrival order: queueing behind earlier jobs + service
        t_free, lat = 0.0, []
        for a, sv in zip(arrivals, service):
            t_free = max(a, t_free) + sv
            lat.append(t_free - a)
        return lat, t_free

    pack_stats = {}
    longest = []

    def served_run(arrivals):
        ids, lat, run = [], [], []
        with JobServer(workers=args.workers) as sv:
            sv.set_packing(args.packing == "on")
            t0 = time.perf_counter()
            for j, m, f, a in zip(stream, models, forces, arrivals):
                now = time.perf_counter() - t0
                if a > now:
                    time.sleep(a - now)
                ids.append(sv.submit_fea(m, tol=1e-8, max_iter=500_000, precision=served, forces=f)
                           if j.kind == "fea" else sv.submit_simp(m, options(j, served), forces=f))
            for jid in ids:
                info = sv.wait(jid)
                if info.state != DONE:
                    raise SystemExit(f"job {jid} failed with status {info.status}")
                lat.append(info.latency_ms / 1e3)
                run.append(info.run_ms / 1e3)
            elapsed = time.perf_counter() - t0
            pack_stats.update(sv.pack_stats())
            longest[:] = [max(run)]
            return lat, elapsed

    pct = lambda v: {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99))}
    out = {"jobs": len(stream), "workers": args.workers, "elements": [20_000, args.max_elements],
           "simp_iterations": args.simp_iterations,
           "server_solver": {"fp64": "plain FP64 Jacobi-PCG",
                             "gmg": "FP64 flexible PCG + FP32 multigrid V-cycle"}[args.serve_precision]}
    if one_at_a_time:
        service64 = alone(M.PRECISION_FP64)
        print(f"[serve] one at a time, FP64: {sum(service64):.1f} s", file=sys.stderr, flush=True)
        servicemg = alone(M.PRECISION_MIXED_GMG)
        print(f"[serve] one at a time, multigrid: {sum(servicemg):.1f} s", file=sys.stderr, flush=True)
    _, burst = served_run(np.zeros(len(stream)))
    capacity = len(stream) / burst
    st = dict(pack_stats)
    out["saturation_longest_job_run_s"] = longest[0]
    out["packing"] = {"on": args.packing == "on", "batches": st.get("batches", 0),
                      "mean_solves_per_batch": st["member_batches"] / st["batches"] if st.get("batches") else 0.0,
                      "graph_batches": st.get("graph_batches", 0), "solves": st.get("solves", 0),
                      "broker_host_ms": {k: st.get(k, 0.0) for k in ("build_ms", "enqueue_ms", "wait_ms")}}
    print(f"[serve] server at saturation: {capacity:.2f} jobs/s", file=sys.stderr, flush=True)
    rate = 0.8 * capacity
    rng = W.SplitMix64(args.seed + 1)
    arrivals = np.cumsum(-np.log1p(-rng.uniform(len(stream))) / rate)
    lat, makespan = served_run(arrivals)
    print(f"[serve] server at 80% of capacity: p99 {np.percentile(lat, 99):.3f} s", file=sys.stderr, flush=True)
    out["capacity_jobs_per_s"] = capacity
    out["saturation_makespan_s"] = burst
    out["at_80pct_capacity"] = {"offered_rate_jobs_per_s": rate, "jobs_per_s": len(stream) / makespan,
                                "latency_s": pct(lat)}
    if one_at_a_time:
        lat64, mk64 = replay(arrivals, service64)
        latmg, mkmg = replay(arrivals, servicemg)
        out["one_at_a_time_fp64"] = {"capacity_jobs_per_s": len(stream) / sum(service64),
                                     "mean_service_s": float(np.mean(service64)),
                                     "replay_at_same_arrivals": {"latency_s": pct(lat64), "makespan_s": mk64}}
        out["one_at_a_time_multigrid"] = {"capacity_jobs_per_s": len(stream) / sum(servicemg),
                                          "mean_service_s": float(np.mean(servicemg)),
                                          "replay_at_same_arrivals": {"latency_s": pct(latmg), "makespan_s": mkmg}}
    return out


def mode_packed(args):
    """--pack systems of the serving mix (fresh full-box grids of 20k-250k elements, densities U[0.01, 1], one random
    face clamped, 1-4 patch loads): one cold multigrid solve to 1e-8 each, one at a time (mycel_pcg_solve_d), then
    all of them together in packed launches (mycel_pcg_solve_packed_d), --steps times each after a warm-up. The
    packed solutions must equal the solutions alone bit for bit. Reports solves/s both ways."""
    import numpy as np
    import torch

    import mycel_amd as M
    from mycel_amd import FemOperator, PcgSolver
    from mycel_amd import workloads as W

    dev = torch.device("cuda:0")
    stream = W.serving_stream(args.pack, seed=args.seed, max_el=args.max_elements, simp_iterations=1)
    probs = [j.problem for j in stream]
    ops, bs, xs = [], [], []
    for k, p in enumerate(probs):
        p.density = 0.01 + 0.99 * W.SplitMix64(args.seed + 17 * k).uniform(p.n_elem)
        fm = p.fe_model()
        ops.append(FemOperator(fm))
        bs.append(torch.tensor(np.where(fm.ConstrainedDofs, 0.0, fm.AssembleForces()), device=dev))
        xs.append(torch.zeros(p.n_dofs, dtype=torch.float64, device=dev))
    solver = PcgSolver()
    G = M.PRECISION_MIXED_GMG

    def alone():
        res = []
        for op, b, x in zip(ops, bs, xs):
            x.zero_()
            res.append(solver.solve_device(op, b, x, tol=args.tol, maxIterations=100000, precision=G))
        torch.cuda.synchronize()
        return res

    def packed():
        for x in xs:
            x.zero_()
        res = solver.solve_packed_device(ops, bs, xs, tol=args.tol, maxIterations=100000)
        torch.cuda.synchronize()
        return res

    for _ in range(max(1, args.warmup)):
        alone()
        packed()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ra = alone()
    t_alone = (time.perf_counter() - t0) / args.steps
    xa = [x.cpu().numpy().copy() for x in xs]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rp = packed()
    t_packed = (time.perf_counter() - t0) / args.steps
    same = all(np.array_equal(x.cpu().numpy(), y) for x, y in zip(xs, xa)) and \
        all(a.Iterations == b.Iterations and a.FinalResidual == b.FinalResidual for a, b in zip(ra, rp))
    if not same:
        raise SystemExit("packed solutions differ from the solutions alone")
    n = len(ops)
    out = {
        "metric": "multigrid solves/s of independent serving-size systems, packed vs one at a time",
        "value": n / t_packed, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_packed * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"packed-{n}-serving-systems", "elements": [int(p.n_elem) for p in probs],
                   "tol": args.tol},
        "one_at_a_time_solves_per_s": n / t_alone, "packed_speedup": t_alone / t_packed,
        "iterations": [r.Iterations for r in rp], "bit_identical": same,
    }
    for op in ops:
        op.close()
    print(json.dumps(out), flush=True)


def mode_serve(args):
    sv = serving(args, one_at_a_time=not args.no_alone)
    out = {
        "metric": METRIC,
        "value": sv["capacity_jobs_per_s"],
        "unit": "jobs/s",
        "n_gpus": 1,
        "steps": sv["jobs"],
        "warmup": 0,
        "ms_per_step": sv["saturation_makespan_s"] / sv["jobs"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.serve_precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"serving-stream-{sv['jobs']}-jobs", "workers": args.workers,
                   "server_solver": sv["server_solver"], "elements": sv["elements"],
                   "simp_iterations": sv["simp_iterations"],
                   "value": "server capacity: every job submitted at once, jobs / makespan"},
    }
    out.update({k: sv[k] for k in ("saturation_longest_job_run_s", "at_80pct_capacity", "packing",
                                   "one_at_a_time_fp64", "one_at_a_time_multigrid") if k in sv})
    print(json.dumps(out), flush=True)


def mode_cases(args):
    import numpy as np
    import torch

    import mycel_amd as M
    from mycel_amd import OptimizerOptions, TopologyOptimizer
    from mycel_amd import workloads as W

    p, cases = W.multi_rhs(seed=99)
    f = W.case_forces(p, cases)
    n_it = args.cases_iterations

    def run(precision, iters, islands=0.0):
        o = OptimizerOptions(VolumeFraction=0.25, FilterRadiusMm=2.0, Penal=3.0, MoveLimit=0.2, MaxIterations=iters,
                             SolverToleranceInitial=1e-8, SolverToleranceFinal=1e-8, SolverMaxIterations=500_000,
                             ProtectBoundaryHinges=False, Precision=precision, IslandDeflation=islands)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = TopologyOptimizer.RunCases(p.fe_model(), f, o)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, r

    run(M.PRECISION_MIXED_GMG, 1)  # warm-up: module load, graph capture
    el, r = run(M.PRECISION_MIXED_GMG, n_it)
    eli, ri = run(M.PRECISION_MIXED_GMG, n_it, 1e-3)
    el64, r64 = run(M.PRECISION_FP64, 1)
    snaps = r.Snapshots
    out = {
        "metric": "SIMP iterations/s, 16 load cases on 100x50x50 (summed compliance)",
        "value": len(snaps) / el,
        "unit": "simp_iterations/s",
        "n_gpus": 1,
        "steps": len(snaps),
        "warmup": 1,
        "ms_per_step": el / len(snaps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "multi-rhs-100x50x50-16-cases", "grid": [p.nx, p.ny, p.nz], "box_dofs": p.n_dofs,
                   "load_cases": len(cases), "volume_fraction": 0.25, "rmin": 2.0, "penal": 3.0, "tol": 1e-8,
                   "solver": "FP64 flexible PCG + FP32 multigrid V-cycle, warm-started per case"},
        "pcg_iterations_per_simp_iteration": [s.SolverIterations for s in snaps],
        "ms_per_solve": el * 1e3 / (len(snaps) * len(cases)),
        "compliance_history": [s.Compliance for s in snaps],
        "island_deflation": {"simp_iterations_per_s": len(ri.Snapshots) / eli,
                             "pcg_iterations_per_simp_iteration": [s.SolverIterations for s in ri.Snapshots],
                             "compliance_history": [s.Compliance for s in ri.Snapshots]},
        "fp64_jacobi_first_simp_iteration": {"s": el64, "pcg_iterations": r64.Snapshots[0].SolverIterations,
                                             "compliance": r64.Snapshots[0].Compliance},
        "first_iteration_compliance_rel_diff_vs_fp64":
            abs(snaps[0].Compliance - r64.Snapshots[0].Compliance) / abs(r64.Snapshots[0].Compliance),
    }
    print(json.dumps(out), flush=True)


COPY_BYTES = 256 << 20  # the device copy in --mode kernels
# The K*u kernel ApplyK runs on the holed box (the automatic form: element-centric, 3 node layers per tile).
KU_KERNEL = "k_apply_k_ec<double,kApplyK,3>"


def _csv_rows(directory, suffix):
    import csv
    import glob

    rows = []
    for f in sorted(glob.glob(os.path.join(directory, "**", "*" + suffix), recursive=True)):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    return rows


def _short(name: str) -> str:
    """Kernel name without namespaces and argument lists: k_apply_k<double, 1, true>."""
    name = name.replace("mycel::(anonymous namespace)::", "").replace("void ", "")
    depth, out = 0, []
    for ch in name:
        if ch == "(" and depth == 0:
            break
        depth += ch == "<"
        depth -= ch == ">"
        out.append(ch)
    return "".join(out).strip()


def _pmc_bytes(directory, counter):
    """Median KiB per dispatch for each kernel, and the scale that makes the largest copy exact."""
    import statistics
    from collections import defaultdict

    per = defaultdict(list)
    copy_kib = 0.0
    for r in _csv_rows(directory, "counter_collection.csv"):
        if r["Counter_Name"] != counter:
            continue
        v = float(r["Counter_Value"])
        if "copyBuffer" in r["Kernel_Name"]:
            copy_kib = max(copy_kib, v)
        else:
            per[_short(r["Kernel_Name"])].append(v)
    scale = (COPY_BYTES / 1024.0) / copy_kib if copy_kib > 0 else float("nan")
    return {k: (statistics.median(v), len(v), statistics.fmean(v)) for k, v in per.items()}, scale, copy_kib


def _mg_table(args, lines, fetch=None, write=None, fs=1.0, ws=1.0):
    """The multigrid outer iteration of the kernels-mode solve: per kernel, launches per iteration, algorithmic
    bytes (mg_iteration_model), the kernel trace's average duration and the PMC bytes (mean over dispatches)."""
    from mycel_amd import workloads as W

    p = W.matvec_pcg_workstation()
    fm = p.fe_model()
    na, nn = fm.ActiveElementCount, fm.ActiveNodeCount
    stats = {}
    for r in _csv_rows(args.stats_dir, "kernel_stats.csv"):
        stats[_short(r["Name"])] = (int(r["Calls"]), float(r["AverageNs"]) / 1e3)
    lines += ["", "Multigrid outer iteration (the solve's kernels; launches per iteration from the V-cycle's structure):",
              "", "kernel | launches/iteration | algorithmic bytes/launch | avg us (trace) | algorithmic GB/s | "
              "PMC bytes/launch | PMC / algorithmic", "--- | ---: | ---: | ---: | ---: | ---: | ---:"]
    tot_us = tot_b = tot_pmc = 0.0
    pmc_ok = fetch is not None
    for prefix, n, b in mg_iteration_model(p.nx, p.ny, p.nz, nn, na):
        def match(k):  # the single-solve instantiations (HALT = false), or the flexible step 4 itself
            return k.startswith(prefix) and (prefix.endswith("true") or "true>" not in k[len(prefix):])
        hits = [(c, us) for k, (c, us) in stats.items() if match(k)]
        calls = sum(c for c, _ in hits)
        us = sum(c * u for c, u in hits) / calls if calls else float("nan")
        pmc = float("nan")
        if pmc_ok:
            # Each counter over every instantiation of the kernel in its own run: the autotuner may pick another
            # tile depth in each profiled run (the depths give the same bits and move the same bytes).
            fk, wk = [k for k in fetch if match(k)], [k for k in write if match(k)]
            fc, wc = sum(fetch[k][1] for k in fk), sum(write[k][1] for k in wk)
            if fc and wc:
                pmc = 1024 * (fs * sum(fetch[k][2] * fetch[k][1] for k in fk) / fc +
                              ws * sum(write[k][2] * write[k][1] for k in wk) / wc)
        per = b / n
        lines.append(f"{prefix.rstrip('<')} | {n} | {per:,.0f} | {us:.2f} | {per / (us * 1e3):,.0f} | "
                     f"{pmc:,.0f} | {pmc / per:.2f}")
        tot_us += n * us
        tot_b += b
        tot_pmc += n * pmc
    lines.append(f"**iteration** | {sum(n for _, n, _ in mg_iteration_model(p.nx, p.ny, p.nz, nn, na))} | "
                 f"{tot_b:,.0f} | {tot_us:.1f} | {tot_b / (tot_us * 1e3):,.0f} | {tot_pmc:,.0f} | {tot_pmc / tot_b:.2f}")
    return tot_pmc if pmc_ok else None



def mode_summarize(args):
    lines = []
    if args.stats_dir:
        lines.append("kernel | calls | avg us | min us | max us | % of GPU time")
        lines.append("--- | ---: | ---: | ---: | ---: | ---:")
        for r in _csv_rows(args.stats_dir, "kernel_stats.csv")[:24]:
            lines.append(f"{_short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.3f} | "
                         f"{float(r['MinNs']) / 1e3:.3f} | {float(r['MaxNs']) / 1e3:.3f} | {float(r['Percentage']):.2f}")
    traffic = {}
    if args.fetch_dir and args.write_dir:
        fetch, fs, fcopy = _pmc_bytes(args.fetch_dir, "FETCH_SIZE")
        write, ws, wcopy = _pmc_bytes(args.write_dir, "WRITE_SIZE")
        lines.append("")
        lines.append(f"calibration: 256 MiB device copy read FETCH_SIZE {fcopy:.1f} KiB -> scale {fs:.4f}; "
                     f"WRITE_SIZE {wcopy:.1f} KiB -> scale {ws:.4f}")
        lines.append("")
        lines.append("kernel | dispatches | FETCH KiB (raw median) | WRITE KiB (raw median) | HBM bytes/launch (scaled)")
        lines.append("--- | ---: | ---: | ---: | ---:")
        for k in sorted(set(fetch) & set(write)):
            b = fetch[k][0] * 1024 * fs + write[k][0] * 1024 * ws
            lines.append(f"{k} | {fetch[k][1]} | {fetch[k][0]:.1f} | {write[k][0]:.1f} | {b:.0f}")
            if k.startswith("k_apply_k"):
                traffic[k] = (b, fetch[k][1])
    mg_pmc = None
    if args.stats_dir:
        if args.fetch_dir and args.write_dir:
            mg_pmc = _mg_table(args, lines, fetch, write, fs, ws)
        else:
            _mg_table(args, lines)
    print("\n".join(lines))
    if args.traffic_json and traffic:
        # Plain K*u launches (mode 0, no dot partials) of either form, keyed without the template tail: the
        # dispatch-weighted mean over the tile depths (the autotuner's trials and the depth it picked).
        acc = {}
        for k, (v, cnt) in traffic.items():
            name, targs = k.split("<", 1)
            parts = [a.strip() for a in targs.rstrip(">").split(",")]
            if len(parts) >= 3 and parts[1] == "0" and parts[2] == "false":
                a = acc.setdefault(f"{name}<{parts[0]}>@{WORKLOAD}", [0.0, 0])
                a[0] += v * cnt
                a[1] += cnt
        ku = {k: a[0] / a[1] for k, a in acc.items()}
        if mg_pmc:
            ku[f"mg_iteration@{WORKLOAD}"] = mg_pmc
        with open(args.traffic_json, "w") as fh:
            json.dump(ku, fh, indent=1, sort_keys=True)


def main():
    args = parse()
    if args.mode == "summarize":
        mode_summarize(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Independent jobs would shard across GPUs with no communication.", file=sys.stderr)
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    {"solve": mode_solve, "kernels": mode_kernels, "serve": mode_serve, "cases": mode_cases,
     "sweep": mode_sweep, "precision": mode_precision, "packed": mode_packed}[args.mode](args)


if __name__ == "__main__":
    main()
