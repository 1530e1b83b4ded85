"""The reference CLI's box-path caller over the GPU library: `python -m mycel_amd.cli analyze <model.json>` and
`python -m mycel_amd.cli optimize <model.json> [--out <dir>]` (src/Mycel.Cli/Program.cs:18-127).

LoadModel (Program.cs:129-337) is restated for JSON models that describe a box grid ("box": [nx, ny, nz]); models
that name an STL surface need the voxeliser, which is not part of this library, and are rejected. Geometry follows
the reference's float32 arithmetic (voxel size, node positions, box bounds); support and load boxes select nodes by
position, a load's nodes carry area voxelSize^2 each, and "preserve" boxes mark elements by their centres.

`optimize` writes what the reference CLI writes: metrics.csv with the 12-column header and .NET InvariantCulture
formats ("{Iteration},{Compliance:G8},{VolumeFraction:G8},{MaxChange:G8},{GreyLevel:G6},{SolverIterations},
{SolverResidual:G6},{SolveMs:F2},{SensitivityMs:F2},{FilterMs:F2},{UpdateMs:F2},{TotalMs:F2}", UTF-8 with a byte-order
mark, as StreamWriter(Encoding.UTF8) writes it) and density.bin, the final physical density as raw little-endian
float32 per box element (Program.cs:101-120). `--precision` (an addition) selects the solver:
fp64 (the reference's Jacobi-PCG arithmetic, the default), reference-order, or gmg.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from typing import List, Optional, Tuple

import numpy as np

from . import _lib as L
from .fem import FeMaterial, FeModel, GridSpec, LoadCase, StaticAnalysis, StaticAnalysisOptions, SurfacePatch
from .optimization import OptimizerOptions, TopologyOptimizer
from .persistence import fmt_exp, fmt_fixed, fmt_g

METRICS_HEADER = ("Iteration,Compliance,VolumeFraction,MaxChange,GreyLevel,SolverIterations,SolverResidual,SolveMs,"
                  "SensitivityMs,FilterMs,UpdateMs,TotalMs")
PRECISIONS = {"fp64": L.PRECISION_FP64, "reference-order": L.PRECISION_FP64_ORDERED, "gmg": L.PRECISION_MIXED_GMG}

_f32 = np.float32


def _nodes_in_box(nx: int, ny: int, nz: int, h: np.float32, b) -> List[int]:
    """NodesInBox (Program.cs:323-337): nodes whose float32 position (i h, j h, k h) lies in the closed box,
    ascending k, j, i."""
    x0, y0, z0, x1, y1, z1 = (_f32(v) for v in b[:6])
    i = np.arange(nx + 1, dtype=np.float32) * h
    j = np.arange(ny + 1, dtype=np.float32) * h
    k = np.arange(nz + 1, dtype=np.float32) * h
    ii = np.flatnonzero((i >= x0) & (i <= x1))
    jj = np.flatnonzero((j >= y0) & (j <= y1))
    kk = np.flatnonzero((k >= z0) & (k <= z1))
    return [int(a + (nx + 1) * (c + (ny + 1) * d)) for d in kk for c in jj for a in ii]


def load_model(json_path: str) -> Tuple[FeModel, OptimizerOptions]:
    """LoadModel (Program.cs:129-321), box path."""
    with open(json_path, "r", encoding="utf-8-sig") as fh:
        root = json.load(fh)
    h = _f32(1.0)
    if "resolution" in root:
        h = _f32(root["resolution"])
    elif "voxelSize" in root:
        h = _f32(root["voxelSize"])
    mat = root.get("material", {})
    e0, nu, emin_ratio = float(mat.get("e0", 3000.0)), float(mat.get("nu", 0.3)), float(mat.get("eminRatio", 1e-6))
    if root.get("stl"):
        raise L.ArgumentError(L.E_ARGUMENT, "STL models need the voxeliser; only box models are supported here")
    nx, ny, nz = (int(v) for v in root["box"][:3])
    grid = GridSpec(nx, ny, nz, float(h))
    domain = np.ones(nx * ny * nz, np.uint8)

    fixed: List[int] = []
    for sup in root.get("supports", []):
        fix = [sup.get("fixX", True), sup.get("fixY", True), sup.get("fixZ", True)]
        if "box" in sup:
            for n in _nodes_in_box(nx, ny, nz, h, sup["box"]):
                fixed.extend(3 * n + c for c in range(3) if fix[c])
    loads = []
    area = float(h * h)  # voxelSize * voxelSize in float32, widened (Program.cs:243-255)
    for ld in root.get("loads", []):
        force = np.array(ld["force"][:3], np.float32)
        if "box" in ld:
            nodes = _nodes_in_box(nx, ny, nz, h, ld["box"])
            loads.append(LoadCase(SurfacePatch(np.array(nodes, np.int32), np.full(len(nodes), area)), force))
    passive: Optional[np.ndarray] = None
    if "preserve" in root:
        passive = np.zeros(nx * ny * nz, np.uint8)
        cx = (np.arange(nx, dtype=np.float32) + _f32(0.5)) * h  # GridSpec.ElementCenter
        cy = (np.arange(ny, dtype=np.float32) + _f32(0.5)) * h
        cz = (np.arange(nz, dtype=np.float32) + _f32(0.5)) * h
        for pr in root["preserve"]:
            if "box" not in pr:
                continue
            b = [_f32(v) for v in pr["box"][:6]]
            ix = np.flatnonzero((cx >= b[0]) & (cx <= b[3]))
            iy = np.flatnonzero((cy >= b[1]) & (cy <= b[4]))
            iz = np.flatnonzero((cz >= b[2]) & (cz <= b[5]))
            for c in iz:
                for bb in iy:
                    passive[ix + nx * (bb + ny * c)] = 1
    model = FeModel(grid, domain, FeMaterial(e0, nu, emin_ratio), np.array(fixed, np.int32), loads)
    opt = root.get("options", {})
    options = OptimizerOptions(VolumeFraction=float(opt.get("volumeFraction", 0.4)),
                               FilterRadiusMm=float(opt.get("filterRadiusMm", 2.0)),
                               MaxIterations=int(opt.get("maxIterations", 100)),
                               MoveLimit=float(opt.get("moveLimit", 0.2)), Penal=float(opt.get("penal", 3.0)),
                               EminRatio=emin_ratio, PassiveMask=passive)
    return model, options


def metrics_row(s) -> str:
    """One metrics.csv row of the reference CLI (Program.cs:106-110)."""
    return ",".join([str(int(s.Iteration)), fmt_g(s.Compliance, 8), fmt_g(s.VolumeFraction, 8), fmt_g(s.MaxChange, 8),
                     fmt_g(s.GreyLevel, 6), str(int(s.SolverIterations)), fmt_g(s.SolverResidual, 6),
                     fmt_fixed(s.SolveMs, 2), fmt_fixed(s.SensitivityMs, 2), fmt_fixed(s.FilterMs, 2),
                     fmt_fixed(s.UpdateMs, 2), fmt_fixed(s.TotalMs, 2)])


def run_analyze(json_path: str, precision: int = L.PRECISION_FP64) -> int:
    """RunAnalyze (Program.cs:61-82)."""
    model, _ = load_model(json_path)
    print(f"Running static analysis on model ({model.ActiveElementCount:,} active elements, {model.TotalDofs:,} DOFs)...")
    t0 = time.perf_counter()
    r = StaticAnalysis.Run(model, StaticAnalysisOptions(Tolerance=1e-6, Precision=precision))
    ms = (time.perf_counter() - t0) * 1e3
    max_vm = max(0.0, float(np.max(r.VonMisesPerElement))) if r.VonMisesPerElement.size else 0.0
    print(f"Compliance:       {fmt_fixed(r.Compliance, 6)}")
    print(f"Max Displacement: {fmt_fixed(r.MaxDisplacement, 6)} mm")
    print(f"Max von Mises:    {fmt_fixed(max_vm, 6)} MPa")
    print(f"Solve Time:       {fmt_fixed(ms, 1)} ms ({r.SolverInfo.Iterations} iters, "
          f"res: {fmt_exp(r.SolverInfo.FinalResidual, 2)})")
    return 0


def run_optimize(json_path: str, out_dir: Optional[str] = None, precision: int = L.PRECISION_FP64) -> int:
    """RunOptimize (Program.cs:84-127): SIMP, then metrics.csv and density.bin in out_dir."""
    model, options = load_model(json_path)
    options.Precision = precision
    out_dir = out_dir or os.path.join("runs", os.path.splitext(os.path.basename(json_path))[0])
    os.makedirs(out_dir, exist_ok=True)
    print(f"Optimizing model ({model.ActiveElementCount:,} active elements, {model.TotalDofs:,} DOFs, "
          f"target vol: {fmt_fixed(100.0 * options.VolumeFraction, 0)} %)...")
    print("Iter | Compliance | Volume  | dRho   | Grey % | PCG Iters (Residual) | Time (ms)")
    print("-----+------------+---------+--------+--------+----------------------+----------")

    def progress(s):  # Program.cs:94-97
        print(f"{s.Iteration:4d} | {fmt_fixed(s.Compliance, 3):>10} | {fmt_fixed(s.VolumeFraction, 4):>7} | "
              f"{fmt_fixed(s.MaxChange, 4):>6} | {fmt_fixed(s.GreyLevel, 1):>5}% | {s.SolverIterations:4d} it "
              f"({fmt_exp(s.SolverResidual, 1):>7}) | {fmt_fixed(s.TotalMs, 1):>8}", flush=True)

    result = TopologyOptimizer.Run(model, options, progress=progress)
    with open(os.path.join(out_dir, "metrics.csv"), "w", encoding="utf-8-sig", newline="") as fh:
        fh.write(METRICS_HEADER + "\n")
        for s in result.Snapshots:
            fh.write(metrics_row(s) + "\n")
    np.asarray(result.FinalPhysicalDensity, "<f4").tofile(os.path.join(out_dir, "density.bin"))
    print("-----+------------+---------+--------+--------+----------------------+----------")
    print(f"Optimization complete in {result.TotalIterations} iterations.")
    print(f"Final Compliance: {fmt_fixed(result.FinalCompliance, 4)}, "
          f"Final Volume: {fmt_fixed(100.0 * result.FinalVolumeFraction, 2)} %")
    print(f"Artifacts saved to {out_dir}: metrics.csv, density.bin")
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m mycel_amd.cli", description="Mycel CLI on the MI355X library")
    ap.add_argument("command", choices=("analyze", "optimize"))
    ap.add_argument("model")
    ap.add_argument("--out", "-o", default=None)
    ap.add_argument("--precision", choices=tuple(PRECISIONS), default="fp64")
    a = ap.parse_args(argv)
    try:
        if a.command == "analyze":
            return run_analyze(a.model, PRECISIONS[a.precision])
        return run_optimize(a.model, a.out, PRECISIONS[a.precision])
    except Exception as e:  # Program.cs:48-52
        print(f"Error: {e}", file=sys.stderr)
        return 2


if __name__ == "__main__":
    sys.exit(main())
