"""The CLI caller's box-path model loading and output formats (src/Mycel.Cli/Program.cs:101-337), on the CPU."""
import json
import os

import numpy as np

from mycel_amd import cli
from mycel_amd import workloads as W


def _write(tmp_path, obj):
    path = os.path.join(str(tmp_path), "model.json")
    with open(path, "w") as fh:
        json.dump(obj, fh)
    return path


EXAMPLE = {  # examples/cantilever.json of the reference
    "box": [30, 10, 4], "voxelSize": 1.0, "material": {"e0": 3000.0, "nu": 0.3, "eminRatio": 1e-6},
    "supports": [{"box": [-0.1, -0.1, -0.1, 0.1, 10.1, 4.1], "fixX": True, "fixY": True, "fixZ": True}],
    "loads": [{"box": [29.9, -0.1, -0.1, 30.1, 10.1, 0.1], "force": [0, 0, -1000]}],
    "options": {"volumeFraction": 0.4, "filterRadiusMm": 2.0, "maxIterations": 50, "penal": 3.0, "moveLimit": 0.2},
}


def test_load_model_box_example(tmp_path):
    model, opt = cli.load_model(_write(tmp_path, EXAMPLE))
    p = W.Problem(30, 10, 4)
    clamp = [p.node(0, j, k) for k in range(5) for j in range(11)]
    assert np.array_equal(model.FixedDofs, W.fix_all(clamp))
    (lc,) = model.Loads
    assert list(lc.Patch.Nodes) == [p.node(30, j, 0) for j in range(11)]
    assert np.all(lc.Patch.NodeArea == 1.0)
    f = model.AssembleForces()
    assert abs(f[2::3].sum() + 1000.0) < 1e-9 and np.count_nonzero(f) == 11
    assert (opt.VolumeFraction, opt.FilterRadiusMm, opt.MaxIterations, opt.MoveLimit, opt.Penal) == (0.4, 2.0, 50, 0.2, 3.0)
    assert opt.ProtectBoundaryHinges and opt.SolverToleranceInitial == 1e-4 and opt.SolverToleranceFinal == 1e-6
    assert model.Material.EminRatio == 1e-6 and model.Material.E0 == 3000.0


def test_load_model_float32_boxes_fix_flags_and_preserve(tmp_path):
    """Bounds compare against float32 node positions (a box edge at 0.3 with h = 0.1 selects by the float32
    products i * h); fix flags default to true; preserve boxes select elements by their centres."""
    obj = {"box": [10, 4, 3], "voxelSize": 0.1,
           "supports": [{"box": [0, 0, 0, 0.3, 0.4, 0.3], "fixZ": False}],
           "loads": [{"box": [1.0, 0, 0, 1.0, 0.4, 0.0], "force": [1, 2, 3]}],
           "preserve": [{"box": [0.0, 0.0, 0.0, 0.21, 0.4, 0.3]}]}
    model, opt = cli.load_model(_write(tmp_path, obj))
    h = np.float32(0.1)
    xs = np.flatnonzero(np.arange(11, dtype=np.float32) * h <= np.float32(0.3))
    assert xs.tolist() == [0, 1, 2, 3] or xs.tolist() == [0, 1, 2]
    fixed = set(model.FixedDofs.tolist())
    nodes = [i + 11 * (j + 5 * k) for k in range(4) for j in range(5) for i in xs]
    assert fixed == {3 * n + c for n in nodes for c in (0, 1)}
    assert np.allclose(model.Loads[0].Patch.NodeArea, float(h * h))
    pm = opt.PassiveMask.reshape(3, 4, 10)
    assert pm[:, :, :2].all() and not pm[:, :, 2:].any()  # centres 0.05, 0.15 <= 0.21 < 0.25


def test_metrics_row_formats():
    class S:
        Iteration, Compliance, VolumeFraction, MaxChange, GreyLevel = 7, 114757.48123, 0.30000000265, 0.012345678, 12.3456789
        SolverIterations, SolverResidual, SolveMs, SensitivityMs, FilterMs, UpdateMs, TotalMs = 123, 8.81e-5, 1.005, 0.04, 0.02, 0.19, 3.659
    row = cli.metrics_row(S)
    assert row == "7,114757.48,0.3,0.012345678,12.3457,123,8.81E-05,1.00,0.04,0.02,0.19,3.66"
