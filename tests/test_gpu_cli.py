"""The CLI caller on the GPU, pinned by the SIMP cantilever configuration: `optimize` on a box JSON describing the
60x20x20 cantilever (the CLI keeps the reference defaults: hinge protection on, tolerance schedule 1e-4 -> 1e-6)
writes the 12-column metrics.csv and float32 density.bin; with the reference-order solver they carry exactly the
oracle's values for the same run (tests/golden/c1_simp_hinges.npz)."""
import json
import os

import numpy as np
import pytest

import mycel_amd as M
from mycel_amd import cli
from mycel_amd.persistence import fmt_g

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C1_JSON = {
    "box": [60, 20, 20], "voxelSize": 1.0, "material": {"e0": 1.0, "nu": 0.3, "eminRatio": 1e-9},
    "supports": [{"box": [-0.1, -0.1, -0.1, 0.1, 20.1, 20.1]}],
    "loads": [{"box": [59.9, -0.1, -0.1, 60.1, 20.1, 0.1], "force": [0, 0, -1]}],
    "options": {"volumeFraction": 0.3, "filterRadiusMm": 1.5, "maxIterations": 30, "penal": 3.0, "moveLimit": 0.2},
}


def test_cli_optimize_c1_writes_reference_artifacts(gpu, tmp_path):
    path = os.path.join(str(tmp_path), "c1.json")
    with open(path, "w") as fh:
        json.dump(C1_JSON, fh)
    out = os.path.join(str(tmp_path), "run")
    assert cli.run_optimize(path, out, M.PRECISION_FP64_ORDERED) == 0
    with np.load(os.path.join(GOLDEN, "c1_simp_hinges.npz")) as z:
        ref = {k: z[k] for k in z.files}
    raw = open(os.path.join(out, "metrics.csv"), "rb").read()
    assert raw.startswith(b"\xef\xbb\xbf")  # StreamWriter(Encoding.UTF8) writes a byte-order mark
    lines = raw[3:].decode("utf-8").splitlines()
    assert lines[0] == cli.METRICS_HEADER and len(lines) == 31
    rows = [ln.split(",") for ln in lines[1:]]
    assert all(len(r) == 12 for r in rows)
    assert [r[1] for r in rows] == [fmt_g(c, 8) for c in ref["compliance"]]
    assert [int(r[5]) for r in rows] == ref["solver_iterations"].tolist()
    dens = np.fromfile(os.path.join(out, "density.bin"), "<f4")
    assert dens.size == 60 * 20 * 20
    assert np.array_equal(dens, ref["rho_bar"].astype(np.float32))


def test_cli_analyze_example(gpu, tmp_path, capsys):
    path = os.path.join(str(tmp_path), "cantilever.json")
    obj = dict(C1_JSON, box=[30, 10, 4], material={"e0": 3000.0, "nu": 0.3, "eminRatio": 1e-6},
               supports=[{"box": [-0.1, -0.1, -0.1, 0.1, 10.1, 4.1]}],
               loads=[{"box": [29.9, -0.1, -0.1, 30.1, 10.1, 0.1], "force": [0, 0, -1000]}])
    with open(path, "w") as fh:
        json.dump(obj, fh)
    assert cli.main(["analyze", path, "--precision", "reference-order"]) == 0
    text = capsys.readouterr().out
    model, _ = cli.load_model(path)
    import oracle as O

    m = O.Model(30, 10, 4, None, model.FixedDofs, 1.0, 3000.0, 0.3, 1e-6, 3.0, None)
    sa = m.static_analysis(model.AssembleForces(), tol=1e-6)
    from mycel_amd.persistence import fmt_fixed

    assert f"Compliance:       {fmt_fixed(sa['compliance'], 6)}" in text
    assert f"Max Displacement: {fmt_fixed(sa['max_displacement'], 6)} mm" in text
