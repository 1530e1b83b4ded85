"""Run folders (mycel_amd.persistence) against the reference's formats: Persistence/RunWriter.cs,
Persistence/RunReader.cs, Persistence/RunModels.cs.

The byte layouts are pinned by the offsets the reference writes at (RunWriter.cs:233-246, :275-278,
:311-315), read back here with explicit offsets rather than the module's own struct formats; the checksum by
the published FNV-1a (32-bit) test vectors; the quantization by an independent numpy restatement of
MathF.Round(clamp(x, 0, 1) * 255f, AwayFromZero); the metrics row by .NET's InvariantCulture format
strings ("G17", "F6", "F4", "E4" with three exponent digits, "F2"). The reference does not build here (no
.NET), so no file written by it is available: compatibility rests on these pins.
"""
import json
import os

import numpy as np
import pytest

import mycel_amd.persistence as P
from mycel_amd.optimization import IterationSnapshot, OptimizerOptions


def _fnv1a(data: bytes) -> int:
    h = 2166136261
    for b in data:
        h = ((h ^ b) * 16777619) & 0xFFFFFFFF
    return h


def _q_restated(p):
    v = np.clip(np.asarray(p, np.float32), np.float32(0), np.float32(1)) * np.float32(255)
    return np.where(np.isnan(v), 0, np.floor(v.astype(np.float64) + 0.5)).astype(np.uint8)


def _i32(b, o):
    return int.from_bytes(b[o:o + 4], "little", signed=True)


def _f64(b, o):
    return float(np.frombuffer(b[o:o + 8], "<f8")[0])


def _snap(it, rng, n, **kw):
    s = IterationSnapshot(Iteration=it, Compliance=float(rng.uniform(10, 1e5)), VolumeFraction=float(rng.uniform()),
                          MaxChange=float(rng.uniform()), GreyLevel=float(rng.uniform(0, 100)),
                          SolverIterations=int(rng.integers(1, 5000)), SolverResidual=float(rng.uniform(1e-9, 1e-5)),
                          TotalMs=float(rng.uniform(1, 500)), Lambda=float(rng.uniform(0.1, 10)))
    s.PhysicalDensity = rng.uniform(-0.1, 1.1, n).astype(np.float32)
    s.DesignDensity = rng.uniform(0, 1, n).astype(np.float32)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_fnv1a_published_vectors():
    for data, h in ((b"", 0x811C9DC5), (b"a", 0xE40C292C), (b"foobar", 0xBF9CF968)):
        assert P.checksum_bytes(np.frombuffer(data, np.uint8)) == h
        assert _fnv1a(data) == h


def test_quantize_matches_restatement():
    rng = np.random.default_rng(3)
    p = np.concatenate([np.array([-1, 0, 0.5, 1, 2, np.nan, 1 / 255, 0.25, 254.5 / 255, -0.0], np.float32),
                        rng.uniform(-0.2, 1.2, 5000).astype(np.float32)])
    q, h = P.quantize(p)
    assert np.array_equal(q, _q_restated(p))
    assert list(q[:8]) == [0, 0, 128, 255, 255, 0, 1, 64]
    assert h == _fnv1a(q.tobytes()) == P.compute_quantized_checksum(p)


def test_brotli_round_trip_and_truncation():
    data = np.random.default_rng(0).integers(0, 7, 300_000, dtype=np.uint8).tobytes()
    blob = P.brotli_compress(data)
    assert len(blob) < len(data)
    assert P.brotli_decompress(blob) == data
    with pytest.raises(EOFError):
        P.brotli_decompress(blob[: len(blob) // 2])


def test_metrics_row_net_formats():
    row = P.format_metrics_row(7, 114757.48, 0.3, 0.012345678, 12.34567, 123, 1.5e-9, 12.5, 4096)
    assert row == "7,114757.48,0.300000,0.012346,12.3457,123,1.5000E-009,12.50,4096"
    assert P.fmt_g17(0.1) == "0.10000000000000001"  # 17 significant digits
    assert P.fmt_g17(1e-9) == "1.0000000000000001E-09"  # G: two exponent digits
    assert P.fmt_g17(1e20) == "1E+20"
    assert P.fmt_g17(7.0) == "7"
    assert P.fmt_exp(114757.48, 4) == "1.1476E+005"  # E: three exponent digits
    assert P.fmt_exp(0.0, 4) == "0.0000E+000"
    assert P.fmt_fixed(0.0078125, 6) == "0.007813"  # an exact binary midpoint rounds away from zero
    assert P.fmt_fixed(-0.0078125, 6) == "-0.007813"
    for x in (114757.48, 1 / 3, 2.0 ** -30, 6.02e23):
        assert float(P.fmt_g17(x)) == x  # G17 round-trips every double


def test_topo_header_offsets():
    rng = np.random.default_rng(1)
    nx, ny, nz = 5, 4, 3
    s = _snap(42, rng, nx * ny * nz)
    blob, h = P.encode_iteration(nx, ny, nz, s.Iteration, s.Compliance, s.VolumeFraction, s.MaxChange, s.GreyLevel,
                                 s.SolverIterations, s.TotalMs, s.PhysicalDensity)
    assert blob[:4] == b"TOPO"
    assert [_i32(blob, o) for o in (4, 8, 12, 16, 20)] == [1, nx, ny, nz, 42]
    assert [_f64(blob, o) for o in (24, 32, 40, 48)] == [s.Compliance, s.VolumeFraction, s.MaxChange, s.GreyLevel]
    assert _i32(blob, 56) == s.SolverIterations and _f64(blob, 60) == s.TotalMs
    assert int.from_bytes(blob[68:72], "little") == h == _fnv1a(_q_restated(s.PhysicalDensity).tobytes())
    assert P.brotli_decompress(blob[72:]) == _q_restated(s.PhysicalDensity).tobytes()


def _reference_checkpoint_read(blob):
    """RunReader.ReadLatestCheckpoint (RunReader.cs:228-250): 16-byte header, then exactly len float32 design
    and len float32 physical values from the decompressed stream; anything after them is never read."""
    assert blob[:4] == b"CHKP"
    it, n = _i32(blob, 8), _i32(blob, 12)
    raw = P.brotli_decompress(blob[16:])
    return it, np.frombuffer(raw[:4 * n], "<f4"), np.frombuffer(raw[4 * n:8 * n], "<f4")


def test_checkpoint_plain_and_extended():
    rng = np.random.default_rng(2)
    d = rng.uniform(0, 1, 600).astype(np.float32)
    p = rng.uniform(0, 1, 600).astype(np.float32)
    plain = P.encode_checkpoint(9, d, p)
    assert _i32(plain, 4) == 1
    it, rd, rp = _reference_checkpoint_read(plain)
    assert it == 9 and np.array_equal(rd, d) and np.array_equal(rp, p)
    g = rng.uniform(0, 1, 333)
    u = rng.standard_normal(1200)
    ext = P.encode_checkpoint(9, d, p, design64=g, displacements=u, lam=0.123456789012345)
    assert _i32(ext, 4) == 2
    it, rd, rp = _reference_checkpoint_read(ext)  # still readable by the reference
    assert it == 9 and np.array_equal(rd, d) and np.array_equal(rp, p)
    ck = P.decode_checkpoint(ext)
    assert np.array_equal(ck.DesignDensity64, g) and np.array_equal(ck.Displacements, u)
    assert ck.Lambda == 0.123456789012345
    u2 = rng.standard_normal((3, 50))
    ck2 = P.decode_checkpoint(P.encode_checkpoint(1, d, p, displacements=u2, lam=1.0))
    assert ck2.Displacements.shape == (3, 50) and np.array_equal(ck2.Displacements, u2)
    mism = P.decode_checkpoint(P.encode_checkpoint(1, d, p[:10]))  # RunWriter.cs:283-290
    assert np.array_equal(mism.PhysicalDensity, np.zeros(600, np.float32))


def test_estimate_disk_bytes():
    # RunWriter.cs:120-132 by hand: frames, max(512, 0.22 n + 72), max(512, 0.1 n), 2 n, 80 per iteration.
    assert P.RunWriter.EstimateDiskBytes(1000, 10) == 512 + 2000 + 10 * 512 + 800
    assert P.RunWriter.EstimateDiskBytes(1000, 10, 3) == 512 + 2000 + (1 + 3 + 1) * 512 + 800
    assert P.RunWriter.EstimateDiskBytes(100000, 4, 2) == 10000 + 200000 + 3 * 22072 + 320


def _writer(tmp_path, nx=6, ny=5, nz=4, keep=1, resume=False, domain=None, passive=None, name="run"):
    m = P.RunMetadata.for_grid(nx, ny, nz, h=0.5, options=OptimizerOptions(), keep_every_n=keep)
    n = nx * ny * nz
    dom = domain if domain is not None else (np.arange(n) % 7 != 0).astype(np.uint8)
    return P.RunWriter(str(tmp_path / name), m, dom, passive, is_resume=resume), dom


def test_writer_reader_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    w, dom = _writer(tmp_path, keep=2)
    n = 6 * 5 * 4
    pas = np.zeros(n, np.uint8)
    snaps = [_snap(i, rng, n) for i in range(1, 6)]
    roll = 2166136261
    for s in snaps:
        w.WriteSnapshot(s, isFinal=(s.Iteration == 5))
        roll = ((roll ^ _fnv1a(_q_restated(s.PhysicalDensity).tobytes())) * 16777619) & 0xFFFFFFFF
    assert w.WrittenSnapshotCount == 5 and w.RollingStreamChecksum == roll
    w.FinalizeRun(P.RunStatus.Completed)
    w.Dispose()

    r = P.RunReader.Open(str(tmp_path / "run"))
    assert r.Metadata.Status == P.RunStatus.Completed and r.Metadata.CompletedIterations == 5
    assert r.GetAvailableIterations() == [1, 2, 4, 5]  # 1, every 2nd, and the final one (RunWriter.cs:139)
    d, p = r.ReadDomainMasks()
    assert np.array_equal(d, dom) and np.array_equal(p, pas)
    for s in snaps:
        if s.Iteration not in (1, 2, 4, 5):
            continue
        hdr, dens = r.ReadIteration(s.Iteration)
        assert (hdr.Version, hdr.Nx, hdr.Ny, hdr.Nz, hdr.Iteration) == (1, 6, 5, 4, s.Iteration)
        assert (hdr.Compliance, hdr.VolumeFraction, hdr.MaxChange, hdr.GreyLevel, hdr.SolverIterations,
                hdr.IterationMs) == (s.Compliance, s.VolumeFraction, s.MaxChange, s.GreyLevel, s.SolverIterations,
                                     s.TotalMs)
        assert np.array_equal(dens, _q_restated(s.PhysicalDensity).astype(np.float32) * np.float32(1 / 255))
    assert r.ReadIteration(3)[0].Iteration == 2  # nearest committed, the earlier on a tie (RunReader.cs:374-398)
    ck = r.ReadLatestCheckpoint()
    assert ck.Iteration == 5 and np.array_equal(ck.DesignDensity, snaps[-1].DesignDensity)
    assert np.array_equal(ck.PhysicalDensity, snaps[-1].PhysicalDensity) and ck.Displacements is None
    rows = r.ReadMetrics()
    assert [x.Iteration for x in rows] == [1, 2, 3, 4, 5]
    for row, s in zip(rows, snaps):
        assert row.Compliance == s.Compliance  # G17 round-trips
        assert abs(row.VolumeFraction - s.VolumeFraction) <= 5e-7 and abs(row.SolverResidual - s.SolverResidual) <= 1e-4 * s.SolverResidual
        assert row.SolverIterations == s.SolverIterations and row.DiskBytes > 0
    raw = open(tmp_path / "run" / "metrics.csv", "rb").read()
    assert raw.startswith(b"\xef\xbb\xbfiteration,compliance,")  # StreamWriter(Encoding.UTF8) preamble
    meta = json.loads(open(tmp_path / "run" / "run.json", encoding="utf-8-sig").read())
    assert meta["Status"] == "Completed" and meta["Grid"]["Nx"] == 6 and meta["Options"]["KeepEveryN"] == 2
    assert meta["StartedAtUtc"].endswith("Z") and len(meta["StartedAtUtc"]) == 28


def test_checksum_mismatch_detected(tmp_path):
    rng = np.random.default_rng(6)
    w, _ = _writer(tmp_path)
    w.WriteSnapshot(_snap(1, rng, 120))
    w.Dispose()
    f = tmp_path / "run" / "iterations" / "it_000001.bin.br"
    b = bytearray(open(f, "rb").read())
    b[68] ^= 0xFF
    open(f, "wb").write(bytes(b))
    with pytest.raises(ValueError, match="Checksum mismatch"):
        P.RunReader.Open(str(tmp_path / "run")).ReadIteration(1)


def test_open_interrupted_and_resume(tmp_path):
    rng = np.random.default_rng(7)
    w, dom = _writer(tmp_path)
    for i in (1, 2, 3):
        w.WriteSnapshot(_snap(i, rng, 120))
    w.Dispose()
    run = tmp_path / "run"
    meta = json.loads(open(run / "run.json", encoding="utf-8-sig").read())
    meta["CompletedIterations"] = 1
    open(run / "run.json", "w", encoding="utf-8").write(json.dumps(meta))
    open(run / "iterations" / "it_000004.bin.br.tmp", "wb").write(b"partial")
    r = P.RunReader.Open(str(run))
    assert r.Metadata.Status == P.RunStatus.Interrupted  # Running with no owner (RunReader.cs:67-70)
    assert r.Metadata.CompletedIterations == 3 and r.GetAvailableIterations() == [1, 2, 3]
    assert P.RunReader.Open(str(run), is_live_active_run=True).Metadata.Status == P.RunStatus.Running
    # Resume: stray .tmp files go, the domain file and the metrics header are kept as they are.
    dom_before = open(run / "domain.bin.br", "rb").read()
    w2, _ = _writer(tmp_path, resume=True, domain=np.ones(120, np.uint8))
    assert not os.path.exists(run / "iterations" / "it_000004.bin.br.tmp")
    w2.WriteSnapshot(_snap(4, rng, 120))
    w2.Dispose()
    assert open(run / "domain.bin.br", "rb").read() == dom_before
    text = open(run / "metrics.csv", encoding="utf-8-sig").read()
    assert text.count("iteration,compliance") == 1 and len(P.RunReader.Open(str(run)).ReadMetrics()) == 4


def test_domain_header_offsets(tmp_path):
    n = 6 * 5 * 4
    pas = (np.arange(n) % 3 == 0).astype(np.uint8)
    w, dom = _writer(tmp_path, passive=pas)
    w.Dispose()
    b = open(tmp_path / "run" / "domain.bin.br", "rb").read()
    assert b[:4] == b"DOMN" and [_i32(b, o) for o in (4, 8, 12, 16)] == [1, 6, 5, 4]
    raw = P.brotli_decompress(b[20:])
    assert raw == dom.tobytes() + pas.tobytes()


def test_list_runs_newest_first_and_skips_malformed(tmp_path):
    rng = np.random.default_rng(8)
    for name in ("2026-01-01_10-00-00_a", "2026-01-02_10-00-00_b"):
        w, _ = _writer(tmp_path, name=name)
        w.WriteSnapshot(_snap(1, rng, 120))
        w.Dispose()
    os.makedirs(tmp_path / "2026-01-03_bad")
    open(tmp_path / "2026-01-03_bad" / "run.json", "w").write("{not json")
    runs = P.RunReader.ListRuns(str(tmp_path))
    assert [os.path.basename(d) for d, _ in runs] == ["2026-01-02_10-00-00_b", "2026-01-01_10-00-00_a"]
    assert all(m.Status == P.RunStatus.Interrupted for _, m in runs)


def test_resume_options_from_checkpoint():
    d = np.linspace(0, 1, 50).astype(np.float32)
    ck = P.decode_checkpoint(P.encode_checkpoint(12, d, d, design64=np.linspace(0, 1, 40), displacements=np.ones(30),
                                                 lam=0.5))
    o = P.resume_options(ck, OptimizerOptions(MaxIterations=40))
    assert o.StartIteration == 13 and o.MaxIterations == 40 and o.InitialLambda == 0.5
    assert np.array_equal(o.InitialDesign64, np.linspace(0, 1, 40)) and np.array_equal(o.WarmStartDisplacements, np.ones(30))
    o2 = P.resume_options(P.decode_checkpoint(P.encode_checkpoint(3, d, d)), OptimizerOptions())
    assert o2.StartIteration == 4 and np.array_equal(o2.InitialDensity, d) and o2.InitialDesign64 is None
