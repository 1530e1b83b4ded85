"""Run folders: the on-disk formats of a topology optimization run, so GPU runs interoperate with the App's
scrub and resume (Persistence/RunWriter.cs, Persistence/RunReader.cs, Persistence/RunModels.cs).

A run folder holds

- ``run.json``: the run metadata (RunModels.cs:54-71), rewritten atomically after every snapshot;
- ``metrics.csv``: one row per iteration, 9 columns (RunWriter.cs:86-91,170-185);
- ``domain.bin.br``: "DOMN" header + Brotli(domain mask, passive mask), one byte per box element
  (RunWriter.cs:297-334);
- ``iterations/it_NNNNNN.bin.br``: "TOPO" 72-byte header + Brotli(8-bit physical density), kept for
  iteration 1, every KeepEveryN-th iteration and the final one (RunWriter.cs:137-149,201-261);
- ``checkpoint_latest.bin.br``: "CHKP" 16-byte header + Brotli(float32 design, float32 physical density)
  (RunWriter.cs:263-295).

Every file is written to ``<name>.tmp`` and renamed into place, so a reader never sees a partial file.

Beyond the reference: the checkpoint may carry an extension after its two float32 arrays, an "FP64" block
with the FP64 active-compact design density, the FP64 displacements and the OC multiplier, so that a run
resumes from exactly the state it stopped in (TopologyOptimizer.Run with InitialDensity, StartIteration,
WarmStartDisplacements and InitialLambda). The reference's reader takes the first 2 * len * 4 decompressed
bytes and ignores the rest (RunReader.cs:236-245), so an extended checkpoint stays readable by it.

Brotli comes from the system's libbrotlienc / libbrotlidec (public one-shot and streaming API); the writer
uses quality 1, window 22, the setting of .NET's BrotliStream at CompressionLevel.Fastest. The 8-bit
quantization and the FNV-1a checksums run in libmycel_hip.so (mycel_run_quantize, mycel_run_checksum).
"""
from __future__ import annotations

import ctypes as C
import datetime as _dt
import json
import math
import os
import struct
from dataclasses import dataclass, field
from decimal import ROUND_HALF_UP, Decimal
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib as L

__all__ = ["RunStatus", "RunMetadata", "RunWriter", "RunReader", "IterationHeader", "IterationMetricRow",
           "Checkpoint", "brotli_compress", "brotli_decompress", "compute_quantized_checksum", "format_metrics_row",
           "resume_options"]

TOPO_HEADER = 72
CHKP_HEADER = 16
DOMN_HEADER = 20
FNV_OFFSET = 2166136261
FNV_PRIME = 16777619
_EXT_MAGIC = b"FP64"
_LOOP_MAGIC = b"SIMP"


# ---- Brotli ---------------------------------------------------------------------------------------

class _Brotli:
    def __init__(self):
        try:
            self.enc = C.CDLL("libbrotlienc.so.1")
            self.dec = C.CDLL("libbrotlidec.so.1")
        except OSError as e:  # pragma: no cover - the image ships both
            raise RuntimeError(f"libbrotlienc/libbrotlidec not found: run folders need Brotli ({e})") from e
        self.enc.BrotliEncoderMaxCompressedSize.restype = C.c_size_t
        self.enc.BrotliEncoderMaxCompressedSize.argtypes = [C.c_size_t]
        self.enc.BrotliEncoderCompress.restype = C.c_int
        self.enc.BrotliEncoderCompress.argtypes = [C.c_int, C.c_int, C.c_int, C.c_size_t, C.c_void_p,
                                                   C.POINTER(C.c_size_t), C.c_void_p]
        self.dec.BrotliDecoderCreateInstance.restype = C.c_void_p
        self.dec.BrotliDecoderCreateInstance.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        self.dec.BrotliDecoderDestroyInstance.restype = None
        self.dec.BrotliDecoderDestroyInstance.argtypes = [C.c_void_p]
        self.dec.BrotliDecoderDecompressStream.restype = C.c_int
        self.dec.BrotliDecoderDecompressStream.argtypes = [C.c_void_p, C.POINTER(C.c_size_t),
                                                           C.POINTER(C.c_void_p), C.POINTER(C.c_size_t),
                                                           C.POINTER(C.c_void_p), C.c_void_p]

    def compress(self, data: bytes, quality: int = 1, lgwin: int = 22) -> bytes:
        src = C.create_string_buffer(bytes(data), len(data))
        cap = int(self.enc.BrotliEncoderMaxCompressedSize(len(data))) or (len(data) + 1024)
        dst = C.create_string_buffer(cap)
        n = C.c_size_t(cap)
        if not self.enc.BrotliEncoderCompress(quality, lgwin, 0, len(data), src, C.byref(n), dst):
            raise IOError("Brotli compression failed")
        return dst.raw[:n.value]

    def decompress(self, data: bytes, expected: int = 0) -> bytes:
        """Decode a whole Brotli stream (output size unknown in advance, as for an extended checkpoint)."""
        st = self.dec.BrotliDecoderCreateInstance(None, None, None)
        if not st:
            raise MemoryError("Brotli decoder")
        try:
            src = C.create_string_buffer(bytes(data), len(data))
            avail_in = C.c_size_t(len(data))
            next_in = C.c_void_p(C.addressof(src))
            chunk = max(1 << 16, expected + 64)
            out = bytearray()
            while True:
                buf = C.create_string_buffer(chunk)
                avail_out = C.c_size_t(chunk)
                next_out = C.c_void_p(C.addressof(buf))
                r = self.dec.BrotliDecoderDecompressStream(st, C.byref(avail_in), C.byref(next_in),
                                                           C.byref(avail_out), C.byref(next_out), None)
                out += buf.raw[:chunk - avail_out.value]
                if r == 1:  # BROTLI_DECODER_RESULT_SUCCESS
                    return bytes(out)
                if r == 3:  # NEEDS_MORE_OUTPUT
                    continue
                if r == 2:  # NEEDS_MORE_INPUT: the stream ended early
                    raise EOFError("Unexpected end of compressed stream.")
                raise IOError("corrupt Brotli stream")
        finally:
            self.dec.BrotliDecoderDestroyInstance(st)


_BROTLI: Optional[_Brotli] = None


def _brotli() -> _Brotli:
    global _BROTLI
    if _BROTLI is None:
        _BROTLI = _Brotli()
    return _BROTLI


def brotli_compress(data: bytes) -> bytes:
    return _brotli().compress(data)


def brotli_decompress(data: bytes, expected: int = 0) -> bytes:
    return _brotli().decompress(data, expected)


# ---- 8-bit frames and checksums ---------------------------------------------------------------------

def quantize(phys) -> Tuple[np.ndarray, int]:
    """RunWriter.cs:215-229 and :376-388: the 8-bit frame of a float32 density and its FNV-1a checksum."""
    a = np.ascontiguousarray(phys, np.float32)
    q = np.empty(a.size, np.uint8)
    h = C.c_uint32()
    L.check(L.load().mycel_run_quantize(L.ptr(a) if a.size else None, int(a.size), L.ptr(q) if a.size else None,
                                        C.byref(h)))
    return q, int(h.value)


def compute_quantized_checksum(phys) -> int:
    """RunWriter.ComputeQuantizedChecksum (RunWriter.cs:376-388)."""
    return quantize(phys)[1]


def checksum_bytes(q) -> int:
    """FNV-1a over frame bytes, as RunReader.ReadIteration recomputes it (RunReader.cs:213-223)."""
    a = np.ascontiguousarray(q, np.uint8)
    h = C.c_uint32()
    L.check(L.load().mycel_run_checksum(L.ptr(a) if a.size else None, int(a.size), C.byref(h)))
    return int(h.value)


# ---- .NET number and date formatting (InvariantCulture) -------------------------------------------

def _nonfinite(x: float) -> Optional[str]:
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    return None


def fmt_fixed(x: float, digits: int) -> str:
    """'F<digits>': the exact binary value rounded half away from zero, as .NET formats it."""
    s = _nonfinite(x)
    if s is not None:
        return s
    d = Decimal(x).quantize(Decimal(1).scaleb(-digits), rounding=ROUND_HALF_UP)
    return f"{d:.{digits}f}"


def fmt_exp(x: float, digits: int) -> str:
    """'E<digits>': d.ddddE+ddd (at least three exponent digits), mantissa rounded half away from zero."""
    s = _nonfinite(x)
    if s is not None:
        return s
    if x == 0.0:
        return ("-" if math.copysign(1.0, x) < 0 else "") + "0." + "0" * digits + "E+000"
    d = Decimal(x)
    e = d.adjusted()
    m = d.scaleb(-e).quantize(Decimal(1).scaleb(-digits), rounding=ROUND_HALF_UP)
    if abs(m) >= 10:
        e += 1
        m = d.scaleb(-e).quantize(Decimal(1).scaleb(-digits), rounding=ROUND_HALF_UP)
    return f"{m:.{digits}f}E{'+' if e >= 0 else '-'}{abs(e):03d}"


def fmt_g(x: float, precision: int) -> str:
    """'G<precision>': that many significant digits, trailing zeros dropped; fixed-point when the scientific
    exponent is greater than -5 and less than the precision, scientific (E+dd / E-dd) otherwise."""
    s = _nonfinite(x)
    if s is not None:
        return s
    if x == 0.0:
        return "-0" if math.copysign(1.0, x) < 0 else "0"
    q = Decimal(1).scaleb(-(precision - 1))
    d = Decimal(x)
    e = d.adjusted()
    m = d.scaleb(-e).quantize(q, rounding=ROUND_HALF_UP)
    if abs(m) >= 10:
        e += 1
        m = d.scaleb(-e).quantize(q, rounding=ROUND_HALF_UP)
    digits = str(abs(m)).replace(".", "").rstrip("0") or "0"
    sign = "-" if x < 0 else ""
    if e <= -5 or e >= precision:
        mant = digits[0] + ("." + digits[1:] if len(digits) > 1 else "")
        return f"{sign}{mant}E{'+' if e >= 0 else '-'}{abs(e):02d}"
    if e >= 0:
        if len(digits) <= e + 1:
            return sign + digits + "0" * (e + 1 - len(digits))
        return sign + digits[:e + 1] + "." + digits[e + 1:]
    return sign + "0." + "0" * (-e - 1) + digits


def fmt_g17(x: float) -> str:
    """'G17' (the round-trip form RunWriter uses for the compliance)."""
    return fmt_g(x, 17)


def format_metrics_row(iteration: int, compliance: float, volume_fraction: float, max_change: float,
                       grey_level: float, solver_iterations: int, solver_residual: float, total_ms: float,
                       disk_bytes: int) -> str:
    """One metrics.csv row: "{0},{1:G17},{2:F6},{3:F6},{4:F4},{5},{6:E4},{7:F2},{8}" (RunWriter.cs:170-181)."""
    return ",".join([str(int(iteration)), fmt_g17(compliance), fmt_fixed(volume_fraction, 6),
                     fmt_fixed(max_change, 6), fmt_fixed(grey_level, 4), str(int(solver_iterations)),
                     fmt_exp(solver_residual, 4), fmt_fixed(total_ms, 2), str(int(disk_bytes))])


METRICS_HEADER = ("iteration,compliance,volume_fraction,max_change,grey_level,solver_iterations,solver_residual,"
                  "total_ms,disk_bytes")


def _utc_now() -> _dt.datetime:
    return _dt.datetime.now(_dt.timezone.utc)


def _fmt_time(t: _dt.datetime) -> str:
    """DateTime.UtcNow as System.Text.Json writes it: yyyy-MM-ddTHH:mm:ss.fffffffZ."""
    t = t.astimezone(_dt.timezone.utc)
    return t.strftime("%Y-%m-%dT%H:%M:%S.") + f"{t.microsecond:06d}0Z"


def _parse_time(s: str) -> _dt.datetime:
    s = s.rstrip("Z")
    base, _, frac = s.partition(".")
    t = _dt.datetime.strptime(base, "%Y-%m-%dT%H:%M:%S")
    return t.replace(microsecond=int((frac + "000000")[:6]), tzinfo=_dt.timezone.utc)


# ---- Run metadata (RunModels.cs) -------------------------------------------------------------------

class RunStatus:
    """RunModels.cs:12-19, serialized by name."""

    Running = "Running"
    Completed = "Completed"
    Stopped = "Stopped"
    Failed = "Failed"
    Interrupted = "Interrupted"
    ALL = (Running, Completed, Stopped, Failed, Interrupted)


def _pla() -> Dict:
    """MaterialDefinition.CreatePreset(Pla) (Workflow/MaterialPresets.cs:94-104), with the get-only
    Effective* properties System.Text.Json also writes (at 100 % infill they equal E, YieldXy, YieldZ)."""
    return {"PresetKind": 0, "E": 3000.0, "Nu": 0.36, "YieldXy": 50.0, "YieldZ": 30.0, "ConservativeZ": False,
            "Density": 1.24, "InfillPercent": 100.0, "InfillPattern": 0, "EffectiveE": 3000.0,
            "EffectiveYieldXy": 50.0, "EffectiveYieldZ": 30.0}


@dataclass
class RunMetadata:
    """RunMetadata (RunModels.cs:54-71). Grid, Options and NodeSets are the DTO records (:21-52) as
    dictionaries with the same property names."""

    FormatVersion: int = 1
    AppVersion: str = "1.0.0"
    RunName: str = "run"
    InputStlSha256: str = ""
    Grid: Dict = field(default_factory=lambda: {"OriginX": 0.0, "OriginY": 0.0, "OriginZ": 0.0, "H": 1.0,
                                                "Nx": 1, "Ny": 1, "Nz": 1})
    Material: Dict = field(default_factory=_pla)
    Options: Dict = field(default_factory=lambda: {"VolumeFraction": 0.3, "FilterRadiusMm": 1.2, "Penal": 3.0,
                                                   "MoveLimit": 0.2, "MaxIterations": 100, "EminRatio": 1e-6,
                                                   "Continuation": False, "EnableHeaviside": False,
                                                   "ProtectBoundaryHinges": True, "KeepEveryN": 1})
    NodeSets: Dict = field(default_factory=lambda: {"FixedDofs": [], "Loads": []})
    Status: str = RunStatus.Running
    StartedAtUtc: _dt.datetime = field(default_factory=_utc_now)
    UpdatedAtUtc: _dt.datetime = field(default_factory=_utc_now)
    CompletedIterations: int = 0
    TotalDiskBytes: int = 0
    OriginalSolidCompliance: float = 0.0
    ErrorMessage: Optional[str] = None

    @staticmethod
    def for_grid(nx: int, ny: int, nz: int, h: float = 1.0, origin=(0.0, 0.0, 0.0), options=None,
                 run_name: str = "run", keep_every_n: int = 1, fixed_dofs=None) -> "RunMetadata":
        m = RunMetadata(RunName=run_name)
        m.Grid = {"OriginX": float(origin[0]), "OriginY": float(origin[1]), "OriginZ": float(origin[2]),
                  "H": float(h), "Nx": int(nx), "Ny": int(ny), "Nz": int(nz)}
        if options is not None:  # an OptimizerOptions (OptimizerOptionsDto, RunModels.cs:38-48)
            m.Options = {"VolumeFraction": float(options.VolumeFraction),
                         "FilterRadiusMm": float(options.FilterRadiusMm), "Penal": float(options.Penal),
                         "MoveLimit": float(options.MoveLimit), "MaxIterations": int(options.MaxIterations),
                         "EminRatio": float(options.EminRatio), "Continuation": bool(options.Continuation),
                         "EnableHeaviside": bool(options.EnableHeaviside),
                         "ProtectBoundaryHinges": bool(options.ProtectBoundaryHinges), "KeepEveryN": 1}
        m.Options["KeepEveryN"] = int(keep_every_n)
        if fixed_dofs is not None:
            m.NodeSets = {"FixedDofs": [int(d) for d in fixed_dofs], "Loads": []}
        return m

    def to_json(self) -> str:
        d = {"FormatVersion": self.FormatVersion, "AppVersion": self.AppVersion, "RunName": self.RunName,
             "InputStlSha256": self.InputStlSha256, "Grid": self.Grid, "Material": self.Material,
             "Options": self.Options, "NodeSets": self.NodeSets, "Status": self.Status,
             "StartedAtUtc": _fmt_time(self.StartedAtUtc), "UpdatedAtUtc": _fmt_time(self.UpdatedAtUtc),
             "CompletedIterations": self.CompletedIterations, "TotalDiskBytes": self.TotalDiskBytes,
             "OriginalSolidCompliance": self.OriginalSolidCompliance, "ErrorMessage": self.ErrorMessage}
        return json.dumps(d, indent=2)

    @staticmethod
    def from_json(text: str) -> "RunMetadata":
        d = json.loads(text)
        m = RunMetadata()
        for k in ("FormatVersion", "AppVersion", "RunName", "InputStlSha256", "CompletedIterations",
                  "TotalDiskBytes", "OriginalSolidCompliance", "ErrorMessage"):
            if k in d:
                setattr(m, k, d[k])
        for k in ("Grid", "Material", "Options", "NodeSets"):
            if isinstance(d.get(k), dict):
                getattr(m, k).update(d[k])
        st = d.get("Status", RunStatus.Running)
        if isinstance(st, int):  # an enum written by number
            st = RunStatus.ALL[st]
        if st not in RunStatus.ALL:
            raise ValueError(f"unknown run status {st!r}")
        m.Status = st
        for k in ("StartedAtUtc", "UpdatedAtUtc"):
            if k in d and d[k]:
                setattr(m, k, _parse_time(d[k]))
        return m

    @property
    def total_elements(self) -> int:
        return int(self.Grid["Nx"]) * int(self.Grid["Ny"]) * int(self.Grid["Nz"])


@dataclass
class IterationHeader:
    """RunModels.cs:73-85."""

    Version: int
    Nx: int
    Ny: int
    Nz: int
    Iteration: int
    Compliance: float
    VolumeFraction: float
    MaxChange: float
    GreyLevel: float
    SolverIterations: int
    IterationMs: float
    Checksum: int


@dataclass
class IterationMetricRow:
    """RunModels.cs:87-96."""

    Iteration: int
    Compliance: float
    VolumeFraction: float
    MaxChange: float
    GreyLevel: float
    SolverIterations: int
    SolverResidual: float
    TotalMs: float
    DiskBytes: int


@dataclass
class Checkpoint:
    """What ReadLatestCheckpoint returns (RunReader.cs:228-263), plus the FP64 extension when present."""

    Iteration: int
    DesignDensity: np.ndarray                 # float32, box-sized
    PhysicalDensity: np.ndarray               # float32, box-sized
    DesignDensity64: Optional[np.ndarray] = None   # FP64, active-compact
    Displacements: Optional[np.ndarray] = None     # FP64 box DOFs (or (n_cases, DOFs))
    Lambda: Optional[float] = None
    # The loop state after the checkpointed iteration (the "SIMP" block of the extension): its max change, the
    # Heaviside beta of the next iteration and the last <= 6 compliances, oldest first.
    MaxChange: Optional[float] = None
    BetaNext: Optional[float] = None
    ComplianceHistory: Optional[List[float]] = None


# ---- file helpers -----------------------------------------------------------------------------------

def _atomic_write(path: str, data: bytes) -> None:
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(data)
        fh.flush()
        os.fsync(fh.fileno())
    os.replace(tmp, path)


def _dir_bytes(folder: str) -> int:
    total = 0
    for root, _, files in os.walk(folder):
        for f in files:
            if f.lower().endswith(".tmp"):
                continue
            try:
                total += os.path.getsize(os.path.join(root, f))
            except OSError:
                pass
    return total


def _cleanup_tmp(folder: str) -> None:
    if not os.path.isdir(folder):
        return
    for f in os.listdir(folder):
        if f.endswith(".tmp"):
            try:
                os.remove(os.path.join(folder, f))
            except OSError:
                pass


def _box_f32(a, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    if a is not None:
        v = np.asarray(a, np.float32).ravel()
        out[:min(n, v.size)] = v[:n]
    return out


def encode_iteration(nx: int, ny: int, nz: int, iteration: int, compliance: float, volume_fraction: float,
                     max_change: float, grey_level: float, solver_iterations: int, total_ms: float,
                     physical_density) -> Tuple[bytes, int]:
    """The bytes of it_NNNNNN.bin.br (RunWriter.cs:201-261) and the frame checksum."""
    total = nx * ny * nz
    q, h = quantize(_box_f32(physical_density, total))
    hdr = b"TOPO" + struct.pack("<iiiiidddd", 1, nx, ny, nz, iteration, compliance, volume_fraction, max_change,
                                grey_level) + struct.pack("<idI", solver_iterations, total_ms, h)
    assert len(hdr) == TOPO_HEADER
    return hdr + brotli_compress(q.tobytes()), h


def encode_checkpoint(iteration: int, design, physical, design64=None, displacements=None,
                      lam: Optional[float] = None, loop_state=None) -> bytes:
    """checkpoint_latest.bin.br (RunWriter.cs:263-295); with design64 / displacements / lam the FP64 block is
    appended after the two float32 arrays (version 2), and with loop_state = (max_change, beta_next,
    compliance_history) a "SIMP" block after it."""
    d = np.ascontiguousarray(design, np.float32).ravel()
    p = np.ascontiguousarray(physical, np.float32).ravel()
    if p.size != d.size:
        p = np.zeros(d.size, np.float32)  # RunWriter.cs:283-290
    ext = design64 is not None or displacements is not None or lam is not None
    payload = d.tobytes() + p.tobytes()
    if ext:
        g = np.ascontiguousarray(design64 if design64 is not None else np.zeros(0), np.float64).ravel()
        u = np.ascontiguousarray(displacements if displacements is not None else np.zeros(0), np.float64)
        cases = u.shape[0] if u.ndim == 2 else 1
        payload += _EXT_MAGIC + struct.pack("<dqqi", float(lam) if lam is not None else float("nan"), g.size,
                                            u.size, cases) + g.tobytes() + u.ravel().tobytes()
        if loop_state is not None:
            mc, beta, hist = loop_state
            hist = [float(c) for c in list(hist)[-6:]]
            payload += _LOOP_MAGIC + struct.pack("<ddi6d", float(mc), float(beta), len(hist),
                                                 *(hist + [0.0] * (6 - len(hist))))
    return b"CHKP" + struct.pack("<iii", 2 if ext else 1, iteration, d.size) + brotli_compress(payload)


def decode_checkpoint(blob: bytes) -> Checkpoint:
    if len(blob) < CHKP_HEADER or blob[:4] != b"CHKP":
        raise ValueError("not a checkpoint")
    _ver, it, n = struct.unpack_from("<iii", blob, 4)
    raw = brotli_decompress(blob[CHKP_HEADER:], 8 * n)
    if len(raw) < 8 * n:
        raise EOFError("Unexpected end of compressed stream.")
    ck = Checkpoint(it, np.frombuffer(raw[:4 * n], np.float32).copy(),
                    np.frombuffer(raw[4 * n:8 * n], np.float32).copy())
    rest = raw[8 * n:]
    if rest[:4] == _EXT_MAGIC:
        lam, ng, nu, cases = struct.unpack_from("<dqqi", rest, 4)
        o = 4 + struct.calcsize("<dqqi")
        g = np.frombuffer(rest[o:o + 8 * ng], np.float64).copy()
        u = np.frombuffer(rest[o + 8 * ng:o + 8 * (ng + nu)], np.float64).copy()
        if g.size != ng or u.size != nu:
            raise EOFError("truncated checkpoint extension")
        ck.DesignDensity64 = g if ng else None
        ck.Displacements = (u.reshape(cases, -1) if cases > 1 else u) if nu else None
        ck.Lambda = None if math.isnan(lam) else lam
        tail = rest[o + 8 * (ng + nu):]
        if tail[:4] == _LOOP_MAGIC:
            vals = struct.unpack_from("<ddi6d", tail, 4)
            ck.MaxChange, ck.BetaNext = vals[0], vals[1]
            ck.ComplianceHistory = list(vals[3:3 + max(0, min(6, vals[2]))])
    return ck


# ---- writer (RunWriter.cs) -------------------------------------------------------------------------

class RunWriter:
    """Single-writer persistence of a run folder (RunWriter.cs:19-404). Use as the progress callback of
    TopologyOptimizer.Run: ``w.WriteSnapshot(snapshot)`` per iteration, then ``w.FinalizeRun(status)``."""

    def __init__(self, run_folder: str, metadata: RunMetadata, domain_mask, passive_mask=None,
                 is_resume: bool = False):
        if not run_folder or not str(run_folder).strip():
            raise L.ArgumentError(L.E_ARGUMENT, "runFolder")
        if metadata is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "metadata")
        self.RunFolder = os.path.abspath(run_folder)
        self._iter_dir = os.path.join(self.RunFolder, "iterations")
        self._run_json = os.path.join(self.RunFolder, "run.json")
        self._metrics = os.path.join(self.RunFolder, "metrics.csv")
        self._domain = os.path.join(self.RunFolder, "domain.bin.br")
        self._checkpoint = os.path.join(self.RunFolder, "checkpoint_latest.bin.br")
        self.Metadata = metadata
        self._keep = max(1, int(metadata.Options.get("KeepEveryN", 1)))
        self.WrittenSnapshotCount = 0
        # The last six compliances, carried in the checkpoint's loop state; a resumed writer continues them.
        self._history: List[float] = []
        self._history_iter = -1
        if is_resume and os.path.exists(self._checkpoint):
            try:
                with open(self._checkpoint, "rb") as fh:
                    self._history = list(decode_checkpoint(fh.read()).ComplianceHistory or [])
            except (OSError, ValueError, EOFError, struct.error):
                self._history = []
        self.RollingStreamChecksum = FNV_OFFSET
        os.makedirs(self._iter_dir, exist_ok=True)
        _cleanup_tmp(self.RunFolder)
        _cleanup_tmp(self._iter_dir)
        if not is_resume or not os.path.exists(self._domain):
            self._write_domain(domain_mask, passive_mask)
        header = (not is_resume or not os.path.exists(self._metrics) or os.path.getsize(self._metrics) == 0)
        fresh = not os.path.exists(self._metrics) or os.path.getsize(self._metrics) == 0
        self._csv = open(self._metrics, "a", encoding="utf-8", newline="")
        if fresh:
            self._csv.write("\ufeff")  # StreamWriter with Encoding.UTF8 starts a new file with its preamble
        if header:
            self._csv.write(METRICS_HEADER + "\n")
            self._csv.flush()
        self.RunningDiskBytes = _dir_bytes(self.RunFolder)
        metadata.Status = RunStatus.Running
        metadata.UpdatedAtUtc = _utc_now()
        metadata.TotalDiskBytes = self.RunningDiskBytes
        self._write_metadata()

    @staticmethod
    def EstimateDiskBytes(total_box_elements: int, max_iterations: int, keep_every_n: int = 1) -> int:
        """RunWriter.cs:120-132."""
        keep = max(1, keep_every_n)
        frames = max_iterations if keep == 1 else 1 + max_iterations // keep + (1 if max_iterations % keep else 0)
        per_frame = max(512, int(total_box_elements * 0.22) + 72)
        domain = max(512, int(total_box_elements * 0.10))
        chk = int(total_box_elements * 2.0)
        return domain + chk + frames * per_frame + max_iterations * 80

    def _write_domain(self, domain_mask, passive_mask) -> None:
        nx, ny, nz = (int(self.Metadata.Grid[k]) for k in ("Nx", "Ny", "Nz"))
        total = nx * ny * nz
        dom = np.zeros(total, np.uint8)
        if domain_mask is not None:
            d = np.asarray(domain_mask, np.uint8).ravel()
            dom[:min(total, d.size)] = d[:total]
        pas = np.zeros(total, np.uint8)
        if passive_mask is not None:
            p = np.asarray(passive_mask, np.uint8).ravel()
            pas[:min(total, p.size)] = p[:total]
        blob = b"DOMN" + struct.pack("<iiii", 1, nx, ny, nz) + brotli_compress(dom.tobytes() + pas.tobytes())
        _atomic_write(self._domain, blob)

    def _write_metadata(self) -> None:
        tmp = self._run_json + ".tmp"
        with open(tmp, "w", encoding="utf-8-sig") as fh:  # File.WriteAllText(..., Encoding.UTF8) writes a BOM
            fh.write(self.Metadata.to_json())
        os.replace(tmp, self._run_json)

    def WriteSnapshot(self, snapshot, isFinal: bool = False, displacements=None, design64=None) -> None:
        """RunWriter.WriteSnapshotAsync (RunWriter.cs:134-196). snapshot: an IterationSnapshot with float32
        box-sized PhysicalDensity and DesignDensity. With displacements / design64 (FP64) the checkpoint
        carries the FP64 extension (with the snapshot's Lambda)."""
        if snapshot is None:
            raise L.ArgumentNullError(L.E_ARGUMENT_NULL, "snapshot")
        g = self.Metadata.Grid
        nx, ny, nz = int(g["Nx"]), int(g["Ny"]), int(g["Nz"])
        it = int(snapshot.Iteration)
        keep = self._keep <= 1 or it == 1 or it % self._keep == 0 or isFinal
        phys = snapshot.PhysicalDensity
        if keep:
            blob, h = encode_iteration(nx, ny, nz, it, snapshot.Compliance, snapshot.VolumeFraction,
                                       snapshot.MaxChange, snapshot.GreyLevel, snapshot.SolverIterations,
                                       snapshot.TotalMs, phys)
            _atomic_write(os.path.join(self._iter_dir, f"it_{it:06d}.bin.br"), blob)
        else:
            h = compute_quantized_checksum(_box_f32(phys, nx * ny * nz) if phys is not None else np.zeros(0))
        design = getattr(snapshot, "DesignDensity", None)
        d = np.asarray(design if design is not None else np.zeros(0), np.float32)
        p = np.asarray(phys if phys is not None else np.zeros(0), np.float32)
        if it != self._history_iter:  # a final re-write of the same iteration does not add to the history
            self._history = (self._history + [float(snapshot.Compliance)])[-6:]
            self._history_iter = it
        ext = displacements is not None or design64 is not None
        lam = getattr(snapshot, "Lambda", None) if ext else None
        loop = (snapshot.MaxChange, getattr(snapshot, "BetaNext", 1.0), self._history) if ext else None
        _atomic_write(self._checkpoint, encode_checkpoint(it, d, p, design64, displacements, lam, loop))
        self.RollingStreamChecksum = ((self.RollingStreamChecksum ^ h) * FNV_PRIME) & 0xFFFFFFFF
        self.WrittenSnapshotCount += 1
        total = _dir_bytes(self.RunFolder)
        self.RunningDiskBytes = total
        if self._csv is not None:
            self._csv.write(format_metrics_row(it, snapshot.Compliance, snapshot.VolumeFraction, snapshot.MaxChange,
                                               snapshot.GreyLevel, snapshot.SolverIterations,
                                               snapshot.SolverResidual, snapshot.TotalMs, total) + "\n")
            self._csv.flush()
        self.Metadata.CompletedIterations = max(self.Metadata.CompletedIterations, it)
        self.Metadata.UpdatedAtUtc = _utc_now()
        self.Metadata.TotalDiskBytes = total
        self._write_metadata()

    def __call__(self, snapshot) -> None:
        self.WriteSnapshot(snapshot)

    def FinalizeRun(self, final_status: str, error_message: Optional[str] = None) -> None:
        """RunWriter.cs:187-196."""
        if final_status not in RunStatus.ALL:
            raise L.ArgumentError(L.E_ARGUMENT, f"unknown run status {final_status!r}")
        if self._csv is not None:
            self._csv.flush()
        self.Metadata.Status = final_status
        self.Metadata.ErrorMessage = error_message
        self.Metadata.UpdatedAtUtc = _utc_now()
        self.Metadata.TotalDiskBytes = _dir_bytes(self.RunFolder)
        self.RunningDiskBytes = self.Metadata.TotalDiskBytes
        self._write_metadata()

    def Dispose(self) -> None:
        if self._csv is not None:
            self._csv.close()
            self._csv = None

    close = Dispose

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.Dispose()


# ---- reader (RunReader.cs) -------------------------------------------------------------------------

class RunReader:
    """Random-access reader of a run folder (RunReader.cs:20-413; mesh extraction is not on this path)."""

    def __init__(self, run_folder: str, metadata: RunMetadata):
        self.RunFolder = run_folder
        self._iter_dir = os.path.join(run_folder, "iterations")
        self.Metadata = metadata

    @staticmethod
    def Open(run_folder: str, is_live_active_run: bool = False) -> "RunReader":
        """RunReader.cs:52-85: a run left Running with no live owner reads as Interrupted, and
        CompletedIterations follows the committed iteration files."""
        if not run_folder or not str(run_folder).strip():
            raise L.ArgumentError(L.E_ARGUMENT, "runFolder")
        full = os.path.abspath(run_folder)
        if not os.path.isdir(full):
            raise FileNotFoundError(f"Run folder not found: {full}")
        path = os.path.join(full, "run.json")
        if not os.path.exists(path):
            raise FileNotFoundError("run.json not found in run folder.")
        with open(path, encoding="utf-8-sig") as fh:
            meta = RunMetadata.from_json(fh.read())
        if not is_live_active_run and meta.Status == RunStatus.Running:
            meta.Status = RunStatus.Interrupted
        r = RunReader(full, meta)
        avail = r.GetAvailableIterations()
        if avail and meta.CompletedIterations < avail[-1]:
            meta.CompletedIterations = avail[-1]
        return r

    @staticmethod
    def ListRuns(base_runs_dir: str) -> List[Tuple[str, RunMetadata]]:
        """RunReader.cs:87-120: newest folder name first; malformed folders are skipped."""
        out: List[Tuple[str, RunMetadata]] = []
        if not os.path.isdir(base_runs_dir):
            return out
        for d in sorted((os.path.join(base_runs_dir, x) for x in os.listdir(base_runs_dir)), reverse=True):
            p = os.path.join(d, "run.json")
            if not os.path.isdir(d) or not os.path.exists(p):
                continue
            try:
                with open(p, encoding="utf-8-sig") as fh:
                    m = RunMetadata.from_json(fh.read())
            except (ValueError, OSError):
                continue
            if m.Status == RunStatus.Running:
                m.Status = RunStatus.Interrupted
            out.append((d, m))
        return out

    def GetAvailableIterations(self) -> List[int]:
        """Committed it_NNNNNN.bin.br files, sorted; .tmp files never count (RunReader.cs:126-150)."""
        if not os.path.isdir(self._iter_dir):
            return []
        its = []
        for name in os.listdir(self._iter_dir):
            if name.startswith("it_") and name.endswith(".bin.br") and not name.lower().endswith(".tmp"):
                num = name[3:-7]
                if num.lstrip("+-").isdigit():
                    its.append(int(num))
        return sorted(its)

    def ReadDomainMasks(self) -> Tuple[np.ndarray, np.ndarray]:
        """RunReader.cs:152-172."""
        path = os.path.join(self.RunFolder, "domain.bin.br")
        if not os.path.exists(path):
            raise FileNotFoundError("domain.bin.br not found.")
        with open(path, "rb") as fh:
            blob = fh.read()
        if blob[:4] != b"DOMN":
            raise ValueError("Invalid magic in domain.bin.br.")
        _, nx, ny, nz = struct.unpack_from("<iiii", blob, 4)
        total = nx * ny * nz
        raw = brotli_decompress(blob[DOMN_HEADER:], 2 * total)
        if len(raw) < 2 * total:
            raise EOFError("Unexpected end of compressed stream.")
        return (np.frombuffer(raw[:total], np.uint8).copy(), np.frombuffer(raw[total:2 * total], np.uint8).copy())

    def _resolve(self, requested: int) -> int:
        """The exact committed iteration, else the nearest one, the earliest on a tie (RunReader.cs:374-398)."""
        if os.path.exists(os.path.join(self._iter_dir, f"it_{requested:06d}.bin.br")):
            return requested
        avail = self.GetAvailableIterations()
        if not avail:
            return requested
        best = avail[0]
        for a in avail[1:]:
            if abs(a - requested) < abs(best - requested):
                best = a
        return best

    def ReadIteration(self, iteration: int) -> Tuple[IterationHeader, np.ndarray]:
        """RunReader.cs:174-224: header, then the frame, with the checksum verified."""
        it = self._resolve(iteration)
        name = f"it_{it:06d}.bin.br"
        path = os.path.join(self._iter_dir, name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"Iteration file not found for iteration {iteration}.")
        with open(path, "rb") as fh:
            blob = fh.read()
        if blob[:4] != b"TOPO":
            raise ValueError(f"Invalid magic header in {name}.")
        ver, nx, ny, nz, itn, comp, vol, chg, grey = struct.unpack_from("<iiiiidddd", blob, 4)
        s_it, ms, h = struct.unpack_from("<idI", blob, 56)
        total = nx * ny * nz
        raw = brotli_decompress(blob[TOPO_HEADER:], total)
        if len(raw) < total:
            raise EOFError("Unexpected end of compressed stream.")
        q = np.frombuffer(raw[:total], np.uint8)
        got = checksum_bytes(q)
        if got != h:
            raise ValueError(f"Checksum mismatch in {name}: expected {h}, got {got}.")
        dens = q.astype(np.float32) * np.float32(1.0 / 255.0)
        return IterationHeader(ver, nx, ny, nz, itn, comp, vol, chg, grey, s_it, ms, h), dens

    def ReadIterationDensity(self, iteration: int) -> np.ndarray:
        return self.ReadIteration(iteration)[1]

    def ReadLatestCheckpoint(self) -> Optional[Checkpoint]:
        """RunReader.cs:228-263: the checkpoint if it reads, else the latest iteration frame (design =
        physical), else None."""
        path = os.path.join(self.RunFolder, "checkpoint_latest.bin.br")
        if os.path.exists(path):
            try:
                with open(path, "rb") as fh:
                    return decode_checkpoint(fh.read())
            except (ValueError, EOFError, IOError, struct.error):
                pass
        avail = self.GetAvailableIterations()
        if not avail:
            return None
        d = self.ReadIterationDensity(avail[-1])
        return Checkpoint(avail[-1], d.copy(), d)

    def ReadMetrics(self) -> List[IterationMetricRow]:
        """RunReader.cs:265-311: rows that parse; a header line and short or malformed lines are skipped."""
        path = os.path.join(self.RunFolder, "metrics.csv")
        rows: List[IterationMetricRow] = []
        if not os.path.exists(path):
            return rows
        with open(path, encoding="utf-8-sig") as fh:
            for n, line in enumerate(fh):
                line = line.rstrip("\r\n")
                if n == 0 and line.lower().startswith("iteration"):
                    continue
                if not line.strip():
                    continue
                parts = line.split(",")
                if len(parts) < 8:
                    continue
                try:
                    disk = int(parts[8]) if len(parts) >= 9 else 0
                except ValueError:
                    disk = 0
                try:
                    rows.append(IterationMetricRow(int(parts[0]), _pf(parts[1]), _pf(parts[2]), _pf(parts[3]),
                                                   _pf(parts[4]), int(parts[5]), _pf(parts[6]), _pf(parts[7]),
                                                   disk))
                except ValueError:
                    continue
        return rows


def _pf(s: str) -> float:
    t = s.strip()
    if t in ("NaN", "Infinity", "-Infinity", "∞", "-∞"):
        return float(t.replace("∞", "inf").replace("Infinity", "inf"))
    return float(t)


def resume_options(checkpoint: Checkpoint, options):
    """OptimizerOptions that continue a run from a checkpoint: StartIteration = checkpoint + 1, the design
    (FP64 from the extension, else the float32 DesignDensity, as the reference's InitialDensity), and, from
    the extension, the warm-start displacements and the OC multiplier."""
    import dataclasses

    o = dataclasses.replace(options)
    o.StartIteration = int(checkpoint.Iteration) + 1
    if checkpoint.MaxChange is not None:  # continue the tolerance schedule, beta and the stop test exactly
        o.ResumeMaxChange = float(checkpoint.MaxChange)
        o.ResumeBeta = float(checkpoint.BetaNext if checkpoint.BetaNext is not None else 1.0)
        o.ResumeComplianceHistory = list(checkpoint.ComplianceHistory or [])
    if checkpoint.DesignDensity64 is not None:
        o.InitialDesign64 = checkpoint.DesignDensity64
    else:
        o.InitialDensity = checkpoint.DesignDensity
    if checkpoint.Displacements is not None:
        o.WarmStartDisplacements = checkpoint.Displacements
    if checkpoint.Lambda is not None:
        o.InitialLambda = float(checkpoint.Lambda)
    return o
