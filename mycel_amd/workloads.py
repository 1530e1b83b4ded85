"""Synthetic problems: the benchmark workloads and the reference tests' fixture models.

Random inputs come from one project-written generator (SplitMix64 -> 53-bit uniform doubles,
Box-Muller normals) so every consumer builds bit-identical inputs from a seed.

Box-path semantics follow the reference CLI's LoadModel (src/Mycel.Cli/Program.cs:168-255): a
patch's nodes carry area h^2 each, so a patch force is split equally over its nodes.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea & Flood 2014) with 53-bit uniform doubles."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & _M64)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            inc = np.uint64(0x9E3779B97F4A7C15)
            steps = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + steps * inc
            self.state = self.state + np.uint64(n) * inc
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]

    def integers(self, n: int, hi: int) -> np.ndarray:
        """Uniform integers in [0, hi)."""
        return np.minimum((self.uniform(n) * hi).astype(np.int64), hi - 1)


@dataclass
class Problem:
    """A structured Hex8 problem in the reference's conventions (box DOFs, ascending active elements)."""

    nx: int
    ny: int
    nz: int
    h: float = 1.0
    domain: Optional[np.ndarray] = None  # uint8 box elements (None = full box)
    fixed_dofs: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    loads: List[Tuple[np.ndarray, np.ndarray, np.ndarray]] = field(default_factory=list)  # (nodes, areas, force f32)
    e0: float = 1.0
    nu: float = 0.3
    emin_ratio: float = 1e-9
    penal: float = 3.0
    density: Optional[np.ndarray] = None  # box FP64 (None = 1.0)
    name: str = ""

    def __post_init__(self):
        if self.domain is None:
            self.domain = np.ones(self.nx * self.ny * self.nz, np.uint8)

    @property
    def n_elem(self):
        return self.nx * self.ny * self.nz

    @property
    def n_nodes(self):
        return (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    @property
    def n_dofs(self):
        return 3 * self.n_nodes

    @property
    def active(self) -> np.ndarray:
        return np.flatnonzero(self.domain).astype(np.int32)

    def node(self, i, j, k):
        return i + (self.nx + 1) * (j + (self.ny + 1) * k)

    def patch_nodes(self) -> np.ndarray:
        return np.concatenate([l[0] for l in self.loads]).astype(np.int32) if self.loads else np.zeros(0, np.int32)

    def density_box(self) -> np.ndarray:
        return np.ones(self.n_elem) if self.density is None else np.asarray(self.density, np.float64)

    def fe_model(self):
        """The mycel_amd FeModel of this problem."""
        from .fem import FeMaterial, FeModel, GridSpec, LoadCase, SurfacePatch

        g = GridSpec(self.nx, self.ny, self.nz, self.h)
        loads = [LoadCase(SurfacePatch(n, a), f) for n, a, f in self.loads]
        return FeModel(g, self.domain, FeMaterial(self.e0, self.nu, self.emin_ratio), self.fixed_dofs, loads,
                       self.density, self.penal)


def face_nodes_x(p: Problem, i: int) -> np.ndarray:
    j, k = np.meshgrid(np.arange(p.ny + 1), np.arange(p.nz + 1), indexing="xy")
    return (i + (p.nx + 1) * (j.reshape(-1) + (p.ny + 1) * k.reshape(-1))).astype(np.int32)


def fix_all(nodes) -> np.ndarray:
    """FeModel.FixDofs(nodes, true, true, true) (Fem/FeModel.cs:189-200)."""
    nodes = np.asarray(nodes, np.int64)
    return (3 * nodes[:, None] + np.arange(3)[None, :]).reshape(-1).astype(np.int32)


def _clamped_x0(p: Problem) -> None:
    """All DOFs of the x = 0 face fixed (ordered k, j as the reference fixtures build them)."""
    nodes = [p.node(0, j, k) for k in range(p.nz + 1) for j in range(p.ny + 1)]
    p.fixed_dofs = fix_all(nodes)


# --------------------------------------------------------------------------------------------
# Benchmark workloads
# --------------------------------------------------------------------------------------------
def cantilever_simp(nx=60, ny=20, nz=20, emin_ratio=1e-9) -> Problem:
    """SIMP cantilever: x = 0 clamped; total load -1 in z over the bottom edge nodes (nx, j, 0)."""
    p = Problem(nx, ny, nz, emin_ratio=emin_ratio, name=f"cantilever {nx}x{ny}x{nz}")
    _clamped_x0(p)
    tip = np.array([p.node(nx, j, 0) for j in range(ny + 1)], np.int32)
    p.loads = [(tip, np.ones(tip.size), np.array([0, 0, -1], np.float32))]
    return p


def cylinder_mask(nx=80, ny=40, nz=40, radius=12.0, cx=40.0, cz=20.0) -> np.ndarray:
    """Box minus a through-hole along y: an element is removed when its centre lies strictly inside the
    cylinder of the given radius about the axis (x = cx, z = cz)."""
    ix = np.arange(nx) + 0.5
    iz = np.arange(nz) + 0.5
    inside = ((ix[None, :] - cx) ** 2 + (iz[:, None] - cz) ** 2) < radius * radius  # [z, x]
    dom = np.ones((nz, ny, nx), np.uint8)
    dom[:, :, :] = (~inside)[:, None, :].astype(np.uint8)
    return dom.reshape(-1)


def matvec_pcg_workstation(seed: int = 1234, nx=80, ny=40, nz=40) -> Problem:
    """80x40x40 box with a radius-12 through-hole (about 110k active elements, 359k active DOFs);
    i.i.d. densities U[0.01, 1]; x = 0 clamped; 8 seeded 5x5-node patches on the x = nx face with
    N(0, I3) forces."""
    dom = cylinder_mask(nx, ny, nz)
    p = Problem(nx, ny, nz, domain=dom, name=f"holed box {nx}x{ny}x{nz}")
    _clamped_x0(p)
    rng = SplitMix64(seed)
    act = p.active
    rho = np.ones(p.n_elem)
    rho[act] = 0.01 + 0.99 * rng.uniform(act.size)
    p.density = rho
    loads = []
    for _ in range(8):
        j0, k0 = rng.integers(2, ny - 3)
        fvec = rng.normal(3).astype(np.float32)
        nodes = np.array([p.node(nx, j0 + a, k0 + b) for b in range(5) for a in range(5)], np.int32)
        loads.append((nodes, np.ones(nodes.size), fvec))
    p.loads = loads
    return p


def bimodal_density(p: Problem, seed: int = 77, sigma: float = 2.0, solid_fraction: float = 0.3,
                    void: float = 1e-3) -> np.ndarray:
    """Seeded N(0,1) noise -> Gaussian blur (sigma voxels, reflective) -> top 30% solid (1), rest `void`."""
    from scipy.ndimage import gaussian_filter

    rng = SplitMix64(seed)
    noise = rng.normal(p.n_elem).reshape(p.nz, p.ny, p.nx)
    sm = gaussian_filter(noise, sigma=sigma, mode="reflect")
    thr = np.quantile(sm, 1.0 - solid_fraction)
    return np.where(sm >= thr, 1.0, void).reshape(-1)


def precision_variant(seed: int = 77) -> Problem:
    """The SIMP cantilever mesh with a spatially correlated bimodal density field (~1e-9 contrast)."""
    p = cantilever_simp()
    p.density = bimodal_density(p, seed)
    p.name = "bimodal cantilever 60x20x20"
    return p


def multi_rhs(seed: int = 99, nx=100, ny=50, nz=50, n_rhs=16) -> Tuple[Problem, List[np.ndarray]]:
    """100x50x50, x = 0 clamped, n_rhs load cases of 1-4 seeded 4x4-node patches on the five other faces."""
    p = Problem(nx, ny, nz, name=f"multi-rhs {nx}x{ny}x{nz}")
    _clamped_x0(p)
    rng = SplitMix64(seed)
    cases = []
    for _ in range(n_rhs):
        m = int(rng.integers(1, 4)[0]) + 1
        f = []
        for _ in range(m):
            face = int(rng.integers(1, 5)[0]) + 1  # 1:+x 2:-y 3:+y 4:-z 5:+z
            a0, b0 = rng.integers(2, 1 << 30)
            fvec = rng.normal(3).astype(np.float32)
            if face == 1:
                j0, k0 = a0 % (ny - 2), b0 % (nz - 2)
                nodes = [p.node(nx, j0 + a, k0 + b) for b in range(4) for a in range(4)]
            elif face in (2, 3):
                jj = 0 if face == 2 else ny
                i0, k0 = 1 + a0 % (nx - 3), b0 % (nz - 2)
                nodes = [p.node(i0 + a, jj, k0 + b) for b in range(4) for a in range(4)]
            else:
                kk = 0 if face == 4 else nz
                i0, j0 = 1 + a0 % (nx - 3), b0 % (ny - 2)
                nodes = [p.node(i0 + a, j0 + b, kk) for b in range(4) for a in range(4)]
            nodes = np.array(nodes, np.int32)
            f.append((nodes, np.ones(nodes.size), fvec))
        cases.append(f)
    return p, cases


def case_forces(p: Problem, cases) -> np.ndarray:
    """(n_cases, n_dofs) assembled RHS vectors, one per load case of multi_rhs (FeModel.AssembleForces of
    each case on its own)."""
    keep = p.loads
    try:
        out = []
        for c in cases:
            p.loads = c
            out.append(p.fe_model().AssembleForces())
        return np.stack(out)
    finally:
        p.loads = keep


# --------------------------------------------------------------------------------------------
# Fixture models of the reference tests (tests/Mycel.Core.Tests)
# --------------------------------------------------------------------------------------------
def bar_model(length=100.0, width=10.0, height=10.0, h=2.5, e0=3000.0, nu=0.3, emin_ratio=1e-6):
    """CreateBarModel (Fem/FemValidationTests.cs:13-59): x = 0 clamped; tip patch at x = L with
    tributary areas (half edges, quarter corners). Returns (problem, tip_nodes, tip_areas)."""
    nx, ny, nz = int(round(length / h)), int(round(width / h)), int(round(height / h))
    p = Problem(nx, ny, nz, h=float(np.float32(h)), e0=e0, nu=nu, emin_ratio=emin_ratio, name="bar")
    _clamped_x0(p)
    nodes, areas = [], []
    for k in range(nz + 1):
        for j in range(ny + 1):
            nodes.append(p.node(nx, j, k))
            dy = 0.5 * h if j in (0, ny) else h
            dz = 0.5 * h if k in (0, nz) else h
            areas.append(dy * dz)
    return p, np.array(nodes, np.int32), np.array(areas)


def benchmark_cantilever(target: int, e0=3000.0):
    """V10 cantilever (Fem/FemValidationTests.cs:632-726): nz = round(cbrt(target/5)), nx = 5 nz,
    h = 1, tip load (0, 0, -1000) spread over all x = nx face nodes with area h^2."""
    nz = int(round((target / 5.0) ** (1.0 / 3.0)))
    p = Problem(5 * nz, nz, nz, e0=e0, nu=0.3, emin_ratio=1e-6, name=f"V10 {5 * nz}x{nz}x{nz}")
    _clamped_x0(p)
    tip = np.array([p.node(p.nx, j, k) for k in range(nz + 1) for j in range(nz + 1)], np.int32)
    p.loads = [(tip, np.ones(tip.size), np.array([0, 0, -1000], np.float32))]
    return p


def ablation_cantilever(nx=40, ny=20, nz=4, emin_ratio=1e-6, e0=3000.0):
    """CreateCantileverModel (Optimization/OptimizationTests.cs:389-423): x = 0 clamped; load
    (0, 0, -1000) over the bottom tip edge (nx, j, 0)."""
    p = Problem(nx, ny, nz, e0=e0, nu=0.3, emin_ratio=emin_ratio, name=f"cantilever {nx}x{ny}x{nz}")
    _clamped_x0(p)
    tip = np.array([p.node(nx, j, 0) for j in range(ny + 1)], np.int32)
    p.loads = [(tip, np.ones(tip.size), np.array([0, 0, -1000], np.float32))]
    return p


def tiny_model(nx=6, ny=4, nz=3, e0=3000.0, emin_ratio=1e-6):
    """CreateTinyTestModel (Optimization/OptimizationTests.cs:18-66): two point loads."""
    p = Problem(nx, ny, nz, e0=e0, nu=0.3, emin_ratio=emin_ratio, name="tiny")
    _clamped_x0(p)
    p.loads = [(np.array([p.node(nx, 0, 0)], np.int32), np.ones(1), np.array([0, 0, -100], np.float32)),
               (np.array([p.node(nx, ny, nz)], np.int32), np.ones(1), np.array([50, -30, 0], np.float32))]
    return p


def l_shape_domain(nx=8, ny=6, nz=5) -> np.ndarray:
    """T2's L-shaped domain (Optimization/OptimizationTests.cs:216-236)."""
    d = np.zeros((nz, ny, nx), np.uint8)
    ix = np.arange(nx)[None, None, :]
    iy = np.arange(ny)[None, :, None]
    d[:] = ((ix < nx // 2) | (iy < ny // 2)).astype(np.uint8)
    return d.reshape(-1)


# --------------------------------------------------------------------------------------------
# Serving stream
# --------------------------------------------------------------------------------------------
@dataclass
class StreamJob:
    kind: str  # "fea" or "simp"
    problem: Problem
    arrival_s: float
    volume_fraction: float = 0.3
    rmin: float = 1.5
    simp_iterations: int = 40


_FACES = ("-x", "+x", "-y", "+y", "-z", "+z")


def _face_nodes(p: Problem, face: str) -> np.ndarray:
    nx, ny, nz = p.nx, p.ny, p.nz
    if face in ("-x", "+x"):
        i = 0 if face == "-x" else nx
        return np.array([p.node(i, j, k) for k in range(nz + 1) for j in range(ny + 1)], np.int32)
    if face in ("-y", "+y"):
        j = 0 if face == "-y" else ny
        return np.array([p.node(i, j, k) for k in range(nz + 1) for i in range(nx + 1)], np.int32)
    k = 0 if face == "-z" else nz
    return np.array([p.node(i, j, k) for j in range(ny + 1) for i in range(nx + 1)], np.int32)


def _patch_on_face(p: Problem, face: str, rng: SplitMix64, size: int = 4) -> np.ndarray:
    axis = face[1]
    dims = {"x": (p.ny, p.nz), "y": (p.nx, p.nz), "z": (p.nx, p.ny)}[axis]
    a0 = int(rng.integers(1, max(1, dims[0] - size + 2))[0])
    b0 = int(rng.integers(1, max(1, dims[1] - size + 2))[0])
    fix = 0 if face[0] == "-" else {"x": p.nx, "y": p.ny, "z": p.nz}[axis]
    nodes = []
    for b in range(size):
        for a in range(size):
            u, v = min(a0 + a, dims[0]), min(b0 + b, dims[1])
            if axis == "x":
                nodes.append(p.node(fix, u, v))
            elif axis == "y":
                nodes.append(p.node(u, fix, v))
            else:
                nodes.append(p.node(u, v, fix))
    return np.unique(np.array(nodes, np.int32))


def stream_problem(rng: SplitMix64, min_el: int, max_el: int) -> Problem:
    """A full box grid with log-uniform element count, aspect 1:1:1, 2:1:1 or 3:1:1, one random face
    clamped and 1-4 random 4x4-node patch loads with N(0, I3) forces on the other faces."""
    n = float(np.exp(np.log(min_el) + (np.log(max_el) - np.log(min_el)) * rng.uniform(1)[0]))
    aspect = 1 + int(rng.integers(1, 3)[0])
    s = max(2, int(round((n / aspect) ** (1.0 / 3.0))))
    p = Problem(aspect * s, s, s, name=f"stream box {aspect * s}x{s}x{s}")
    fixed_face = _FACES[int(rng.integers(1, 6)[0])]
    p.fixed_dofs = fix_all(_face_nodes(p, fixed_face))
    others = [f for f in _FACES if f != fixed_face]
    loads = []
    for _ in range(1 + int(rng.integers(1, 4)[0])):
        face = others[int(rng.integers(1, len(others))[0])]
        nodes = _patch_on_face(p, face, rng)
        loads.append((nodes, np.ones(nodes.size), rng.normal(3).astype(np.float32)))
    p.loads = loads
    return p


def serving_stream(n_jobs: int = 512, seed: int = 5, min_el: int = 20_000, max_el: int = 250_000,
                   rate_hz: float = 1.0, simp_iterations: int = 40):
    """Jobs on fresh full-box grids, half single FEA solves and half SIMP runs (volume fraction U[0.2, 0.4],
    rmin U[1.5, 2.0]), with Poisson arrivals at rate_hz."""
    rng = SplitMix64(seed)
    t = 0.0
    jobs = []
    for q in range(n_jobs):
        t += float(-np.log1p(-rng.uniform(1)[0]) / rate_hz)
        p = stream_problem(rng, min_el, max_el)
        kind = "fea" if q % 2 == 0 else "simp"
        vf = 0.2 + 0.2 * float(rng.uniform(1)[0])
        rmin = 1.5 + 0.5 * float(rng.uniform(1)[0])
        jobs.append(StreamJob(kind, p, t, vf, rmin, simp_iterations))
    return jobs
