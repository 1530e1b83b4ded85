"""Shared helpers for the tests: oracle models from problems, and independent numpy restatements used
to pin the oracle (dense assembly, brute-force filter)."""
from __future__ import annotations

import numpy as np

import oracle as O
from mycel_amd import workloads as W

CX = np.array([0, 1, 1, 0, 0, 1, 1, 0])
CY = np.array([0, 0, 1, 1, 0, 0, 1, 1])
CZ = np.array([0, 0, 0, 0, 1, 1, 1, 1])


def oracle_model(p: W.Problem, density=None):
    return O.Model(p.nx, p.ny, p.nz, p.domain, p.fixed_dofs, p.h, p.e0, p.nu, p.emin_ratio, p.penal,
                   p.density if density is None else density)


def oracle_forces(p: W.Problem):
    return O.assemble_forces(p.n_nodes, p.loads) if p.loads else np.zeros(p.n_dofs)


def element_dofs(p: W.Problem, e: int) -> np.ndarray:
    ix, rem = e % p.nx, e // p.nx
    iy, iz = rem % p.ny, rem // p.ny
    nodes = [(ix + CX[a]) + (p.nx + 1) * ((iy + CY[a]) + (p.ny + 1) * (iz + CZ[a])) for a in range(8)]
    return np.array([3 * n + c for n in nodes for c in range(3)])


def dense_k(p: W.Problem, ke: np.ndarray, density=None) -> np.ndarray:
    """Dense global K assembled element by element in numpy (independent of the oracle's colouring)."""
    rho = p.density_box() if density is None else density
    emin = p.e0 * p.emin_ratio
    K = np.zeros((p.n_dofs, p.n_dofs))
    for e in np.flatnonzero(p.domain):
        d = element_dofs(p, e)
        E = emin + rho[e] ** p.penal * (p.e0 - emin)
        K[np.ix_(d, d)] += E * ke
    return K


def brute_filter(p: W.Problem, rmin: float):
    """Dense W from the definition H = max(0, r - dist), row-normalised (DensityFilter.cs:9-14)."""
    act = p.active
    coords = np.array([(e % p.nx, (e // p.nx) % p.ny, e // (p.nx * p.ny)) for e in act], float)
    d = np.sqrt(((coords[:, None, :] - coords[None, :, :]) ** 2).sum(-1))
    H = np.where(d < rmin, rmin - d, 0.0)
    return H / H.sum(1, keepdims=True)


def rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def random_field(seed, n, lo=0.0, hi=1.0):
    return lo + (hi - lo) * W.SplitMix64(seed).uniform(n)
