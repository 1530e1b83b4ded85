"""C ABI of libmycel_hip.so without a GPU: the library loads, exports every symbol the header declares,
and its host-side set-up routines (Ke, force assembly, passive region, argument validation) agree
with the oracle. No call here touches the device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from mycel_amd import _lib as L
from mycel_amd import workloads as W

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "mycel_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"MYCEL_API\s+[\w\s\*]*?\b(mycel_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), "ctypes signatures out of sync with the header"


def test_version_and_error_string():
    lib = L.load()
    assert lib.mycel_version() >= 10000
    st = lib.mycel_hex8_ke(0.0, 0.3, L.ptr(np.zeros(576)))
    assert st == L.E_OUT_OF_RANGE
    assert b"positive" in lib.mycel_last_error()


def test_ke_bitwise_equals_oracle(oracle):
    lib = L.load()
    for h, nu in ((1.0, 0.3), (2.5, 0.3), (0.5, 0.28), (10.0, 0.0)):
        ke = np.zeros(576)
        L.check(lib.mycel_hex8_ke(h, nu, L.ptr(ke)))
        assert np.array_equal(ke, oracle.hex8_ke(h, nu).ravel())
        bc = np.zeros(144)
        L.check(lib.mycel_hex8_bc(h, L.ptr(bc)))
        assert np.array_equal(bc, oracle.hex8_bc(h).ravel())


def test_assemble_forces_equals_oracle(oracle):
    p = W.matvec_pcg_workstation(seed=5, nx=20, ny=12, nz=12)
    fm = p.fe_model()
    f = fm.AssembleForces()
    assert np.array_equal(f, oracle.assemble_forces(p.n_nodes, p.loads))
    p2, tip, areas = W.bar_model()
    p2.loads = [(tip, areas, np.array([3, -2, 1000], np.float32)), (tip[:3], np.zeros(3), np.array([1, 1, 1], np.float32))]
    assert np.array_equal(p2.fe_model().AssembleForces(), oracle.assemble_forces(p2.n_nodes, p2.loads))


def test_assemble_forces_errors():
    from mycel_amd import FeMaterial, FeModel, GridSpec, InvalidOperationError, LoadCase, SurfacePatch

    g = GridSpec(2, 2, 2)
    m = FeModel(g, np.ones(8), FeMaterial(1, 0.3), [0, 1, 2], [LoadCase(SurfacePatch([], []), [0, 0, 1])])
    with pytest.raises(InvalidOperationError):
        m.AssembleForces()
    m = FeModel(g, np.ones(8), FeMaterial(1, 0.3), [0, 1, 2], [LoadCase(SurfacePatch([999], [1.0]), [0, 0, 1])])
    with pytest.raises(InvalidOperationError):
        m.AssembleForces()


def test_passive_region_equals_oracle(oracle):
    lib = L.load()
    g = L.Grid(9, 7, 5, 1.0)
    rng = W.SplitMix64(3)
    nodes = L.i32(rng.integers(40, 10 * 8 * 6))
    for r in (1, 2):
        m = np.zeros(9 * 7 * 5, np.uint8)
        L.check(lib.mycel_passive_region(C.byref(g), L.ptr(nodes), nodes.size, r, L.ptr(m)))
        assert np.array_equal(m, oracle.passive_region(9, 7, 5, nodes, r))


def test_femodel_validation_matches_reference():
    from mycel_amd import ArgumentError, ArgumentOutOfRangeError, FeMaterial, FeModel, GridSpec

    g = GridSpec(3, 2, 2)
    with pytest.raises(ArgumentError):
        FeModel(g, np.ones(5), FeMaterial(1, 0.3), [])
    with pytest.raises(ArgumentOutOfRangeError):
        FeModel(g, np.ones(12), FeMaterial(1, 0.3), [10_000])
    with pytest.raises(ArgumentOutOfRangeError):
        GridSpec(0, 1, 1)
    m = FeModel(g, np.ones(12), FeMaterial(1, 0.3), FeModel.FixDofs([0, 1], True, False, True))
    assert list(m.FixedDofs) == [0, 2, 3, 5]
    assert m.ActiveFixedDofCount == 4


def test_workload_counts():
    p = W.matvec_pcg_workstation()
    fm = p.fe_model()
    assert p.n_elem == 128_000 and p.n_dofs == 408_483
    assert fm.ActiveElementCount == 110_080
    assert fm.ActiveNodeCount == 119_720
    c1 = W.cantilever_simp()
    assert c1.n_elem == 24_000 and c1.n_dofs == 80_703 and c1.fixed_dofs.size == 1_323
    assert np.isclose(c1.fe_model().AssembleForces().sum(), -1.0, rtol=0, atol=1e-15)


def test_splitmix64_reference_vector():
    # First outputs of SplitMix64 seeded with 0 (the generator's published test sequence).
    z = W.SplitMix64(0).next_u64(3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
