"""The OC update that allocates nothing (mycel_oc_update_noalloc_d), against mycel_oc_update.

Both run the same kernels; the no-allocation form keeps its scratch in a fixed device array, so two calls in a
row (the second on a smaller design vector) must each match the allocating entry point bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from mycel_amd import _lib as L
from mycel_amd import workloads as W

pytestmark = pytest.mark.gpu


def _case(n, seed):
    rng = W.SplitMix64(seed)
    rho = 0.25 + 0.1 * rng.uniform(n)
    dc = -(1e-3 + rng.uniform(n)) * 10.0 ** (-2.0 * rng.uniform(n))
    dc[rng.integers(max(1, n // 50), n)] = 0.0  # some non-negative sensitivities take the lower bound
    s = 0.5 + rng.uniform(n)
    passive = (rng.uniform(n) < 0.05).astype(np.uint8)
    return rho, dc, s, passive


def _alloc_update(lib, rho, dc, s, passive, vf, lam0):
    out = np.empty_like(rho)
    lam = C.c_double()
    L.check(lib.mycel_oc_update(L.ptr(rho), L.ptr(dc), L.ptr(s), L.ptr(passive), L.ptr(out), rho.size, vf, 0.2, 1e-3,
                                lam0, 60, 1e-8, C.byref(lam)))
    return out, lam.value


def _noalloc_update(lib, rho, dc, s, passive, vf, lam0, in_place=False):
    dev = torch.device("cuda:0")
    t_rho = torch.from_numpy(rho).to(dev)
    t_dc = torch.from_numpy(dc).to(dev)
    t_s = torch.from_numpy(s).to(dev)
    t_p = torch.from_numpy(passive).to(dev)
    t_out = t_rho if in_place else torch.full_like(t_rho, float("nan"))
    torch.cuda.synchronize()
    lam = C.c_double()
    L.check(lib.mycel_oc_update_noalloc_d(L.device_ptr(t_rho), L.device_ptr(t_dc), L.device_ptr(t_s),
                                          L.device_ptr(t_p), L.device_ptr(t_out), rho.size, vf, 0.2, 1e-3, lam0, 60,
                                          1e-8, None, C.byref(lam)))
    return t_out.cpu().numpy(), lam.value


def test_noalloc_matches_allocating_entry_point(gpu):
    torch.cuda.set_device(0)
    # Two calls in a row, the second smaller; the second also writes its result over its input.
    for n, seed, lam0, in_place in ((300_001, 11, 1.0, False), (40_000, 12, 3.7e-4, True)):
        rho, dc, s, passive = _case(n, seed)
        ref, lam_ref = _alloc_update(gpu, rho, dc, s, passive, 0.3, lam0)
        got, lam = _noalloc_update(gpu, rho.copy(), dc, s, passive, 0.3, lam0, in_place=in_place)
        print(f"n={n}: lambda {lam:.17g} vs {lam_ref:.17g}", flush=True)
        assert 1e-6 < lam_ref < 1e6  # reached by bisection, not at a bracket limit
        assert lam == lam_ref
        assert np.array_equal(got, ref)


def test_noalloc_rejects_oversized_and_null(gpu):
    lam = C.c_double()
    st = gpu.mycel_oc_update_noalloc_d(C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), None, C.c_void_p(16),
                                       L.OC_NOALLOC_MAX_ELEMENTS + 1, 0.3, 0.2, 1e-3, 1.0, 60, 1e-8, None,
                                       C.byref(lam))
    assert st == L.E_OUT_OF_RANGE
    assert b"MYCEL_OC_NOALLOC_MAX_ELEMENTS" in gpu.mycel_last_error()
    st = gpu.mycel_oc_update_noalloc_d(None, None, None, None, None, 10, 0.3, 0.2, 1e-3, 1.0, 60, 1e-8, None,
                                       C.byref(lam))
    assert st == L.E_ARGUMENT_NULL
