"""hb_eval_metrics_noalloc (no device allocation; every intermediate buffer in one
fixed device array) against hb_eval_metrics on the same device arrays: identical
MAE, MAPE, Spearman rho and top-k overlap on consecutive calls, the second smaller
than the first, so state left in the fixed array by one call cannot leak into the
next."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return C.CDLL(name)
        except OSError:
            continue
    pytest.skip("libamdhip64 not found")


class DeviceArrays:
    """Test-owned device copies of host float64 arrays."""

    def __init__(self, hip, *arrays):
        self.hip = hip
        self.ptrs = []
        for a in arrays:
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
            self.ptrs.append(p)
            assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # H2D

    def f64(self, i):
        return C.cast(self.ptrs[i], C.POINTER(C.c_double))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        for p in self.ptrs:
            if p.value:
                self.hip.hipFree(p)


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    b = rng.random(n) * 10.0
    b[rng.integers(0, n, n // 50)] = 0.0                      # MAPE skips these
    a = b + rng.normal(0.0, 0.5, n)
    a[rng.integers(0, n, n // 20)] = 1.25                     # ties
    a[: n // 100] = -0.0                                       # -0.0 equals +0.0
    return a, b


def _both(L, lib, dev, n, top):
    ref = np.zeros(4, np.float64)
    got = np.full(4, np.nan, np.float64)
    lib.check(L.hb_eval_metrics(0, dev.f64(0), dev.f64(1), n, top, lib.ptr(ref, C.c_double)))
    print(f"[case] noalloc n={n} top={top}", flush=True)
    lib.check(L.hb_eval_metrics_noalloc(0, dev.f64(0), dev.f64(1), n, top, lib.ptr(got, C.c_double)))
    return ref, got


def test_noalloc_matches_allocating_entry_point_on_consecutive_calls():
    from sparkx_amd import _lib
    L = _lib.load()
    hip = _hip()
    for n, seed, top in ((300_000, 1, 0.01), (5_000, 2, 0.05), (1 << 19, 3, 0.01), (4_000, 4, 1.0)):
        a, b = _inputs(n, seed)
        with DeviceArrays(hip, a, b) as dev:
            ref, got = _both(L, _lib, dev, n, top)
        assert np.array_equal(ref, got), (n, ref, got)
    # the same results on a repeated call (nothing carried over from the larger calls)
    a, b = _inputs(5_000, 2)
    with DeviceArrays(hip, a, b) as dev:
        ref, got = _both(L, _lib, dev, 5_000, 0.05)
        ref2, got2 = _both(L, _lib, dev, 5_000, 0.05)
    assert np.array_equal(ref, got) and np.array_equal(got, got2) and np.array_equal(ref, ref2)


def test_noalloc_errors():
    from sparkx_amd import _lib
    L = _lib.load()
    hip = _hip()
    out = np.zeros(4, np.float64)
    a, b = _inputs(1000, 5)
    # host arrays are refused (this entry point reads its inputs in place on the device)
    rc = L.hb_eval_metrics_noalloc(0, _lib.ptr(a, C.c_double), _lib.ptr(b, C.c_double), a.size, 0.01,
                                   _lib.ptr(out, C.c_double))
    assert rc == _lib.HB_EINVAL and b"device" in L.hb_last_error()
    # more than the fixed array holds
    big = np.zeros(1 << 22, np.float64)
    with DeviceArrays(hip, big, big) as dev:
        rc = L.hb_eval_metrics_noalloc(0, dev.f64(0), dev.f64(1), big.size, 0.01, _lib.ptr(out, C.c_double))
        assert rc == _lib.HB_ERANGE and b"scratch" in L.hb_last_error()
    # a NaN is reported; the next call is unaffected by the NaN slot it left behind
    a2 = a.copy()
    a2[17] = np.nan
    with DeviceArrays(hip, a2, b, a) as dev:
        rc = L.hb_eval_metrics_noalloc(0, dev.f64(0), dev.f64(1), a.size, 0.01, _lib.ptr(out, C.c_double))
        assert rc == _lib.HB_EINVAL and b"NaN at index 17 of argument 1" in L.hb_last_error()
        ref = np.zeros(4, np.float64)
        _lib.check(L.hb_eval_metrics(0, dev.f64(2), dev.f64(1), a.size, 0.01, _lib.ptr(ref, C.c_double)))
        _lib.check(L.hb_eval_metrics_noalloc(0, dev.f64(2), dev.f64(1), a.size, 0.01, _lib.ptr(out, C.c_double)))
        assert np.array_equal(out, ref)
    # bad arguments
    assert L.hb_eval_metrics_noalloc(0, None, None, 10, 0.01, _lib.ptr(out, C.c_double)) == _lib.HB_EINVAL
    assert L.hb_eval_metrics_noalloc(99, None, None, 10, 0.01, _lib.ptr(out, C.c_double)) == _lib.HB_EINVAL
