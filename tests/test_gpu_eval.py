"""GPU parity of the evaluation metrics (hb_eval_ranks / hb_eval_metrics) against
the oracle's utils.rs restatement (src/utils.rs:57-100).

Ranks and the top-k overlap are index work: bit-identical.  Spearman's squared rank
differences are summed exactly on the device, so rho must equal the reference's
formula over the exact integer ssd bit for bit, and the oracle's sequential-sum rho
within 1e-12.  MAE / MAPE are f64 sums in another order: within 1e-12 relative of
the oracle's sequential sums."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def metrics():
    from sparkx_amd import _lib, metrics as M
    _lib.load()
    return M


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(-4, 5, n).astype(np.float64)  # heavy ties
    a[rng.random(n) < 0.05] = -0.0
    a[rng.random(n) < 0.01] = np.inf
    a[rng.random(n) < 0.01] = -np.inf
    b = a + rng.normal(scale=2.0, size=n)
    b[rng.random(n) < 0.05] = 0.0  # MAPE skips these
    return a, b


def _exact_rho(ra, rb):
    d = ra.astype(np.int64) - rb.astype(np.int64)
    assert d.size <= 1 << 21  # d^2 <= 2^42, so the int64 sum below is exact
    ssd = int(np.sum(d * d))
    n = float(d.size)
    return 1.0 - (6.0 * float(ssd)) / (n * max(n * n - 1.0, 1.0))  # float(int): one rounding


@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 1000, 65537, (1 << 20) + 3])
def test_ranks_and_metrics_match_oracle(metrics, oracle, n):
    a, b = _inputs(n, n)
    say("eval n", n)
    ra, rb = metrics.ranks(a), metrics.ranks(b)
    ra_ref, rb_ref = oracle.ranks(a), oracle.ranks(b)
    assert np.array_equal(ra.astype(np.uint64), ra_ref)
    assert np.array_equal(rb.astype(np.uint64), rb_ref)
    m = metrics.evaluate(a, b, 0.01)
    assert m["spearman"] == _exact_rho(ra_ref, rb_ref)
    # the reference's sequential sum of squared rank differences rounds once per term
    # after partial sums pass 2^53 (here from n ~ 2e5 on): within 1e-9 relative of (1 - rho)
    rho_ref = oracle.spearman(a, b)
    assert abs(m["spearman"] - rho_ref) <= 1e-9 * max(abs(1.0 - rho_ref), 1e-300)
    assert (n < 200000) <= (m["spearman"] == rho_ref)  # exact sums agree bit for bit
    # MAE / MAPE over finite values (inf - inf is NaN on both sides)
    af = np.where(np.isinf(a), np.sign(a) * 1e6, a)
    bf = np.where(np.isinf(b), np.sign(b) * 1e6, b)
    me = metrics.evaluate(af, bf, 0.01)
    for key, ref in (("mae", oracle.mae(af, bf)), ("mape", oracle.mape(af, bf))):
        assert abs(me[key] - ref) <= 1e-12 * max(1.0, abs(ref)), (key, me[key], ref)
    for frac in (0.01, 0.25, 1.0):
        got = metrics.top_overlap(a, b, frac)
        assert got == oracle.top_overlap(a, b, frac), frac


def test_metrics_closed_forms(metrics):
    v = np.arange(1000, dtype=np.float64)
    m = metrics.evaluate(v, v, 0.1)
    assert m["spearman"] == 1.0 and m["mae"] == 0.0 and m["top_overlap"] == 1.0
    m = metrics.evaluate(v, -v, 0.1)
    assert m["spearman"] == -1.0 and m["top_overlap"] == 0.0
    assert m["mae"] == float(np.sum(2 * v)) / 1000
    assert metrics.evaluate([7.0], [3.0])["spearman"] == 1.0  # denominator max(n^2 - 1, 1)
    # ties keep index order: constant input ranks are the identity
    assert np.array_equal(metrics.ranks(np.zeros(300)), np.arange(300))
    # -0.0 equals +0.0
    assert np.array_equal(metrics.ranks([0.0, -0.0, 0.0, -0.0]), [0, 1, 2, 3])


def _hip():
    """The HIP runtime the library is linked against, for device buffers in tests."""
    import ctypes.util
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return C.CDLL(name)
        except OSError:
            continue
    pytest.skip("libamdhip64 not found")


def test_device_pointers(metrics, oracle):
    """Device arrays in, host results out (ranks also to a device array)."""
    from sparkx_amd import _lib
    hip = _hip()
    n = 5000
    a, b = _inputs(n, 11)
    ptrs = [C.c_void_p() for _ in range(3)]
    try:
        for pp, size in zip(ptrs, (n * 8, n * 8, n * 4)):
            assert hip.hipMalloc(C.byref(pp), C.c_size_t(size)) == 0
        assert hip.hipMemcpy(ptrs[0], a.ctypes.data_as(C.c_void_p), C.c_size_t(n * 8), 1) == 0  # H2D
        assert hip.hipMemcpy(ptrs[1], b.ctypes.data_as(C.c_void_p), C.c_size_t(n * 8), 1) == 0
        L = _lib.load()
        f64p = C.POINTER(C.c_double)
        out = np.zeros(4, np.float64)
        _lib.check(L.hb_eval_metrics(0, C.cast(ptrs[0], f64p), C.cast(ptrs[1], f64p), n, 0.01,
                                     _lib.ptr(out, C.c_double)))
        _lib.check(L.hb_eval_ranks(0, C.cast(ptrs[0], f64p), n, C.cast(ptrs[2], C.POINTER(C.c_uint32))))
        r = np.zeros(n, np.uint32)
        assert hip.hipMemcpy(r.ctypes.data_as(C.c_void_p), ptrs[2], C.c_size_t(n * 4), 2) == 0  # D2H
    finally:
        for pp in ptrs:
            if pp.value:
                hip.hipFree(pp)
    assert np.array_equal(r.astype(np.uint64), oracle.ranks(a))
    host = metrics.evaluate(a, b, 0.01)
    assert out[2] == host["spearman"]
    assert out[3] == oracle.top_overlap(a, b, 0.01) == host["top_overlap"]


def test_errors(metrics):
    from sparkx_amd._lib import HyperBallError
    a = np.ones(10)
    bad = a.copy()
    bad[7] = np.nan
    with pytest.raises(HyperBallError, match="NaN at index 7 of argument 2"):
        metrics.evaluate(a, bad)
    with pytest.raises(HyperBallError, match="NaN at index 7 of argument 1"):
        metrics.ranks(bad)
    with pytest.raises(HyperBallError, match="not in 1"):
        metrics.ranks(np.zeros(0))
    with pytest.raises(HyperBallError, match="top_frac"):
        metrics.evaluate(a, a, 0.0)
    with pytest.raises(HyperBallError, match="top_frac"):
        metrics.evaluate(a, a, 1.5)
    with pytest.raises(HyperBallError, match="bad device"):
        metrics.evaluate(a, a, 0.5, device=99)
    with pytest.raises(ValueError):
        metrics.evaluate(a, np.ones(9))
