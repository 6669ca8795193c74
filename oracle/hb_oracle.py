"""ctypes front-end of the CPU oracle (libhb_oracle.so).

TEST INFRASTRUCTURE ONLY: the tests check the HIP path against it.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.

Each function mirrors a reference routine (paths relative to the reference
repo root):
  hyperball()          src/centralities/hyper_ball.rs:702-736 and :754-874
  centrality()         src/centralities/hyper_ball.rs:434-555
  exact_closeness()    src/trails/bfs.rs:199-234, src/test_common.rs:217-227
  exact_bfs()          src/trails/bfs.rs:199-234, src/centralities/exact.rs:386-396
  exact_centrality()   src/centralities/exact.rs:138-201,357-441,560-619
  mae/mape/spearman    src/utils.rs:57-100
  dbg_mtx_csr()        gen_dbg_mtx.py:4-22 + parser_mtx/producers.rs:228-234
                       + impl_sort.rs:10-51
  walk_counts(), walk_inputs()
                       no reference counterpart (inputs of the learned surrogate);
                       pinned by brute-force walk enumeration, exact sparse powers
                       A^k deg and a closed-form saturating case (tests/test_cpu_oracle.py)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# HB_ORACLE_LIB selects another build of the oracle (the Makefile's sanitizer builds)
_LIB_PATH = os.environ.get("HB_ORACLE_LIB") or os.path.join(_HERE, "libhb_oracle.so")


class Variant(C.Structure):
    _fields_ = [
        ("hash", C.c_int),
        ("index", C.c_int),
        ("lc_rule", C.c_int),
        ("lc_calc", C.c_int),
        ("sum_order", C.c_int),
        ("alpha", C.c_int),
        ("final_div", C.c_int),
        ("bias", C.c_int),
    ]

    def astuple(self):
        return tuple(getattr(self, f) for f, _ in self._fields_)

    def __repr__(self):
        return "Variant(" + ", ".join(f"{f}={getattr(self, f)}" for f, _ in self._fields_) + ")"


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("final_parity_latest", C.c_int),
        ("increases", C.c_uint64 * 1024),
        ("init_seconds", C.c_double),
        ("loop_seconds", C.c_double),
        ("regime", C.c_uint64 * 3),
        ("iter_seconds", C.c_double * 1024),
    ]


ITER_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_uint8),
                      C.POINTER(C.c_double), C.POINTER(C.c_double))

class MtxInfo(C.Structure):
    _fields_ = [("n", C.c_uint64), ("nnz", C.c_uint64), ("body", C.c_uint64), ("body_line", C.c_uint64),
                ("entries", C.c_uint64), ("arcs", C.c_uint64), ("err_line", C.c_uint64),
                ("symmetric", C.c_int)]


MTX_ERRORS = {-1: "bad header", -2: "malformed header", -3: "unsupported format", -4: "no size line",
              -5: "bad size line", -6: "not square", -7: "bad entry", -8: "zero index",
              -9: "out of bounds", -10: "out of memory"}


class MtxError(ValueError):
    def __init__(self, code: int, line: int):
        super().__init__(f"{MTX_ERRORS.get(code, code)} (line {line})")
        self.code = code
        self.line = line


_lib = None


def build() -> str:
    """Compiles the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        u64p = C.POINTER(C.c_uint64)
        u32p = C.POINTER(C.c_uint32)
        u8p = C.POINTER(C.c_uint8)
        f64p = C.POINTER(C.c_double)
        f32p = C.POINTER(C.c_float)
        vp = C.POINTER(Variant)
        L.hbo_siphash.restype = C.c_uint64
        L.hbo_siphash.argtypes = [C.c_uint64, C.c_uint64, u8p, C.c_size_t, C.c_int, C.c_int]
        L.hbo_hash_usize.restype = C.c_uint64
        L.hbo_hash_usize.argtypes = [vp, C.c_uint64]
        L.hbo_index_rank.restype = C.c_uint32
        L.hbo_index_rank.argtypes = [vp, C.c_int, C.c_uint64, u8p]
        L.hbo_estimate_u8.restype = C.c_float
        L.hbo_estimate_u8.argtypes = [vp, C.c_int, u8p]
        L.hbo_estimate_packed.restype = C.c_float
        L.hbo_estimate_packed.argtypes = [vp, C.c_int, u32p]
        L.hbo_knn_bias.restype = C.c_double
        L.hbo_knn_bias.argtypes = [C.c_int, C.c_double]
        L.hbo_knn_bias_many.argtypes = [C.c_int, f32p, C.c_uint64, f64p]
        L.hbo_node_sample.restype = C.c_int
        L.hbo_node_sample.argtypes = [u64p, u32p, C.c_uint64, C.c_int, vp, C.c_int, C.c_uint64, C.c_int, f64p,
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.hbo_pack6.argtypes = [C.c_int, u8p, u32p]
        L.hbo_unpack6.argtypes = [C.c_int, u32p, u8p]
        L.hbo_hyperball.restype = C.c_int
        L.hbo_hyperball.argtypes = [u64p, u32p, C.c_uint64, C.c_int, vp, C.c_int, C.c_int,
                                    C.c_int, u8p, u8p, f64p, f64p, C.POINTER(Stats),
                                    ITER_CB, C.c_void_p]
        L.hbo_estimates.argtypes = [vp, C.c_int, C.c_uint64, u8p, f32p]
        L.hbo_centrality.argtypes = [C.c_int, C.c_int, C.c_uint64, f32p, f64p, f64p, f64p]
        L.hbo_exact_closeness.argtypes = [u64p, u32p, C.c_uint64, C.c_uint64, C.c_uint64, f64p]
        L.hbo_exact_bfs.argtypes = [u64p, u32p, C.c_uint64, C.c_uint64, C.c_uint64, u64p, f64p, f64p]
        for f in ("hbo_mae", "hbo_mape", "hbo_spearman"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [f64p, f64p, C.c_uint64]
        L.hbo_ranks.argtypes = [f64p, C.c_uint64, u64p]
        L.hbo_dbg_mtx_csr.restype = C.c_uint64
        L.hbo_dbg_mtx_csr.argtypes = [C.c_int, C.c_int, u64p, u32p]
        L.hbo_mtx_header.restype = C.c_int
        L.hbo_mtx_header.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(MtxInfo)]
        L.hbo_parse_mtx.restype = C.c_int
        L.hbo_parse_mtx.argtypes = [C.c_char_p, C.c_uint64, C.POINTER(MtxInfo), u64p, u32p]
        _lib = L
    return _lib


def _p(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def pinned() -> Variant:
    return Variant.in_dll(lib(), "HBO_VARIANT_PINNED")


def variant(**kw) -> Variant:
    v = Variant()
    base = pinned()
    for f, _ in Variant._fields_:
        setattr(v, f, kw.get(f, getattr(base, f)))
    return v


def siphash(msg: bytes, c_rounds=1, d_rounds=3, k0=0, k1=0) -> int:
    buf = np.frombuffer(msg, dtype=np.uint8).copy() if msg else np.zeros(1, np.uint8)
    return int(lib().hbo_siphash(k0, k1, _p(buf, C.c_uint8), len(msg), c_rounds, d_rounds))


def index_rank(p: int, x: int, v: Variant | None = None):
    v = v or pinned()
    r = C.c_uint8()
    idx = lib().hbo_index_rank(C.byref(v), p, x, C.byref(r))
    return int(idx), int(r.value)


def node_sample(offsets, nbrs, p: int, stride: int, reps: int = 3, packed: bool = True):
    """(seconds per repetition, sampled nodes, their arcs): one thread doing the first
    iteration's work for a pseudo-random sample of about one node in `stride`."""
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    nb = np.ascontiguousarray(nbrs, dtype=np.uint32)
    secs = np.zeros(reps, np.float64)
    nodes, arcs = C.c_uint64(), C.c_uint64()
    v = pinned()
    rc = lib().hbo_node_sample(_p(off, C.c_uint64), _p(nb, C.c_uint32), off.size - 1, p, C.byref(v), int(packed),
                               stride, reps, _p(secs, C.c_double), C.byref(nodes), C.byref(arcs))
    if rc != 0:
        raise RuntimeError(f"hbo_node_sample failed: {rc}")
    return secs, nodes.value, arcs.value


def knn_bias(p: int, e: np.ndarray) -> np.ndarray:
    """HLL++ EstimateBias (k = 6 nearest points, ties to the lower one) of each f32
    estimate in e."""
    e = np.ascontiguousarray(e, dtype=np.float32)
    out = np.empty(e.size, np.float64)
    lib().hbo_knn_bias_many(p, _p(e, C.c_float), e.size, _p(out, C.c_double))
    return out


def estimate(p: int, regs: np.ndarray, v: Variant | None = None) -> np.float32:
    v = v or pinned()
    regs = np.ascontiguousarray(regs, dtype=np.uint8)
    return np.float32(lib().hbo_estimate_u8(C.byref(v), p, _p(regs, C.c_uint8)))


def estimate_packed(p: int, words: np.ndarray, v: Variant | None = None) -> np.float32:
    v = v or pinned()
    words = np.ascontiguousarray(words, dtype=np.uint32)
    return np.float32(lib().hbo_estimate_packed(C.byref(v), p, _p(words, C.c_uint32)))


def pack6(p: int, regs: np.ndarray) -> np.ndarray:
    m = 1 << p
    out = np.zeros((m + 4) // 5, np.uint32)
    regs = np.ascontiguousarray(regs, dtype=np.uint8)
    lib().hbo_pack6(p, _p(regs, C.c_uint8), _p(out, C.c_uint32))
    return out


def estimates(p: int, regs: np.ndarray, v: Variant | None = None) -> np.ndarray:
    v = v or pinned()
    regs = np.ascontiguousarray(regs, dtype=np.uint8)
    n = regs.size >> p
    out = np.empty(n, np.float32)
    lib().hbo_estimates(C.byref(v), p, n, _p(regs, C.c_uint8), _p(out, C.c_float))
    return out


@dataclass
class HBResult:
    regs_ref: np.ndarray      # the buffer the reference leaves in its counters file
    regs_last: np.ndarray     # output of the last executed iteration
    dist: np.ndarray
    inv: np.ndarray
    iterations: int
    final_parity_latest: bool
    increases: np.ndarray
    init_seconds: float = 0.0
    loop_seconds: float = 0.0
    # estimates after the union over all (node, iteration): linear counting, raw with
    # the HLL++ bias subtracted, raw as is
    regime: tuple = (0, 0, 0)
    iter_seconds: tuple = ()  # wall time of each iteration


def hyperball(offsets: np.ndarray, nbrs: np.ndarray, p: int, *, v: Variant | None = None,
              threads: int = 1, packed: bool = False, max_iters: int = 0,
              on_iter=None) -> HBResult:
    """Runs the restated HyperBall; on_iter(t, regs_out, dist, inv) gets copies."""
    v = v or pinned()
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    nb = np.ascontiguousarray(nbrs, dtype=np.uint32)
    n = off.size - 1
    m = 1 << p
    regs_ref = np.zeros(n * m, np.uint8)
    regs_last = np.zeros(n * m, np.uint8)
    dist = np.zeros(n, np.float64)
    inv = np.zeros(n, np.float64)
    st = Stats()
    if on_iter is not None:
        def _cb(_u, t, regs, d, i):
            on_iter(t, np.ctypeslib.as_array(regs, shape=(n * m,)).copy(),
                    np.ctypeslib.as_array(d, shape=(n,)).copy(),
                    np.ctypeslib.as_array(i, shape=(n,)).copy())
        cb = ITER_CB(_cb)
    else:
        cb = ITER_CB()
    nbp = _p(nb, C.c_uint32) if nb.size else C.POINTER(C.c_uint32)()
    rc = lib().hbo_hyperball(_p(off, C.c_uint64), nbp, n, p, C.byref(v), threads,
                             int(packed), max_iters, _p(regs_ref, C.c_uint8),
                             _p(regs_last, C.c_uint8), _p(dist, C.c_double),
                             _p(inv, C.c_double), C.byref(st), cb, None)
    if rc != 0:
        raise RuntimeError(f"hbo_hyperball failed: {rc}")
    it = st.iterations
    return HBResult(regs_ref, regs_last, dist, inv, it, bool(st.final_parity_latest),
                    np.array(st.increases[:min(it, 1024)], dtype=np.uint64),
                    st.init_seconds, st.loop_seconds, tuple(int(x) for x in st.regime),
                    tuple(float(x) for x in st.iter_seconds[:min(it, 1024)]))


CLOSENESS, HARMONIC, LIN = 0, 1, 2


def centrality(kind: int, normalize: bool, est: np.ndarray, dist: np.ndarray,
               inv: np.ndarray) -> np.ndarray:
    est = np.ascontiguousarray(est, dtype=np.float32)
    dist = np.ascontiguousarray(dist, dtype=np.float64)
    inv = np.ascontiguousarray(inv, dtype=np.float64)
    out = np.empty(est.size, np.float64)
    lib().hbo_centrality(kind, int(normalize), est.size, _p(est, C.c_float),
                         _p(dist, C.c_double), _p(inv, C.c_double), _p(out, C.c_double))
    return out


def exact_closeness(offsets, nbrs, src_begin=0, src_end=None) -> np.ndarray:
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    nb = np.ascontiguousarray(nbrs, dtype=np.uint32)
    n = off.size - 1
    src_end = n if src_end is None else min(src_end, n)
    out = np.zeros(max(src_end - src_begin, 0), np.float64)
    lib().hbo_exact_closeness(_p(off, C.c_uint64), _p(nb, C.c_uint32), n, src_begin, src_end,
                              _p(out, C.c_double))
    return out


def exact_bfs(offsets, nbrs, src_begin=0, src_end=None, inv=True):
    """(reachable u64, total f64, inv_sum f64 or None) per source, bfs.rs:199-234 + exact.rs:386-396."""
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    nb = np.ascontiguousarray(nbrs, dtype=np.uint32)
    n = off.size - 1
    src_end = n if src_end is None else min(src_end, n)
    cnt = max(src_end - src_begin, 0)
    reach = np.zeros(cnt, np.uint64)
    total = np.zeros(cnt, np.float64)
    invs = np.zeros(cnt, np.float64) if inv else None
    lib().hbo_exact_bfs(_p(off, C.c_uint64), _p(nb, C.c_uint32), n, src_begin, src_end,
                        _p(reach, C.c_uint64), _p(total, C.c_double),
                        _p(invs, C.c_double) if inv else None)
    return reach, total, invs


def exact_centrality(kind, normalization, reach, total, inv_sum, n):
    """exact.rs formulas from the BFS sums; kind 0 closeness (:160-201), 1 harmonic
    (:379-441), 2 Lin (:610-619); normalization None / False / True as the reference's
    Option<bool> (:130-136); n = node count."""
    reach = np.asarray(reach, np.uint64)
    r = reach.astype(np.float64)
    total = np.asarray(total, np.float64)
    norm = float(max(n - 1, 0))
    with np.errstate(divide="ignore", invalid="ignore"):
        if kind == 0:
            zero = (reach <= 1) | (total == 0.0)
            val = total if normalization is None else (np.float64(norm) / total if normalization is False
                                                       else r / total)
        elif kind == 1:
            inv_sum = np.asarray(inv_sum, np.float64)
            zero = reach <= 1
            val = inv_sum if normalization is None else (np.float64(norm) / inv_sum if normalization is False
                                                         else r / inv_sum)
        elif kind == 2:
            zero = total == 0.0
            val = (r * r) / total
        else:
            raise ValueError(kind)
    return np.where(zero, 0.0, val)


def _metric(name, a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return float(getattr(lib(), name)(_p(a, C.c_double), _p(b, C.c_double), a.size))


def mae(a, b):
    return _metric("hbo_mae", a, b)


def mape(a, b):
    return _metric("hbo_mape", a, b)


def spearman(a, b):
    return _metric("hbo_spearman", a, b)


def ranks(v) -> np.ndarray:
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty(v.size, np.uint64)
    lib().hbo_ranks(_p(v, C.c_double), v.size, _p(out, C.c_uint64))
    return out


def top_overlap(pred, truth, frac: float = 0.01) -> float:
    """|top-k(pred) & top-k(truth)| / k, k = max(1, ceil(frac * n)); the top k of each
    array are the first k of its stable descending order (the ranks above)."""
    rp, rt = ranks(pred), ranks(truth)
    n = rt.size
    k = max(1, int(np.ceil(frac * n)))
    return float(np.count_nonzero((rp < k) & (rt < k))) / k


def dbg_mtx_csr(sigma: int, k: int):
    """CSR of gen_dbg_mtx.py(sigma, k) as the reference's MTX reader builds it."""
    n = sigma ** k
    off = np.zeros(n + 1, np.uint64)
    arcs = lib().hbo_dbg_mtx_csr(sigma, k, _p(off, C.c_uint64), C.POINTER(C.c_uint32)())
    nb = np.zeros(arcs, np.uint32)
    lib().hbo_dbg_mtx_csr(sigma, k, _p(off, C.c_uint64), _p(nb, C.c_uint32))
    return off, nb


def parse_mtx(text: bytes):
    """MatrixMarket text -> (offsets u64[n+1], neighbours u32[arcs], MtxInfo), following
    parser_mtx/{mod.rs,producers.rs} and impl_sort.rs; raises MtxError."""
    info = MtxInfo()
    rc = lib().hbo_parse_mtx(text, len(text), C.byref(info), None, None)
    if rc:
        raise MtxError(rc, info.err_line)
    off = np.zeros(info.n + 1, np.uint64)
    nb = np.zeros(max(info.arcs, 1), np.uint32)
    rc = lib().hbo_parse_mtx(text, len(text), C.byref(info), _p(off, C.c_uint64), _p(nb, C.c_uint32))
    if rc:
        raise MtxError(rc, info.err_line)
    return off, nb[:info.arcs], info


U64_MAX = (1 << 64) - 1


def walk_counts(offsets, nbrs):
    """(degree, S_1, S_2, S_3, M_1) as uint64 arrays: S_0 = degree, S_k(u) = sum of
    S_{k-1}(v) over u's arcs u -> v (walks of k + 1 arcs leaving u), saturating at
    2^64 - 1 (min(2^64 - 1, exact sum)); M_1(u) = largest degree among u's heads, 0
    for a node without arcs."""
    off = np.asarray(offsets, np.int64)
    nb = np.asarray(nbrs, np.int64)
    n = off.size - 1
    deg = np.diff(off).astype(np.uint64)
    tails = np.repeat(np.arange(n), np.diff(off))
    sums = []
    s = deg
    for _ in range(3):
        x = s[nb]
        # the sum modulo 2^64 from wrapping prefix sums; exact unless the true sum
        # reaches 2^64, which the f64 bound flags (those rows are redone in Python ints)
        cs = np.concatenate([np.zeros(1, np.uint64), np.cumsum(x, dtype=np.uint64)])
        tot = cs[off[1:]] - cs[off[:-1]]
        approx = np.bincount(tails, weights=x.astype(np.float64), minlength=n)
        for u in np.nonzero(approx >= 2.0 ** 63)[0]:
            exact = sum(int(t) for t in x[off[u]:off[u + 1]])
            tot[u] = np.uint64(min(exact, U64_MAX))
        sums.append(tot)
        s = tot
    m1 = np.zeros(n, np.uint64)
    np.maximum.at(m1, tails, deg[nb])
    return deg, sums[0], sums[1], sums[2], m1


def walk_inputs(offsets, nbrs, passes: int = 3) -> np.ndarray:
    """Surrogate inputs of every node, f32 [n, 6], from L = f64 log1p of the counts,
    each value rounded to f32 once:
      passes 2: L(deg), L(S_1), L(S_2), L(M_1), L(S_2) - L(S_1), L(S_1) - L(deg)
      passes 3: L(deg), L(S_1), L(S_2), L(S_3), L(M_1), L(S_3) - L(S_2)"""
    deg, s1, s2, s3, m1 = walk_counts(offsets, nbrs)
    L = lambda v: np.log1p(v.astype(np.float64))  # noqa: E731
    ld, l1, l2, l3, lm = L(deg), L(s1), L(s2), L(s3), L(m1)
    cols = [ld, l1, l2, lm, l2 - l1, l1 - ld] if passes == 2 else [ld, l1, l2, l3, lm, l3 - l2]
    return np.column_stack(cols).astype(np.float32)
