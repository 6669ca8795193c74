"""GPU graph ingestion (hb_stage_mtx / hb_stage_coo / hb_read_csr) against the oracle's
MatrixMarket reader (oracle/mtx_oracle.c, pinned by tests/test_cpu_mtx.py) and a numpy
edge-list reference: offsets, neighbours, counts, error codes and error lines must be
identical."""
import ctypes as C
import sys

import numpy as np
import pytest

from test_cpu_mtx import HDR, dbg_text, random_mtx

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hb():
    from sparkx_amd import _lib
    _lib.load()
    import sparkx_amd
    return sparkx_amd


def fmt_pairs(i1, j1) -> bytes:
    """'i j\\n' lines (right-aligned, space padded) for 1-based index arrays."""
    i1 = np.asarray(i1, np.int64)
    j1 = np.asarray(j1, np.int64)
    w = max(1, len(str(int(max(i1.max(initial=1), j1.max(initial=1))))))

    def col(x):
        d = np.empty((x.size, w), np.uint8)
        y = x.copy()
        for k in range(w - 1, -1, -1):
            d[:, k] = 48 + y % 10
            y //= 10
        lead = np.cumsum(d != 48, axis=1) == 0
        lead[:, -1] = False
        d[lead] = 32
        return d

    rows = np.concatenate([col(i1), np.full((i1.size, 1), 32, np.uint8), col(j1),
                           np.full((i1.size, 1), 10, np.uint8)], axis=1)
    return rows.tobytes()


def stage_mtx(hb, text):
    """(rc, info, offsets, neighbours) through the C ABI on a fresh context."""
    from sparkx_amd import _lib
    L = _lib.load()
    ctx = C.c_void_p()
    _lib.check(L.hb_create(C.byref(ctx), 0))
    try:
        _lib.check(L.hb_set_layout(ctx, _lib.HB_LAYOUT_INPUT))
        info = _lib.HbMtxInfo()
        rc = L.hb_stage_mtx(ctx, text, len(text), C.byref(info))
        if rc:
            return rc, info, None, None
        n, arcs = int(L.hb_num_nodes(ctx)), int(L.hb_num_arcs(ctx))
        off = np.empty(n + 1, np.uint64)
        nb = np.empty(max(arcs, 1), np.uint64)
        _lib.check(L.hb_read_csr(ctx, _lib.ptr(off, C.c_uint64), _lib.ptr(nb, C.c_uint64)))
        return 0, info, off, nb[:arcs]
    finally:
        L.hb_destroy(ctx)


def check_against_oracle(hb, oracle, text, ctx=""):
    rc, info, off, nb = stage_mtx(hb, text)
    assert rc == 0, ctx
    ref_off, ref_nb, ref = oracle.parse_mtx(text)
    assert (info.n, info.nnz, info.entries, info.arcs, bool(info.symmetric)) == \
        (ref.n, ref.nnz, ref.entries, ref.arcs, bool(ref.symmetric)), ctx
    assert np.array_equal(off, ref_off), ctx
    assert np.array_equal(nb, ref_nb.astype(np.uint64)), ctx
    return info


@pytest.mark.parametrize("sigma,k", [(4, 3), (20, 2), (4, 7), (3, 1)])
def test_mtx_dbg_fixtures(hb, oracle, sigma, k):
    say("mtx dbg", sigma, k)
    check_against_oracle(hb, oracle, dbg_text(sigma, k), (sigma, k))


@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric"])
@pytest.mark.parametrize("field", ["pattern", "integer"])
def test_mtx_random_files(hb, oracle, symmetry, field):
    rng = np.random.default_rng(len(symmetry) * 7 + len(field))
    for n, m in ((1, 3), (7, 40), (300, 2500), (5000, 60000)):
        say("mtx random", symmetry, field, n, m)
        check_against_oracle(hb, oracle, random_mtx(rng, n, m, field, symmetry), (symmetry, field, n))


def test_mtx_token_rules_and_edges(hb, oracle):
    text = (b"%%MatrixMarket MATRIX Coordinate Pattern GENERAL extra\r\n"
            b"\r\n% comment\r\n  4\t4   5  \r\n"
            b"1 2\r\n\t+3\t4 99 junk\r\n% inside\n\n2\n4 4\n1 2")
    info = check_against_oracle(hb, oracle, text, "token rules")
    assert info.entries == 4
    for t in (HDR + b"3 3 0", HDR + b"3 3 0\n", HDR + b"1 1 1\n1 1\n", HDR + b"5 5 2\n5 1\n1 5\n\n\n",
              b"%%MatrixMarket matrix coordinate pattern symmetric\n6 6 3\n6 6\n1 6\n6 1\n"):
        check_against_oracle(hb, oracle, t, t[-20:])


def test_mtx_rmat_text(hb, oracle):
    """An R-MAT 2^16 graph written as a symmetric file (one line per undirected edge)."""
    g = hb.Graph.rmat(16, 16, seed=4)
    off = g.offsets.astype(np.int64)
    u = np.repeat(np.arange(g.size()), np.diff(off))
    v = g.neighbours.astype(np.int64)
    keep = u <= v
    body = fmt_pairs(u[keep] + 1, v[keep] + 1)
    text = (f"%%MatrixMarket matrix coordinate pattern symmetric\n{g.size()} {g.size()} {int(keep.sum())}\n"
            .encode() + body)
    say("mtx rmat16", len(text))
    check_against_oracle(hb, oracle, text, "rmat16")
    _, info, goff, gnb = stage_mtx(hb, text)
    # the file describes the generator's symmetric CSR exactly
    assert np.array_equal(goff, g.offsets.astype(np.uint64))
    assert np.array_equal(gnb, g.neighbours.astype(np.uint64))
    assert info.ms_device > 0


@pytest.mark.parametrize("text,code,line", [
    (b"", 1, 1),
    (b"%%matrixmarket matrix coordinate pattern general\n1 1 0\n", 1, 1),
    (b"%%MatrixMarket matrix coordinate pattern\n1 1 0\n", 2, 1),
    (b"%%MatrixMarket matrix coordinate real general\n1 1 0\n", 3, 1),
    (HDR + b"% only comments\n\n", 4, 0),
    (HDR + b"3 3\n", 5, 2),
    (HDR + b"3 4 1\n1 1\n", 6, 2),
    (HDR + b"3 3 2\n1 2\n2 -1\n", 7, 4),
    (HDR + b"3 3 2\n1 2\n99999999999999999999 1\n", 7, 4),
    (HDR + b"3 3 2\n% c\n0 1\n", 8, 4),
    (HDR + b"3 3 3\n1 2\n1 4\n0 1\n", 9, 4),  # the first bad line is reported
])
def test_mtx_errors(hb, oracle, text, code, line):
    rc, info, _, _ = stage_mtx(hb, text)
    from sparkx_amd import _lib
    assert rc == _lib.HB_EINVAL
    assert info.err_code == code and info.err_line == line
    with pytest.raises(oracle.MtxError) as ei:
        oracle.parse_mtx(text)
    assert ei.value.code == -code
    if code >= 5:
        assert ei.value.line == line


def test_hyperball_from_mtx(hb, oracle):
    """GraphCache::from_mtx_file + HyperBall on the de Bruijn fixture, parsed on the GPU:
    identical to the run on the oracle's CSR."""
    text = dbg_text(20, 2)
    off, nb = oracle.dbg_mtx_csr(20, 2)
    ref = oracle.hyperball(off, nb, 8, threads=4)
    with hb.HyperBall.from_mtx(text, 8) as h:
        assert h.ingest_info.entries == 8000 and h.g.size() == 400
        assert h.stats.iterations == ref.iterations
        from sparkx_amd import _lib
        assert np.array_equal(h.registers(_lib.HB_REGS_REFERENCE).ravel(), ref.regs_ref)
        d, i = h._accumulators()
        assert np.array_equal(d, ref.dist) and np.array_equal(i, ref.inv)


def coo_reference(n, s, d, symmetrize, dedup, no_loops):
    s = s.astype(np.int64)
    d = d.astype(np.int64)
    if no_loops:
        keep = s != d
        s, d = s[keep], d[keep]
    if symmetrize:
        off_diag = s != d
        s, d = np.concatenate([s, d[off_diag]]), np.concatenate([d, s[off_diag]])
    keys = np.sort((s << 32) | d)
    if dedup:
        keys = np.unique(keys)
    u = keys >> 32
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, u + 1, 1)
    return np.cumsum(off).astype(np.uint64), (keys & 0xffffffff).astype(np.uint64)


@pytest.mark.parametrize("flags", range(8))
def test_coo_flags(hb, flags):
    from sparkx_amd.ingest import coo_graph
    sym, dedup, no_loops = bool(flags & 1), bool(flags & 2), bool(flags & 4)
    rng = np.random.default_rng(flags)
    for n, m in ((1, 0), (1, 5), (50, 400), (100000, 300000)):
        s = rng.integers(0, n, m).astype(np.uint32)
        d = rng.integers(0, n, m).astype(np.uint32)
        if m > 10:
            s[:10] = d[:10]  # some loops
            s[10:20], d[10:20] = s[20:30], d[20:30]  # some duplicates
        say("coo", flags, n, m)
        g = coo_graph(n, s, d, symmetrize=sym, dedup=dedup, no_loops=no_loops)
        ref_off, ref_nb = coo_reference(n, s, d, sym, dedup, no_loops)
        assert np.array_equal(g.offsets, ref_off) and np.array_equal(g.neighbours, ref_nb), (flags, n, m)


def test_coo_errors(hb):
    from sparkx_amd.ingest import coo_graph
    with pytest.raises(hb.HyperBallError, match="entry 2 has an id >= n=4"):
        coo_graph(4, [0, 1, 4], [1, 2, 3])
