"""CPU tests of the host side: the C ABI library loads and exports every
declared symbol, the generators build canonical CSRs, and the Python mirror of
HyperBallInner validates its arguments.  No GPU compute calls."""
import ctypes as C
import os
import re
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for hdr in ("sparkx_hb.h", "sparkx_hb_graphgen.h"):
        with open(os.path.join(ROOT, "include", hdr)) as fh:
            text = re.sub(r"/\*.*?\*/", "", fh.read(), flags=re.S)
        names |= set(re.findall(r"\b(hb_[a-z_0-9]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    from sparkx_amd import _lib
    L = _lib.load()
    declared = _declared_functions()
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    nm = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    exported = set(re.findall(r" T (hb_[a-z_0-9]+)", nm))
    assert declared <= exported


def test_library_was_built_from_these_sources():
    """The loaded binary's recorded digest equals the digest of the sources in the tree."""
    from sparkx_amd import _lib, build
    assert _lib.load().hb_build_digest().decode() == build.source_digest()


def test_rust_binding_in_integration_matches_the_header():
    """The Rust binding documented for a maintainer declares hb_stats' fields in the C
    order with the same sizes, and only functions the header declares."""
    c_types = {"int32_t": 4, "uint32_t": 4, "float": 4, "double": 8, "uint64_t": 8}
    rust_types = {"i32": 4, "u32": 4, "f32": 4, "f64": 8, "u64": 8}
    with open(os.path.join(ROOT, "include", "sparkx_hb.h")) as fh:
        h = fh.read()
    body = re.sub(r"/\*.*?\*/", "", h[h.index("typedef struct {"):h.index("} hb_stats;")], flags=re.S)
    c_fields = []
    for decl in body[len("typedef struct {"):].split(";"):
        decl = decl.strip()
        if decl:
            ctype, names = decl.split(None, 1)
            for name in names.split(","):
                m = re.match(r"\s*(\w+)\s*(\[\w+\])?", name)
                c_fields.append((m.group(1), c_types[ctype], bool(m.group(2))))
    with open(os.path.join(ROOT, "INTEGRATION.md")) as fh:
        doc = fh.read()
    rust = doc[doc.index("pub struct HbStats {"):]
    rust = rust[:rust.index("\n}")]
    r_fields = [(name, rust_types[t], arr is not None and arr != "")
                for name, arr, t in re.findall(r"pub (\w+): (\[)?(\w+)", rust)]
    assert r_fields == c_fields
    declared = _declared_functions()
    for fn in re.findall(r"\bfn (hb_\w+)\s*\(", doc):
        assert fn in declared, fn


def test_create_reports_errors_without_crashing():
    from sparkx_amd import _lib
    L = _lib.load()
    assert L.hb_create(None, 0) == _lib.HB_EINVAL
    assert b"NULL" in L.hb_last_error()
    assert L.hb_run(None, 6, 0, 0, None) == _lib.HB_EINVAL
    assert L.hb_stage_csr(None, None, 8, 0, None, 8, 0) == _lib.HB_EINVAL
    assert L.hb_read_registers(None, 0, None) == _lib.HB_EINVAL
    assert L.hb_restore(None, 6, None, None, None) == _lib.HB_EINVAL
    assert b"hb_restore" in L.hb_last_error()
    assert L.hb_num_nodes(None) == 0
    L.hb_destroy(None)


def _check_canonical(g):
    off = g.offsets.astype(np.int64)
    nb = g.neighbours.astype(np.int64)
    n = g.size()
    assert off[0] == 0 and off[-1] == nb.size and np.all(np.diff(off) >= 0)
    src = np.repeat(np.arange(n), np.diff(off))
    assert np.all(nb < n) and not np.any(src == nb)
    key = src * n + nb
    assert np.all(np.diff(key) > 0)  # sorted per node, no duplicates
    assert np.array_equal(np.sort(nb * n + src), key)  # symmetric


def test_erdos_renyi_config_graph():
    from sparkx_amd import Graph
    g = Graph.erdos_renyi(65536, 1 << 20, seed=1)
    assert g.size() == 65536
    assert g.width() == 2096656  # recorded arc count of the ER config (seed 1)
    _check_canonical(g)
    g1 = Graph.erdos_renyi(65536, 1 << 20, seed=1, threads=1)
    assert np.array_equal(g.offsets, g1.offsets) and np.array_equal(g.neighbours, g1.neighbours)
    g2 = Graph.erdos_renyi(65536, 1 << 20, seed=2)
    assert not np.array_equal(g.neighbours[:1000], g2.neighbours[:1000])


def test_rmat_graph_shape():
    from sparkx_amd import Graph
    g = Graph.rmat(14, 16, seed=5)
    assert g.size() == 1 << 14
    _check_canonical(g)
    d = g.degrees()
    assert d.max() > 50 * d.mean()  # heavy-tailed hubs
    g1 = Graph.rmat(14, 16, seed=5, threads=1)
    assert np.array_equal(g.neighbours, g1.neighbours)


def test_kmer_graph_shape():
    from sparkx_amd import Graph
    g = Graph.kmer_dbg(50_000, 8, 0.005, 21, seed=3)
    _check_canonical(g)
    d = g.degrees()
    assert d.max() <= 8 and np.bincount(d).argmax() == 2
    assert 50_000 < g.size() < 2 * 50_000


def test_generators_reject_bad_arguments():
    from sparkx_amd import Graph, HyperBallError
    with pytest.raises(HyperBallError):
        Graph.rmat(0)
    with pytest.raises(HyperBallError):
        Graph.kmer_dbg(10, 8, 0.005, 21)


def test_graph_mirror_bounds():
    from sparkx_amd import Graph
    g = Graph(np.array([0, 1, 2], np.uint64), np.array([1, 0], np.uint64))
    assert g.size() == 2 and g.width() == 2
    assert list(g.neighbours_of(0)) == [1]
    with pytest.raises(IndexError):
        g.neighbours_of(2)
    with pytest.raises(TypeError):
        Graph(np.array([0, 1], np.int32), np.array([0], np.uint32))


def test_hyperball_rejects_bad_precision():
    from sparkx_amd import Graph, HyperBall
    g = Graph(np.array([0, 0], np.uint64), np.zeros(0, np.uint64))
    with pytest.raises(ValueError):
        HyperBall(g, 3)
    with pytest.raises(ValueError):
        HyperBall(g, 17)


def _build_cpp_mirror(tmp_path, name="hyperball_cpp_mirror"):
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
    exe = str(tmp_path / name)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", src,
           "-L" + os.path.join(ROOT, "sparkx_amd"), "-lsparkx_hb", "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


def test_cpp_mirror_compiles_and_links(tmp_path):
    """The header-only C++ mirror of HyperBallInner builds against the C ABI (and its
    get_or_compute client too)."""
    assert os.path.exists(_build_cpp_mirror(tmp_path))
    assert os.path.exists(_build_cpp_mirror(tmp_path, "hyperball_cpp_cache"))


def _write_dbg_mtx(path, sigma, k):
    # MatrixMarket symmetric pattern file of the complete de Bruijn graph on
    # sigma^k nodes: node u -> ((u mod sigma^(k-1)) * sigma + c), 1-based
    n, base = sigma ** k, sigma ** (k - 1)
    with open(path, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate pattern symmetric\n% de Bruijn\n")
        fh.write(f"{n} {n} {sigma ** (k + 1)}\n")
        for u in range(n):
            for c in range(sigma):
                fh.write(f"{u + 1} {(u % base) * sigma + c + 1}\n")


def test_mtx_reader_matches_reference_rules(tmp_path, oracle):
    from sparkx_amd.cache import read_mtx
    path = str(tmp_path / "dbg.mtx")
    _write_dbg_mtx(path, 4, 5)
    g = read_mtx(path)
    off, nb = oracle.dbg_mtx_csr(4, 5)
    assert np.array_equal(g.offsets, off) and np.array_equal(g.neighbours, nb.astype(np.uint64))


@pytest.mark.parametrize("kind", ["general", "symmetric", "skew-symmetric", "Symmetric"])
def test_mtx_reader_symmetry_kinds_match_the_oracle(tmp_path, oracle, kind):
    """Every header but "general" emits j -> i too (self-loops once), as the reference's
    producer and the device parser do; comments, blank and single-token lines skip."""
    from sparkx_amd.cache import read_mtx
    rng = np.random.default_rng(len(kind))
    n = 30
    body = "\n".join(f"{a} {b}" for a, b in rng.integers(1, n + 1, size=(80, 2)))
    text = f"%%MatrixMarket matrix coordinate pattern {kind}\n% c\n{n} {n} 80\n\n7\n% mid\n{body}\n"
    path = tmp_path / "g.mtx"
    path.write_text(text)
    g = read_mtx(str(path))
    off, nb, info = oracle.parse_mtx(text.encode())
    assert info.symmetric == (kind.lower() != "general")
    assert np.array_equal(g.offsets, off) and np.array_equal(g.neighbours, nb.astype(np.uint64))


@pytest.mark.parametrize("text,what", [
    ("%%MatrixMarket matrix coordinate pattern general\n3 4 1\n1 2\n", "not square"),
    ("%%MatrixMarket matrix array pattern general\n3 3 1\n1 2\n", "unsupported"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 2\n", "unsupported"),
    ("%%MatrixMarket matrix coordinate pattern\n3 3 1\n1 2\n", "malformed"),
    ("%%matrixmarket matrix coordinate pattern general\n3 3 1\n1 2\n", "not a MatrixMarket"),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n0 2\n", "index 0"),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n1 4\n", "above n"),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n1 x\n", "not a number"),
])
def test_mtx_reader_rejects_what_the_reference_rejects(tmp_path, oracle, text, what):
    from sparkx_amd.cache import read_mtx
    path = tmp_path / "bad.mtx"
    path.write_text(text)
    with pytest.raises(ValueError, match=what):
        read_mtx(str(path))
    with pytest.raises(oracle.MtxError):
        oracle.parse_mtx(text.encode())


def test_cache_ids_and_names():
    from sparkx_amd import cache
    # the id of the reference's committed exact-closeness fixture for this dataset
    assert cache.graph_id("./datasets/graphs/synth_gen_dbg_protein20_k2_edges16k.mtx") == \
        "46dfbb599cc933be7aca48553950fc09d8418d0cecb287a2598df9df16821730"
    gid = "abc"
    assert cache.cache_file("/c", "distances", gid).endswith("/hyperballdistances_abc.mmap")
    assert cache.cache_file("/c", "closeness", gid, cache.centrality_seq("closeness", True)).endswith(
        "/hyperballcloseness_2_abc.mmap")
    assert cache.cache_file("/c", "harmonic", gid, cache.centrality_seq("harmonic", False)).endswith(
        "/hyperballharmonic_1_abc.mmap")
    assert cache.cache_file("/c", "lin", gid, cache.centrality_seq("lin", False)).endswith("/hyperballlin_0_abc.mmap")


def test_graph_cache_roundtrip(tmp_path):
    from sparkx_amd import Graph, cache
    g = Graph.erdos_renyi(500, 2000, seed=3)
    cache.save_graph_cache(g, str(tmp_path), "gid")
    h = cache.load_cached_graph(str(tmp_path), "gid")
    assert h.offsets.dtype == np.dtype("<u8")
    assert np.array_equal(h.offsets, g.offsets) and np.array_equal(h.neighbours, g.neighbours)


def test_drop_cache_follows_the_reference_quirks(tmp_path):
    """hyper_ball.rs:564-593: closeness slots both remove seq 1; harmonic files go
    with the closeness slots; a missing file is an error."""
    from types import SimpleNamespace

    from sparkx_amd import cache
    d, gid = str(tmp_path), "g"

    def touch(kind, seq=None):
        open(cache.cache_file(d, kind, gid, seq), "wb").close()

    for k in ("distances", "inv_distances"):
        touch(k)
    touch("registers", 7)
    touch("closeness", 1)
    touch("harmonic", 1)
    touch("lin", 0)
    info = lambda: (7, 0, 0)  # noqa: E731  (p of the run: the registers file's seq)
    hb = SimpleNamespace(_closeness=[1, None], _harmonic=[None, None], _lin=1, result_info=info)
    cache.drop_cache(hb, d, gid)
    assert os.listdir(d) == []
    for k in ("distances", "inv_distances"):
        touch(k)
    touch("registers", 7)
    touch("closeness", 1)
    touch("closeness", 2)
    hb = SimpleNamespace(_closeness=[1, 1], _harmonic=[None, None], _lin=None, result_info=info)
    with pytest.raises(FileNotFoundError):
        cache.drop_cache(hb, d, gid)  # the second slot names seq 1 again
    assert os.path.exists(cache.cache_file(d, "closeness", gid, 2))


def test_bench_algorithmic_bytes_per_unit():
    """The roofline's algorithmic bytes, built up from the arrays a recomputed node
    touches: its two u32 offsets, its 2^p-byte row read and written, its f32 estimate
    and two f64 accumulators read and written; per neighbour a u32 id and, when the
    neighbour's block is gathered, 2^p bytes."""
    sys.path.insert(0, ROOT)
    import bench
    u32, f32, f64 = (np.dtype(t).itemsize for t in (np.uint32, np.float32, np.float64))
    for p in (4, 7, 10):
        R = 1 << p
        node = 2 * u32 + 2 * R + 2 * f32 + 2 * 2 * f64
        # a dense iteration: one node, three neighbours, all gathered
        assert bench.algorithmic_bytes(0, 3, 1, 1, p, 3) == node + 3 * (u32 + R)
        # selective: two of the three neighbours changed and are gathered
        assert bench.algorithmic_bytes(1, 3, 1, 1, p, 2) == node + 3 * u32 + 2 * R
        # iteration 1 hashes ids: no neighbour block and not the node's own block is read
        assert bench.algorithmic_bytes(0, 3, 1, 1, p, 0, first_nodes=1) == node - R + 3 * u32
        # hub chunks: one partial sketch written per task besides their gathers
        assert bench.algorithmic_bytes(3, 10, 0, 2, p, 10) == 10 * (u32 + R) + 2 * R
    # the dense R-MAT-24 p = 7 iteration quoted beside the kernels
    assert abs(bench.algorithmic_bytes(0, 520768654, 16777216, 0, 7, 520768654) - 73.8e9) < 0.1e9


def test_bench_traffic_files_match_kernel_workload_and_layout():
    """The bench line's `traffic` comes from the committed counter summary of the same
    kernel, workload, precision and node layout, and from no other."""
    sys.path.insert(0, ROOT)
    import bench
    t, src = bench.measured_traffic("hb::k_union<7, 8>", "rmat24", 7, "degree")
    assert t > 0 and src.endswith("k_union_p7_t8_rmat24_traffic.txt")
    assert bench.measured_traffic("hb::k_union<7, 8>", "rmat24", 7, "chain") == (None, None)
    t_chain, src = bench.measured_traffic("hb::k_union_rows<6>", "kmer", 6, "chain")
    t_degree, _ = bench.measured_traffic("hb::k_union<6, 4>", "kmer", 6, "degree")
    assert src.endswith("k_union_rows_p6_kmer_chain_traffic.txt") and 0 < t_chain < t_degree
    assert bench.measured_traffic("hb::k_union<6, 4>", "kmer", 6, "chain") == (None, None)


def test_bench_names_the_kernel_that_runs_the_dense_iterations():
    """In the chain layout at p = 5 and 6 the one-copy bin is run by k_union_rows after
    the first iteration; everywhere else the slot's kernel is k_union<p, team>."""
    sys.path.insert(0, ROOT)
    import bench
    assert bench.kernel_name(0, 6, 4, "chain") == "hb::k_union_rows<6>"
    assert bench.kernel_name(0, 5, 2, "chain") == "hb::k_union_rows<5>"
    assert bench.kernel_name(0, 6, 4, "degree") == "hb::k_union<6, 4>"
    assert bench.kernel_name(0, 7, 8, "chain") == "hb::k_union<7, 8>"
    assert bench.kernel_name(0, 6, 8, "chain") == "hb::k_union<6, 8>"
    assert bench.kernel_name(3, 7, 64, "degree") == "hb::k_chunk<7>"


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says so, before
    touching the library or a device."""
    import subprocess
    import sys as _sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([_sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


@pytest.mark.parametrize("p", list(range(4, 17)))
def test_bias_windows_equal_the_nearest_neighbour_rule(oracle, p):
    """The estimator's tabulated HLL++ bias (window breakpoints, window means and the
    per-bucket starting window the kernels read) gives, for every f32 estimate in
    [m/2, 5m], exactly the bias the k = 6 nearest-neighbour rule gives (the oracle's
    search, expansion and ascending sum).  Estimates reaching the bias branch lie in
    that range: the raw estimate is at least alpha m > m/2 (every register term is at
    most 1) and the branch needs it at most 5m.  The device steps up at most 2 times
    from a bucket's start."""
    from sparkx_amd import _lib
    L = _lib.load()
    m = 1 << p
    mean = np.zeros(195)
    bp = np.zeros(195, np.float32)
    bk = np.zeros(1024, np.uint8)
    _lib.check(L.hb_bias_windows(p, _lib.ptr(mean, C.c_double), _lib.ptr(bp, C.c_float),
                                 _lib.ptr(bk, C.c_uint8)))
    assert np.all(np.diff(bp[1:]) > 0) and np.all(np.diff(bk.astype(int)) >= 0)
    lo, hi = int(np.float32(m / 2).view(np.uint32)), int(np.float32(5 * m).view(np.uint32))
    base = (127 + p - 1) << 8
    for c0 in range(lo, hi + 1, 1 << 22):
        bits = np.arange(c0, min(c0 + (1 << 22), hi + 1), dtype=np.uint32)
        e = bits.view(np.float32)
        b = (bits >> 15).astype(np.int64) - base
        s = np.where(b < 0, 0, bk[np.clip(b, 0, 1023)]).astype(np.int64)
        for _ in range(2):  # the device's walk: while e >= bp[s + 1], s += 1
            s = s + ((s + 1 < 195) & (bp[np.minimum(s + 1, 194)] <= e))
        assert not ((s + 1 < 195) & (bp[np.minimum(s + 1, 194)] <= e)).any(), "more than 2 steps"
        got = mean[s]
        ref = oracle.knn_bias(p, e)
        bad = np.flatnonzero(got != ref)
        assert bad.size == 0, (p, e[bad[:5]], got[bad[:5]], ref[bad[:5]])
    assert L.hb_bias_windows(3, _lib.ptr(mean, C.c_double), _lib.ptr(bp, C.c_float),
                             _lib.ptr(bk, C.c_uint8)) == _lib.HB_EINVAL
