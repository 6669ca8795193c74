"""On-disk compatibility with the reference's graph and result caches.

The reference keeps every array as a raw little-endian memory-mapped file in
CACHE_DIR (src/graph/cache/utils.rs:10-13) named `<suffix>_<id>.mmap` or, for
sequenced files, `<suffix>_<seq>_<id>.mmap` (:318-341), where `id` is the SHA-256
hex digest of the graph's key (:53-67).  The CSR is `index_<id>.mmap`
(u64 offsets[n+1]) and `edges_<id>.mmap` (u64 neighbours); HyperBall results are
`hyperballdistances_<id>.mmap`, `hyperballinvdistances_<id>.mmap` (f64[n]) and the
centralities `hyperballcloseness_<seq>_<id>.mmap`, `hyperballharmonic_<seq>_<id>.mmap`
(seq 1: normalize=false, seq 2: normalize=true) and `hyperballlin_0_<id>.mmap`
(src/centralities/hyper_ball.rs:203-246).  The counters file `hyperball_<id>.mmap`
holds hyperloglog-rs structs whose byte layout is not available offline, so it is
not written; the reference's search_cache_* and the accumulator files interoperate.

MTX input follows the reference's reader (src/graph/cache/parser_mtx/mod.rs:34-162,
producers.rs:228-234, impl_sort.rs:10-51): 1-based entries, `symmetric` headers
emit j->i for i != j, self-loops once, duplicates kept, each list sorted.
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

from .graph import Graph

SUFFIX = {
    "edges": "edges", "index": "index", "distances": "hyperballdistances",
    "inv_distances": "hyperballinvdistances", "closeness": "hyperballcloseness",
    "harmonic": "hyperballharmonic", "lin": "hyperballlin",
}


def graph_id(key: str) -> str:
    """utils.rs:53-67: SHA-256 of the key, lower-case hex."""
    return hashlib.sha256(key.encode()).hexdigest()


def cache_file(cache_dir: str, kind: str, gid: str, seq: int | None = None) -> str:
    name = f"{SUFFIX[kind]}_{gid}.mmap" if seq is None else f"{SUFFIX[kind]}_{seq}_{gid}.mmap"
    return os.path.join(cache_dir, name)


def centrality_seq(kind: str, normalize: bool) -> int:
    """hyper_ball.rs:208-246: closeness/harmonic seq 1 (false) or 2 (true); Lin 0."""
    return 0 if kind == "lin" else (2 if normalize else 1)


def load_cached_graph(cache_dir: str, gid: str) -> Graph:
    """Memory-maps the reference's u64 CSR (index_<id>.mmap, edges_<id>.mmap)."""
    off = np.memmap(cache_file(cache_dir, "index", gid), dtype="<u8", mode="r")
    nbr_path = cache_file(cache_dir, "edges", gid)
    nbr = (np.memmap(nbr_path, dtype="<u8", mode="r") if os.path.getsize(nbr_path)
           else np.zeros(0, np.uint64))
    return Graph(np.asarray(off), np.asarray(nbr), gid)


def save_graph_cache(g: Graph, cache_dir: str, gid: str) -> None:
    os.makedirs(cache_dir, exist_ok=True)
    np.asarray(g.offsets, dtype="<u8").tofile(cache_file(cache_dir, "index", gid))
    np.asarray(g.neighbours, dtype="<u8").tofile(cache_file(cache_dir, "edges", gid))


def write_results(hb, cache_dir: str, gid: str) -> list:
    """Writes the accumulators and every centrality `hb` has computed under the
    reference's file names; returns the paths written."""
    os.makedirs(cache_dir, exist_ok=True)
    dist, inv = hb._accumulators()
    out = []
    for kind, arr in (("distances", dist), ("inv_distances", inv)):
        path = cache_file(cache_dir, kind, gid)
        np.asarray(arr, "<f8").tofile(path)
        out.append(path)
    for norm in (False, True):
        for kind, cached in (("closeness", hb._closeness[int(norm)]), ("harmonic", hb._harmonic[int(norm)])):
            if cached is not None:
                path = cache_file(cache_dir, kind, gid, centrality_seq(kind, norm))
                np.asarray(cached, "<f8").tofile(path)
                out.append(path)
    if hb._lin is not None:
        path = cache_file(cache_dir, "lin", gid, 0)
        np.asarray(hb._lin, "<f8").tofile(path)
        out.append(path)
    return out


def drop_cache(hb, cache_dir: str, gid: str) -> None:
    """HyperBallInner::drop_cache (hyper_ball.rs:564-593) over the files write_results
    wrote: the accumulators, then the closeness files -- the reference names seq 1
    for both normalisation slots (:575-579), so with both cached the second removal
    fails -- then the harmonic files, whose removal is conditioned on the *closeness*
    slots (:582-586), then Lin's.  A missing file raises FileNotFoundError, as the
    reference's `?` returns Err.  The counters file is not written here (its
    hyperloglog-rs layout is unavailable), so it is not removed."""
    for kind in ("distances", "inv_distances"):
        os.remove(cache_file(cache_dir, kind, gid))
    for i in (0, 1):
        if hb._closeness[i] is not None:
            os.remove(cache_file(cache_dir, "closeness", gid, 1))
    for i, seq in ((0, 1), (1, 2)):
        if hb._closeness[i] is not None:
            os.remove(cache_file(cache_dir, "harmonic", gid, seq))
    if hb._lin is not None:
        os.remove(cache_file(cache_dir, "lin", gid, 0))


def read_result(cache_dir: str, kind: str, gid: str, normalize: bool | None = None) -> np.ndarray:
    seq = None if kind in ("distances", "inv_distances") else centrality_seq(kind, bool(normalize))
    return np.fromfile(cache_file(cache_dir, kind, gid, seq), dtype="<f8")


def read_mtx(path: str) -> Graph:
    """CSR of a Matrix Market `coordinate` file with the reference's rules."""
    with open(path) as fh:
        header = fh.readline().lower().split()
        if len(header) < 5 or header[0] != "%%matrixmarket" or header[2] != "coordinate":
            raise ValueError(f"{path}: not a MatrixMarket coordinate file")
        symmetric = header[4] == "symmetric"
        line = fh.readline()
        while line.startswith("%"):
            line = fh.readline()
        nrows, ncols, _ = (int(v) for v in line.split()[:3])
        data = np.loadtxt(fh, dtype=np.int64, usecols=(0, 1), ndmin=2)
    n = max(nrows, ncols)
    i = data[:, 0] - 1
    j = data[:, 1] - 1
    if data.size and (i.min() < 0 or j.min() < 0 or i.max() >= nrows or j.max() >= ncols):
        raise ValueError(f"{path}: entry out of bounds")
    if symmetric:
        off_diag = i != j
        src = np.concatenate([i, j[off_diag]])
        dst = np.concatenate([j, i[off_diag]])
    else:
        src, dst = i, j
    order = np.lexsort((dst, src))
    nbr = dst[order].astype(np.uint64)
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(src, minlength=n), out=off[1:])
    return Graph(off, nbr, os.path.basename(path))
