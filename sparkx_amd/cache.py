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
neither written nor read: the registers go to a file of this library's own,
`hyperballregisters_<p>_<id>.mmap` (u8, node-major, the buffer the reference's
counters file would hold), which the reference never opens.  The accumulator and
centrality files interoperate; the reference's own get_or_compute (hyper_ball.rs:
165-201) finds no counters file and recomputes.

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
    "harmonic": "hyperballharmonic", "lin": "hyperballlin", "registers": "hyperballregisters",
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


def write_centrality(cache_dir: str, kind: str, gid: str, normalize: bool, values) -> str:
    """One centrality file under the reference's name (as its compute_* create it)."""
    os.makedirs(cache_dir, exist_ok=True)
    path = cache_file(cache_dir, kind, gid, centrality_seq(kind, normalize))
    np.asarray(values, "<f8").tofile(path)
    return path


def write_results(hb, cache_dir: str, gid: str) -> list:
    """Writes the accumulators and every centrality `hb` has computed under the
    reference's file names, and the registers under this library's own name; returns
    the paths written."""
    os.makedirs(cache_dir, exist_ok=True)
    dist, inv = hb._accumulators()
    out = []
    p = hb.result_info()[0]
    path = cache_file(cache_dir, "registers", gid, p)
    hb.registers().tofile(path)
    out.append(path)
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
    wrote: the registers (in place of the reference's counters file), the
    accumulators, then the closeness files -- the reference names seq 1 for both
    normalisation slots (:575-579), so with both cached the second removal fails --
    then the harmonic files, whose removal is conditioned on the *closeness* slots
    (:582-586), then Lin's.  A missing file raises FileNotFoundError, as the
    reference's `?` returns Err."""
    os.remove(cache_file(cache_dir, "registers", gid, hb.result_info()[0]))
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
    """CSR of a MatrixMarket file with the reference's rules (parser_mtx/mod.rs:34-162,
    producers.rs:228-234, impl_sort.rs:10-51), the same ones hb_stage_mtx applies on the
    device: the first line starts with "%%MatrixMarket" (case-sensitive) and names
    "matrix coordinate pattern|integer symmetric|skew-symmetric|general" (case-
    insensitive); every kind but "general" also emits j -> i for i != j; rows must equal
    columns; blank and '%' lines and lines with a single token are skipped; indices are
    1-based, 0 and values above n are errors; duplicates are kept and every list is
    sorted ascending.  Raises ValueError naming the first bad line."""
    with open(path, "rb") as fh:
        lines = fh.read().decode("ascii").split("\n")
    head = lines[0].rstrip("\r")
    if not head.startswith("%%MatrixMarket"):
        raise ValueError(f"{path}: line 1: not a MatrixMarket file")
    tok = head.split()
    if len(tok) < 5:
        raise ValueError(f"{path}: line 1: malformed header")
    kind = [t.lower() for t in tok[1:5]]
    if (kind[0] != "matrix" or kind[1] != "coordinate" or kind[2] not in ("pattern", "integer")
            or kind[3] not in ("symmetric", "skew-symmetric", "general")):
        raise ValueError(f"{path}: line 1: unsupported MatrixMarket format {' '.join(tok[1:5])}")
    symmetric = kind[3] != "general"
    ln = 1
    while True:
        if ln >= len(lines):
            raise ValueError(f"{path}: no size line")
        size = lines[ln].split()
        ln += 1
        if size and not size[0].startswith("%"):
            break
    try:
        nrows, ncols = int(size[0]), int(size[1])
    except (IndexError, ValueError):
        raise ValueError(f"{path}: line {ln}: bad size line") from None
    if nrows != ncols:
        raise ValueError(f"{path}: line {ln}: matrix is not square ({nrows} x {ncols})")
    n = nrows
    src, dst = [], []
    for k in range(ln, len(lines)):
        t = lines[k].split()
        if len(t) < 2 or t[0].startswith("%"):
            continue
        a, b = t[0].lstrip("+"), t[1].lstrip("+")
        if not (a.isdigit() and b.isdigit() and t[0].count("+") <= 1 and t[1].count("+") <= 1):
            raise ValueError(f"{path}: line {k + 1}: index is not a number")
        i, j = int(a), int(b)
        if i == 0 or j == 0:
            raise ValueError(f"{path}: line {k + 1}: index 0 (MatrixMarket is 1-based)")
        if i > n or j > n:
            raise ValueError(f"{path}: line {k + 1}: index above n = {n}")
        src.append(i - 1)
        dst.append(j - 1)
    i = np.asarray(src, np.int64)
    j = np.asarray(dst, np.int64)
    if symmetric:
        off_diag = i != j
        i, j = np.concatenate([i, j[off_diag]]), np.concatenate([j, i[off_diag]])
    order = np.lexsort((j, i))
    nbr = j[order].astype(np.uint64)
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(i, minlength=n), out=off[1:])
    return Graph(off, nbr, os.path.basename(path))
