"""Graphs built on the GPU: MatrixMarket text (the reference's GraphCache::from_mtx_file,
src/graph/cache/parser_mtx/mod.rs:34-162, producers.rs:15-234, impl_sort.rs:10-51) and
edge lists with symmetrisation / deduplication / loop removal, through hb_stage_mtx /
hb_stage_coo.  The CSR stays on the device for HyperBall; `read_mtx` / `coo_graph`
also copy it back as 8-byte offsets and neighbours, the layout of the reference's
index_<id> / edges_<id> cache files."""
from __future__ import annotations

import ctypes as C
import os
from typing import Tuple, Union

import numpy as np

from . import _lib
from .graph import Graph

Text = Union[bytes, bytearray, memoryview, str, os.PathLike]


def _bytes(text: Text) -> bytes:
    if isinstance(text, (bytes, bytearray, memoryview)):
        return bytes(text)
    with open(text, "rb") as f:
        return f.read()


def _read_back(L, ctx, name: str) -> Graph:
    n = int(L.hb_num_nodes(ctx))
    arcs = int(L.hb_num_arcs(ctx))
    off = np.empty(n + 1, np.uint64)
    nb = np.empty(max(arcs, 1), np.uint64)
    _lib.check(L.hb_read_csr(ctx, _lib.ptr(off, C.c_uint64), _lib.ptr(nb, C.c_uint64)))
    return Graph(off, nb[:arcs], name)


def stage_mtx_fn(text: Text):
    """A staging callable for HyperBall(_stage_fn=...): parses `text` into the context."""
    data = _bytes(text)

    def stage(L, ctx):
        info = _lib.HbMtxInfo()
        _lib.check(L.hb_stage_mtx(ctx, data, len(data), C.byref(info)))
        return _read_back(L, ctx, "mtx"), info
    return stage


def read_mtx(text: Text, device: int = 0) -> Tuple[Graph, "_lib.HbMtxInfo"]:
    """MatrixMarket text (bytes or a path) -> (Graph with u64 CSR, parse summary),
    parsed on the GPU."""
    L = _lib.load()
    ctx = C.c_void_p()
    _lib.check(L.hb_create(C.byref(ctx), device))
    try:
        _lib.check(L.hb_set_layout(ctx, _lib.HB_LAYOUT_INPUT))  # no renumbering needed here
        return stage_mtx_fn(text)(L, ctx)
    finally:
        L.hb_destroy(ctx)


def coo_graph(n: int, src, dst, *, symmetrize: bool = False, dedup: bool = False, no_loops: bool = False,
              device: int = 0) -> Graph:
    """Edge list (src[k] -> dst[k], ids < n) -> Graph, built on the GPU."""
    s = np.ascontiguousarray(src, np.uint32)
    d = np.ascontiguousarray(dst, np.uint32)
    if s.shape != d.shape:
        raise ValueError("src and dst differ in length")
    flags = ((_lib.HB_COO_SYMMETRIZE if symmetrize else 0) | (_lib.HB_COO_DEDUP if dedup else 0)
             | (_lib.HB_COO_NO_LOOPS if no_loops else 0))
    L = _lib.load()
    ctx = C.c_void_p()
    _lib.check(L.hb_create(C.byref(ctx), device))
    try:
        _lib.check(L.hb_set_layout(ctx, _lib.HB_LAYOUT_INPUT))
        _lib.check(L.hb_stage_coo(ctx, n, _lib.ptr(s, C.c_uint32), _lib.ptr(d, C.c_uint32), s.size, flags))
        return _read_back(L, ctx, "coo")
    finally:
        L.hb_destroy(ctx)


def write_mtx(g: Graph, symmetric: bool = True) -> bytes:
    """MatrixMarket pattern text of a CSR: "symmetric" writes each arc u -> v with
    u <= v once (the graph must hold v -> u for every such arc, as the generators' CSRs
    do), "general" every arc.  Reading it back gives the same CSR."""
    off = g.offsets.astype(np.int64)
    n = g.size()
    u = np.repeat(np.arange(n, dtype=np.int64), np.diff(off))
    v = g.neighbours.astype(np.int64)
    if symmetric:
        keep = u <= v
        u, v = u[keep], v[keep]
    head = (f"%%MatrixMarket matrix coordinate pattern {'symmetric' if symmetric else 'general'}\n"
            f"{n} {n} {u.size}\n").encode()
    if u.size == 0:
        return head
    w = len(str(n))

    def digits(x):  # right-aligned, space-padded decimal columns
        d = np.empty((x.size, w), np.uint8)
        y = x + 1
        for k in range(w - 1, -1, -1):
            d[:, k] = 48 + y % 10
            y //= 10
        lead = np.cumsum(d != 48, axis=1) == 0
        lead[:, -1] = False
        d[lead] = 32
        return d

    out = [head]
    step = 1 << 22
    for a in range(0, u.size, step):
        uu, vv = u[a:a + step], v[a:a + step]
        rows = np.concatenate([digits(uu), np.full((uu.size, 1), 32, np.uint8), digits(vv),
                               np.full((uu.size, 1), 10, np.uint8)], axis=1)
        out.append(rows.tobytes())
    return b"".join(out)
