"""Host-side CSR graph, mirroring the part of the reference's GraphMemoryMap
that the HyperBall path walks (src/graph/mod.rs:569-595 index_node/neighbours,
:648-658 size/width), plus the seeded synthetic generators of
include/sparkx_hb_graphgen.h."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class Graph:
    """Symmetric or directed CSR: offsets[n+1] and neighbours[arcs] (u32 or u64)."""

    def __init__(self, offsets: np.ndarray, neighbours: np.ndarray, name: str = "graph"):
        self.offsets = np.ascontiguousarray(offsets)
        self.neighbours = np.ascontiguousarray(neighbours)
        if self.offsets.dtype not in (np.uint32, np.uint64) or self.neighbours.dtype not in (np.uint32, np.uint64):
            raise TypeError("offsets / neighbours must be uint32 or uint64")
        self.name = name

    # GraphMemoryMap::size (graph/mod.rs:648): number of nodes
    def size(self) -> int:
        return int(self.offsets.size - 1)

    # GraphMemoryMap::width (graph/mod.rs:656): number of arcs
    def width(self) -> int:
        return int(self.neighbours.size)

    # GraphMemoryMap::index_node (graph/mod.rs:569-575)
    def index_node(self, u: int) -> range:
        if not 0 <= u < self.size():
            raise IndexError(f"node {u} out of range (|V| = {self.size()})")
        return range(int(self.offsets[u]), int(self.offsets[u + 1]))

    # GraphMemoryMap::neighbours (graph/mod.rs:580-595): Err for u >= |V|
    def neighbours_of(self, u: int) -> np.ndarray:
        r = self.index_node(u)
        return self.neighbours[r.start:r.stop]

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets.astype(np.int64))

    # ---- seeded generators ---------------------------------------------------
    @staticmethod
    def _take(handle: C.c_void_p, name: str) -> "Graph":
        L = _lib.load()
        try:
            n = L.hb_graph_nodes(handle)
            arcs = L.hb_graph_arcs(handle)
            off = np.ctypeslib.as_array(L.hb_graph_offsets(handle), shape=(n + 1,)).copy()
            nb = (np.ctypeslib.as_array(L.hb_graph_neighbours(handle), shape=(arcs,)).copy()
                  if arcs else np.zeros(0, np.uint32))
        finally:
            L.hb_graph_free(handle)
        return Graph(off, nb, name)

    @staticmethod
    def erdos_renyi(n: int, samples: int, seed: int = 1, threads: int = 0) -> "Graph":
        h = C.c_void_p()
        _lib.check(_lib.load().hb_graph_er(C.byref(h), n, samples, seed, threads))
        return Graph._take(h, f"er_n{n}_m{samples}_s{seed}")

    @staticmethod
    def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1, threads: int = 0) -> "Graph":
        h = C.c_void_p()
        samples = edge_factor << scale
        _lib.check(_lib.load().hb_graph_rmat(C.byref(h), scale, samples, a, b, c, seed, threads))
        return Graph._take(h, f"rmat{scale}_ef{edge_factor}_s{seed}")

    @staticmethod
    def kmer_dbg(genome_len: int = 4_000_000, copies: int = 8, sub_rate: float = 0.005, k: int = 21,
                 seed: int = 1, threads: int = 0) -> "Graph":
        h = C.c_void_p()
        _lib.check(_lib.load().hb_graph_kmer(C.byref(h), genome_len, copies, sub_rate, k, seed, threads))
        return Graph._take(h, f"kmer_L{genome_len}_c{copies}_k{k}_s{seed}")
