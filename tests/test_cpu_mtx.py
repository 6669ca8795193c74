"""The oracle's MatrixMarket ingestion (oracle/mtx_oracle.c) pinned on the CPU:

- against the oracle's independent de Bruijn CSR builder on the text gen_dbg_mtx.py
  writes (the fixture family whose exact closeness the reference ships);
- against scipy.io.mmread (a library reader: it mirrors off-diagonal entries of
  symmetric and skew-symmetric files and keeps duplicates, as the reference's
  emission does) on random files;
- on the reference's token rules (parser_mtx/producers.rs): blank and '%' lines,
  CRLF, tabs, a leading '+', ignored trailing tokens, one-token lines skipped, and
  each error with its line number."""
import io

import numpy as np
import pytest


def dbg_text(sigma: int, k: int) -> bytes:
    """The text gen_dbg_mtx.py (emit_dbg_mtx, lines 4-22) writes for (sigma, k)."""
    n = sigma ** k
    base = sigma ** (k - 1)
    out = [b"%%MatrixMarket matrix coordinate pattern symmetric\n",
           f"% de Bruijn graph: sigma={sigma}, k={k}; nodes={n}, edges={sigma ** (k + 1)}\n".encode(),
           f"{n} {n} {sigma ** (k + 1)}\n".encode()]
    for u in range(n):
        bt = (u % base) * sigma
        out.append("".join(f"{u + 1} {bt + c + 1}\n" for c in range(sigma)).encode())
    return b"".join(out)


def csr_from_pairs(n, rows, cols):
    order = np.lexsort((cols, rows))
    rows, cols = np.asarray(rows)[order], np.asarray(cols)[order]
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, rows + 1, 1)
    return np.cumsum(off).astype(np.uint64), cols.astype(np.uint32)


@pytest.mark.parametrize("sigma,k", [(4, 3), (20, 2), (4, 7), (3, 1)])
def test_mtx_oracle_matches_dbg_builder(oracle, sigma, k):
    off, nb, info = oracle.parse_mtx(dbg_text(sigma, k))
    ref_off, ref_nb = oracle.dbg_mtx_csr(sigma, k)
    assert info.n == sigma ** k and info.entries == sigma ** (k + 1) and info.symmetric
    assert np.array_equal(off, ref_off) and np.array_equal(nb, ref_nb)


def random_mtx(rng, n, m, field, symmetry):
    i = rng.integers(1, n + 1, m)
    j = rng.integers(1, n + 1, m)
    if symmetry == "skew-symmetric":
        keep = i != j  # a skew-symmetric matrix has a zero diagonal
        i, j = i[keep], j[keep]
    lines = [f"%%MatrixMarket matrix coordinate {field} {symmetry}\n", "% generated\n", "%\n",
             f"{n} {n} {len(i)}\n"]
    for a, b in zip(i, j):
        lines.append(f"{a} {b}\n" if field == "pattern" else f"{a} {b} {int(rng.integers(-9, 10))}\n")
    return "".join(lines).encode()


@pytest.mark.parametrize("symmetry", ["general", "symmetric", "skew-symmetric"])
@pytest.mark.parametrize("field", ["pattern", "integer"])
def test_mtx_oracle_matches_scipy(oracle, symmetry, field):
    import scipy.io
    rng = np.random.default_rng(len(symmetry) * 7 + len(field))
    for n, m in ((1, 3), (7, 40), (300, 2500)):
        text = random_mtx(rng, n, m, field, symmetry)
        off, nb, info = oracle.parse_mtx(text)
        coo = scipy.io.mmread(io.BytesIO(text))
        ref_off, ref_nb = csr_from_pairs(n, coo.row.astype(np.int64), coo.col.astype(np.int64))
        assert np.array_equal(off, ref_off) and np.array_equal(nb, ref_nb), (n, m)
        assert info.arcs == coo.row.size


HDR = b"%%MatrixMarket matrix coordinate pattern general\n"


def test_mtx_oracle_token_rules(oracle):
    text = (b"%%MatrixMarket MATRIX Coordinate Pattern GENERAL extra\r\n"
            b"\r\n% comment\r\n  4\t4   5  \r\n"
            b"1 2\r\n\t+3\t4 99 junk\r\n% inside\n\n2\n4 4\n1 2")  # "2": one token, skipped
    off, nb, info = oracle.parse_mtx(text)
    assert info.n == 4 and info.nnz == 5 and not info.symmetric
    assert info.entries == 4 and info.arcs == 4
    assert off.tolist() == [0, 2, 2, 3, 4] and nb.tolist() == [1, 1, 3, 3]
    # no trailing newline after the size line: no entries
    off, nb, info = oracle.parse_mtx(HDR + b"3 3 0")
    assert off.tolist() == [0, 0, 0, 0] and info.arcs == 0


@pytest.mark.parametrize("text,code,line", [
    (b"", -1, 0),
    (b"%%matrixmarket matrix coordinate pattern general\n1 1 0\n", -1, 0),  # prefix is case-sensitive
    (b"%MatrixMarket matrix coordinate pattern general\n1 1 0\n", -1, 0),
    (b"%%MatrixMarket matrix coordinate pattern\n1 1 0\n", -2, 0),
    (b"%%MatrixMarket matrix array real general\n1 1 0\n", -3, 0),
    (b"%%MatrixMarket matrix coordinate real general\n1 1 0\n", -3, 0),
    (HDR + b"% only comments\n\n", -4, 0),
    (HDR + b"3 3\n", -5, 2),
    (HDR + b"3 x 1\n", -5, 2),
    (HDR + b"3 4 1\n1 1\n", -6, 2),
    (HDR + b"3 3 2\n1 2\n2 -1\n", -7, 4),
    (HDR + b"3 3 2\n1 2\nq 1\n", -7, 4),
    (HDR + b"3 3 2\n1 2\n99999999999999999999 1\n", -7, 4),
    (HDR + b"3 3 2\n% c\n0 1\n", -8, 4),
    (HDR + b"3 3 2\n1 2\n1 4\n", -9, 4),
])
def test_mtx_oracle_errors(oracle, text, code, line):
    with pytest.raises(oracle.MtxError) as ei:
        oracle.parse_mtx(text)
    assert ei.value.code == code
    if line:
        assert ei.value.line == line


def test_write_mtx_round_trip(oracle):
    """sparkx_amd.ingest.write_mtx (host text writer) read back by the oracle."""
    from sparkx_amd import Graph
    from sparkx_amd.ingest import write_mtx
    for g, sym in ((Graph.rmat(12, 16, seed=3), True), (Graph.erdos_renyi(2000, 7000, seed=2), False),
                   (Graph(np.zeros(6, np.uint64), np.zeros(0, np.uint32)), True)):
        off, nb, info = oracle.parse_mtx(write_mtx(g, symmetric=sym))
        assert np.array_equal(off, g.offsets.astype(np.uint64))
        assert np.array_equal(nb, g.neighbours.astype(np.uint32))
