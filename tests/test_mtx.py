"""Matrix Market reader (msbfs_mtx_parse / msbfs_mtx_read, the reference's
MtxReader in crates/core/src/mtx.rs): its own known answers
(tests/golden/mtx_kats.json), scipy.io.mmread on random files (an independent
parser; scipy keeps values, so the comparison applies the reference's rules
-- drop the diagonal, mirror symmetric files, deduplicate -- to scipy's
coordinates), write -> read round trips, and exact line numbers of errors in
files large enough to be parsed in several parallel chunks."""
import io
import json
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import gpu_bfs as gb

HERE = os.path.dirname(os.path.abspath(__file__))


def kats():
    with open(os.path.join(HERE, "golden", "mtx_kats.json")) as f:
        return json.load(f)


def neighbors(g, v):
    return g.col_indices[g.row_offsets[v]:g.row_offsets[v + 1]].tolist()


def expect_error(text, kind, **fields):
    with pytest.raises(gb.MtxParseError) as ei:
        gb.parse_mtx(text)
    e = ei.value
    assert e.kind == kind, (e.kind, str(e))
    for k, v in fields.items():
        assert getattr(e, k) == v, (k, getattr(e, k), v)
    return e


def canonical_from_coo(n, rows, cols, symmetric):
    """The reference's rules applied to coordinates read by another parser."""
    r, c = np.asarray(rows, np.int64), np.asarray(cols, np.int64)
    keep = r != c
    r, c = r[keep], c[keep]
    if symmetric:
        r, c = np.concatenate([r, c]), np.concatenate([c, r])
    m = sp.csr_matrix((np.ones(len(r), np.int8), (r, c)), shape=(n, n))
    m.sum_duplicates()
    m.sort_indices()
    return m.indptr.astype(np.uint32), m.indices.astype(np.uint32)


# ---------------------------------------------------------------- known answers

def test_reference_header_variants():
    for case in kats()["header_variants"]["cases"]:
        text = case["banner"] + "\n2 2 1\n1 2\n"
        if case["ok"]:
            g = gb.parse_mtx(text)
            assert g.num_edges == (2 if case["symmetric"] else 1)
        else:
            expect_error(text, case["kind"], line_number=1)


@pytest.mark.parametrize("name", ["symmetric_pattern", "self_loops_and_dedup"])
def test_reference_graph_kats(name):
    k = kats()[name]
    g = gb.parse_mtx(k["text"])
    assert g.num_vertices == k["num_vertices"] and g.num_edges == k["num_edges"]
    assert [neighbors(g, v) for v in range(g.num_vertices)] == k["neighbors"]
    if k.get("symmetric"):
        assert g.is_symmetric()
    assert gb.lib().msbfs_csr_validate(g.num_vertices, g.num_edges,
                                       g.row_offsets.ctypes.data, g.col_indices.ctypes.data) == 0


def test_reference_error_kats():
    k = kats()["zero_index"]
    expect_error(k["text"], "ZeroIndex", line_number=3, row=0, col=1)
    k = kats()["out_of_bounds"]
    e = expect_error(k["text"], "IndexOutOfBounds", line_number=3, row=4, col=1, max_rows=3,
                     max_cols=3)
    assert str(e) == ("Coordinate out of bounds at line 3: entry (4, 1) exceeds matrix "
                      "dimensions (3, 3)")


def test_reference_luxembourg_mock():
    k = kats()["luxembourg_osm_mock"]
    g = gb.parse_mtx(k["text"])
    assert g.num_vertices == 6 and g.num_edges == 10
    deg = g.degrees()
    assert int(deg.max()) == k["max_degree"] and deg[0] == 1 and deg[1] == 2
    assert g.is_symmetric()


# ---------------------------------------------------------------- error texts

def test_error_messages_follow_the_reference():
    e = expect_error("", "InvalidHeader", line_number=1)
    assert str(e) == ("Invalid Matrix Market header at line 1: Empty input: no MatrixMarket "
                      "header found")
    e = expect_error("\n\n", "InvalidHeader", line_number=3)
    e = expect_error("%MatrixMarket matrix coordinate pattern general\n", "InvalidHeader")
    assert "Missing '%%MatrixMarket' banner" in str(e)
    e = expect_error("%%MatrixMarket vector coordinate pattern general\n", "InvalidHeader")
    assert "Unsupported object 'vector'" in str(e)
    expect_error("%%MatrixMarket matrix coordinate pattern\n", "InvalidHeader")
    expect_error("%%MatrixMarket matrix sparse pattern general\n", "InvalidHeader")
    expect_error("%%MatrixMarket matrix coordinate boolean general\n", "InvalidHeader")
    expect_error("%%MatrixMarket matrix coordinate pattern upper\n", "InvalidHeader")
    e = expect_error("%%MatrixMarket matrix coordinate pattern general\n% c\n\n",
                     "InvalidDimensions", line_number=4)
    assert "Unexpected end of file while searching for dimension line" in str(e)
    e = expect_error("%%MatrixMarket matrix coordinate pattern general\n3 x 1\n",
                     "InvalidDimensions", line_number=2)
    assert str(e) == ("Invalid dimensions at line 2: Failed to parse cols count 'x': invalid "
                      "digit found in string")
    e = expect_error("%%MatrixMarket matrix coordinate pattern general\n3 3 99999999999999999999\n",
                     "InvalidDimensions")
    assert "number too large to fit in target type" in str(e)
    e = expect_error("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n\n7\n",
                     "InvalidCoordinateLine", line_number=4)
    assert str(e) == "Invalid coordinate entry at line 4: '7'"
    expect_error("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n-1 2\n",
                 "InvalidCoordinateLine", line_number=3)
    expect_error("%%MatrixMarket matrix coordinate pattern skew-symmetric\n2 3 1\n",
                 "InvalidDimensions", line_number=2)
    with pytest.raises(gb.MtxParseError) as ei:
        gb.read_mtx("/nonexistent/graph.mtx")
    assert ei.value.kind == "Io"


def test_lenient_forms_the_reference_accepts():
    # CRLF line ends, '+' signs, tabs, values of any field, comments between entries,
    # an entry count that does not match, hermitian / skew-symmetric mirrored
    text = ("%%MatrixMarket Matrix Coordinate Complex Hermitian\r\n% c\r\n\r\n"
            "4\t4  9\r\n+1 2 0.5 -1.5\r\n% mid comment\r\n3 4 1e3 2\r\n2 2 7 7\r\n")
    g = gb.parse_mtx(text)
    assert g.num_vertices == 4 and g.num_edges == 4
    assert [neighbors(g, v) for v in range(4)] == [[1], [0], [3], [2]]
    g = gb.parse_mtx("%%MatrixMarket matrix coordinate integer general\n2 5 1\n1 5 3\n")
    assert g.num_vertices == 5 and neighbors(g, 0) == [4]


# ---------------------------------------------------------------- against scipy

@pytest.mark.parametrize("seed,symmetric,field", [(1, False, "pattern"), (2, True, "pattern"),
                                                  (3, False, "real"), (4, True, "integer")])
def test_matches_scipy_mmread(tmp_path, seed, symmetric, field):
    rng = np.random.default_rng(seed)
    n, m = 300, 2500
    r = rng.integers(0, n, m)
    c = rng.integers(0, n, m)
    if symmetric:  # symmetric files store one triangle (incl. the diagonal)
        r, c = np.maximum(r, c), np.minimum(r, c)
    path = tmp_path / "g.mtx"
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} "
                f"{'symmetric' if symmetric else 'general'}\n% seeded\n{n} {n} {m}\n")
        for i in range(m):
            val = "" if field == "pattern" else (f" {rng.integers(-9, 9)}" if field == "integer"
                                                 else f" {rng.normal():.6f}")
            f.write(f"{r[i] + 1} {c[i] + 1}{val}\n")
    ours = gb.read_mtx(path)
    ref = scipy.io.mmread(str(path))
    coo = ref.tocoo() if sp.issparse(ref) else sp.coo_matrix(ref)
    rows, cols = coo.row, coo.col
    if symmetric:  # scipy already mirrored; undo so both sides apply the same rule
        lower_tri = rows >= cols
        rows, cols = rows[lower_tri], cols[lower_tri]
    ro, ci = canonical_from_coo(n, rows, cols, symmetric)
    assert ours.num_vertices == n
    assert np.array_equal(ours.row_offsets, ro) and np.array_equal(ours.col_indices, ci)


# ---------------------------------------------------------------- round trips

def test_round_trip_generated_graphs(tmp_path):
    graphs = [gb.generate_road_like(5000, 70, 300, seed=2), gb.generate_rmat(12, 8, directed=True),
              gb.generate_grid2d(40, 30)]
    for i, g in enumerate(graphs):
        for symmetric in ([False, True] if g.is_symmetric() else [False]):
            p = tmp_path / f"g{i}{int(symmetric)}.mtx"
            gb.write_mtx(g, p, symmetric=symmetric)
            h = gb.read_mtx(p)
            assert h.num_vertices == g.num_vertices and h.num_edges == g.num_edges
            assert np.array_equal(h.row_offsets, g.row_offsets)
            assert np.array_equal(h.col_indices, g.col_indices)


def test_parallel_chunks_keep_order_and_line_numbers(tmp_path):
    # ~16 MB body: parsed in several chunks; the result must equal the
    # edge list built directly, and an error planted deep in the file must
    # carry its exact line number
    rng = np.random.default_rng(7)
    n, m = 200_000, 1_500_000
    r = rng.integers(1, n + 1, m)
    c = rng.integers(1, n + 1, m)
    buf = io.StringIO()
    buf.write("%%MatrixMarket matrix coordinate pattern general\n% big\n")
    buf.write(f"{n} {n} {m}\n")
    lines = np.char.add(np.char.add(r.astype(str), " "), c.astype(str))
    body = "\n".join(lines.tolist()) + "\n"
    text = buf.getvalue() + body
    g = gb.parse_mtx(text)
    ro, ci = canonical_from_coo(n, r - 1, c - 1, False)
    assert np.array_equal(g.row_offsets, ro) and np.array_equal(g.col_indices, ci)
    # plant a zero index at entry k (file line 3 + k + 1)
    k = 1_234_567
    bad = lines.copy()
    bad[k] = "0 5"
    text = buf.getvalue() + "\n".join(bad.tolist()) + "\n"
    expect_error(text, "ZeroIndex", line_number=3 + k + 1, row=0, col=5)
    # two errors: the earlier one in file order wins
    bad[100] = "1 999999999"
    text = buf.getvalue() + "\n".join(bad.tolist()) + "\n"
    expect_error(text, "IndexOutOfBounds", line_number=3 + 100 + 1, row=1, col=999999999)
