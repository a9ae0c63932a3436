"""CPU tests of the C ABI: the library loads, exports every function that
include/radar_graph.h declares, and its host-side entry points (generator,
CSR builder, readers / writers, source pickers, error reporting) behave as
the reference's host code.  No GPU calls."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "radar_graph.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_function(rg):
    missing = [f for f in declared_functions() if not hasattr(rg.lib, f)]
    assert not missing, missing
    assert len(declared_functions()) >= 30


def test_version_and_error_text(rg):
    assert b"gfx950" in rg.lib.rg_version()
    with pytest.raises(rg.RGError) as e:
        rg.HostGraph.from_edges(0, [], [])
    assert e.value.code == rg.RG_EINVAL


def test_bad_input_file(rg, tmp_path):
    p = tmp_path / "bad.adj"
    p.write_text("NotAGraph\n3\n0\n")
    with pytest.raises(rg.RGError) as e:
        rg.HostGraph.read(p, rg.FMT_ADJ)
    assert e.value.code == rg.RG_EFORMAT
    p.write_text("AdjacencyGraph\n3\n2\n0\n1\n")   # too few words (IO.h:243-249)
    with pytest.raises(rg.RGError):
        rg.HostGraph.read(p, rg.FMT_ADJ)
    with pytest.raises(rg.RGError) as e:
        rg.HostGraph.read(tmp_path / "missing.adj", rg.FMT_ADJ)
    assert e.value.code == rg.RG_EIO


def test_reader_dedups_like_ligra(rg, orc, tmp_path):
    # Duplicates and self-loops in the text file are removed per list
    # (IO.h:432-467); the symmetric reader does not add reverse arcs.
    p = tmp_path / "g.adj"
    p.write_text("AdjacencyGraph\n4\n7\n0\n3\n5\n6\n1\n1\n0\n0\n2\n3\n1\n")
    g = rg.HostGraph.read(p, rg.FMT_ADJ, symmetric=True)
    off, nb = g.csr()
    # vertex 0: {1,1,0} -> {1}; 1: {0,2}; 2: {3}; 3: {1}.  The reference's
    # reader prints "Num Edges = 5" for this file.
    assert off.tolist() == [0, 1, 3, 4, 5]
    assert nb.tolist() == [1, 0, 2, 3, 1]


@pytest.mark.parametrize("fmt", ["adj", "bin", "el", "sg"])
@pytest.mark.parametrize("sym", [True, False])
def test_format_round_trip(rg, graphs, tmp_path, fmt, sym):
    g = graphs("rmat12s" if sym else "rmat12d")
    path = str(tmp_path / ("g." + fmt))
    code = {"adj": rg.FMT_ADJ, "bin": rg.FMT_ADJ_BINARY, "el": rg.FMT_EL, "sg": rg.FMT_SG}[fmt]
    g.write(path, code)
    h = rg.HostGraph.read(path, code, symmetric=sym)
    a, b = g.csr(), h.csr()
    if fmt == "el":
        # GAP's reader sizes the graph by the largest id seen (builder.h:209-210).
        n = len(b[0]) - 1
        assert np.array_equal(a[0][: n + 1], b[0]) and np.array_equal(a[1][: b[0][-1]], b[1])
    else:
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        if not sym:
            assert np.array_equal(g.csr(True)[1], h.csr(True)[1])


def test_directed_csc_is_transpose(rg, orc, graphs):
    g = graphs("rmat12d")
    off, nb = g.csr()
    toff, tnb = orc.transpose_csr(off, nb)
    assert np.array_equal(toff, g.offsets(True)) and np.array_equal(tnb, g.neighbors(True))


def test_from_csr_dedup_matches_numpy(rg, orc):
    rng = np.random.default_rng(7)
    n = 300
    s = rng.integers(0, n, 5000).astype(np.uint32)
    d = rng.integers(0, n, 5000).astype(np.uint32)
    g = rg.HostGraph.from_edges(n, s, d, symmetric=False)
    off, nb = orc.csr_from_edges(n, s, d, symmetric=False)
    assert np.array_equal(g.offsets(), off) and np.array_equal(g.neighbors(), nb)
    gs = rg.HostGraph.from_edges(n, s, d, symmetric=True)
    off, nb = orc.csr_from_edges(n, s, d, symmetric=True)
    assert np.array_equal(gs.offsets(), off) and np.array_equal(gs.neighbors(), nb)


def test_edge_cases(rg):
    # A graph with no edges; a single vertex; self-loop only.
    g = rg.HostGraph.from_edges(5, np.array([], np.uint32), np.array([], np.uint32))
    assert g.info() == (5, 0, True)
    g = rg.HostGraph.from_edges(1, np.array([0], np.uint32), np.array([0], np.uint32))
    assert g.m == 0
    with pytest.raises(rg.RGError):
        g.pick_sources(1)          # no vertex with an out-edge
    with pytest.raises(rg.RGError):
        rg.HostGraph.from_edges(3, np.array([5], np.uint32), np.array([0], np.uint32))


def test_bc_picker_alternates_orientation(rg):
    # Directed: odd-numbered BC sources are tested against in-degrees
    # because BC transposes once per source (BC-manyiter-v2.C:217).
    g = rg.HostGraph.generate(10, 4, symmetric=False)
    srcs = g.pick_bc_sources(6)
    out_deg = np.diff(g.offsets())
    in_deg = np.diff(g.offsets(True))
    for i, s in enumerate(srcs):
        assert (in_deg if i % 2 else out_deg)[s] > 0


def test_device_count_without_gpu(rg):
    assert rg.device_count() >= 0


# splitmix64 known answers (published outputs of Vigna's splitmix64.c for
# seeds 1234567 and 0): output k of a generator seeded with x is the
# finaliser of x + k * 0x9E3779B97F4A7C15, i.e. level_hash(v, level) with
# (level << 32 | v) = x + (k - 1) * 0x9E3779B97F4A7C15.
SPLITMIX = [(1234567, [6457827717110365317, 3203168211198807973, 9817491932198370423,
                       4593380528125082431, 16408922859458223821]),
            (0, [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC])]


@pytest.mark.parametrize("seed,outs", SPLITMIX)
def test_level_hash_known_answers(rg, orc, seed, outs):
    g = 0x9E3779B97F4A7C15
    for k, want in enumerate(outs):
        x = (seed + k * g) & ((1 << 64) - 1)
        v, lv = x & 0xFFFFFFFF, x >> 32
        assert rg.level_hash(v, lv) == want
        assert int(orc.level_hash(v, lv)) == want


def test_bfs_summary_by_hand(orc):
    # Path 0 -> 1 -> 2 plus an isolated vertex 3: levels 0, 1, 2, -1.
    off = np.array([0, 1, 2, 2, 2], np.uint64)
    lv = np.array([0, 1, 2, -1], np.int32)
    it, reached, edges, cs = orc.bfs_summary(off, lv)
    assert (it, reached, edges) == (3, 3, 2)
    want = sum(int(orc.level_hash(v, l)) for v, l in [(0, 0), (1, 1), (2, 2)]) % (1 << 64)
    assert cs == want


# ---- GAP's other text formats (reader.h:59-217) ----------------------------------
# A small graph written in each format must read back as the same CSR as its
# edge list through the reference's builder rules (dedup, no self-loops,
# num_nodes = max id + 1); ids are 1-based in METIS and Matrix Market,
# as written in DIMACS .gr.

EDGES = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 3), (1, 2), (4, 0)]


def _expect(orc, edges, sym, shift=0):
    src = np.array([u + shift for u, _ in edges])
    dst = np.array([v + shift for _, v in edges])
    n = int(max(src.max(), dst.max())) + 1
    return orc.csr_from_edges(n, src, dst, sym)


@pytest.mark.parametrize("sym", [True, False])
def test_read_wel(rg, orc, tmp_path, sym):
    p = tmp_path / "g.wel"
    p.write_text("".join(f"{u} {v} {3 * u + v}\n" for u, v in EDGES))
    hg = rg.HostGraph.read(str(p), rg.FMT_WEL, symmetric=sym)
    off, nb = _expect(orc, EDGES, sym)
    assert np.array_equal(hg.offsets(), off) and np.array_equal(hg.neighbors(), nb)


@pytest.mark.parametrize("sym", [True, False])
def test_read_dimacs_gr(rg, orc, tmp_path, sym):
    p = tmp_path / "g.gr"
    body = "c comment line\np sp 6 7\n" + "".join(f"a {u + 1} {v + 1} 7\n" for u, v in EDGES)
    p.write_text(body)
    hg = rg.HostGraph.read(str(p), rg.FMT_GR, symmetric=sym)
    off, nb = _expect(orc, EDGES, sym, shift=1)   # DIMACS ids kept as written
    assert np.array_equal(hg.offsets(), off) and np.array_equal(hg.neighbors(), nb)


@pytest.mark.parametrize("weighted", [False, True])
def test_read_metis(rg, orc, tmp_path, weighted):
    n = 5
    adj = [[] for _ in range(n)]
    for u, v in EDGES:
        adj[u].append(v)
    lines = ["% a METIS file", f"{n} {len(EDGES)}" + (" 1" if weighted else "")]
    for u in range(n):
        if weighted:
            lines.append(" ".join(f"{v + 1} 9" for v in adj[u]))
        else:
            lines.append(" ".join(str(v + 1) for v in adj[u]))
    p = tmp_path / "g.graph"
    p.write_text("\n".join(lines) + "\n")
    hg = rg.HostGraph.read(str(p), rg.FMT_METIS, symmetric=False)
    off, nb = _expect(orc, EDGES, False)
    assert np.array_equal(hg.offsets(), off) and np.array_equal(hg.neighbors(), nb)


@pytest.mark.parametrize("field,symmetry", [("pattern", "general"), ("real", "general"), ("integer", "symmetric")])
def test_read_matrix_market(rg, orc, tmp_path, field, symmetry):
    rows = [f"%%MatrixMarket matrix coordinate {field} {symmetry}", "% comment", f"5 5 {len(EDGES)}"]
    for u, v in EDGES:
        rows.append(f"{u + 1} {v + 1}" + ("" if field == "pattern" else " 2.5"))
    p = tmp_path / "g.mtx"
    p.write_text("\n".join(rows) + "\n")
    hg = rg.HostGraph.read(str(p), rg.FMT_MTX, symmetric=False)
    edges = EDGES + ([(v, u) for u, v in EDGES] if symmetry == "symmetric" else [])
    off, nb = _expect(orc, edges, False)
    assert np.array_equal(hg.offsets(), off) and np.array_equal(hg.neighbors(), nb)


def test_read_formats_reject_bad_headers(rg, tmp_path):
    bad = {"a.mtx": "%%MatrixMarket matrix array real general\n3 3\n",
           "b.mtx": "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 1\n",
           "c.mtx": "%%MatrixMarket matrix coordinate pattern general\n3 4 1\n1 2\n",
           "d.graph": "3 2 11\n2\n1\n\n",
           "e.wel": "0 1\n"}
    fmts = {"mtx": rg.FMT_MTX, "graph": rg.FMT_METIS, "wel": rg.FMT_WEL}
    for name, text in bad.items():
        p = tmp_path / name
        p.write_text(text)
        with pytest.raises(rg.RGError):
            rg.HostGraph.read(str(p), fmts[name.split(".")[1]], symmetric=False)


def test_format_of_suffix(rg):
    f = rg.lib.rg_format_of
    f.restype = rg.C.c_int
    f.argtypes = [rg.C.c_char_p]
    assert [f(x) for x in (b"g.el", b"g.wel", b"g.sg", b"g.gr", b"g.graph", b"g.mtx", b"g.adj")] == \
        [rg.FMT_EL, rg.FMT_WEL, rg.FMT_SG, rg.FMT_GR, rg.FMT_METIS, rg.FMT_MTX, rg.FMT_ADJ]


def _file_info(rg, path, fmt):
    import ctypes as C
    n, m, s, has = C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
    rc = rg.lib.rg_graph_file_info(os.fsencode(str(path)), fmt, C.byref(n), C.byref(m), C.byref(s), C.byref(has))
    return rc, (n.value, m.value, s.value, has.value)


@pytest.mark.parametrize("gname", ["rmat10s", "rmat12d"])
def test_graph_file_info(rg, graphs, tmp_path, gname):
    # rg_graph_file_info reads the binary headers on the host (no GPU): n, m,
    # the orientation (-1 for Ligra binary, whose files do not say) and
    # whether a new_ids sidecar (n uint32) exists; a sidecar of the wrong
    # size and a truncated .sg are refused.
    hg = graphs(gname)
    sg = tmp_path / "g.sg"
    hg.write(sg, rg.FMT_SG)
    assert _file_info(rg, sg, rg.FMT_SG) == (0, (hg.n, hg.m, 1 if hg.symmetric else 0, 0))
    np.arange(hg.n, dtype=np.uint32).tofile(str(sg) + ".newids")
    assert _file_info(rg, sg, rg.FMT_SG) == (0, (hg.n, hg.m, 1 if hg.symmetric else 0, 1))
    np.arange(hg.n - 1, dtype=np.uint32).tofile(str(sg) + ".newids")
    assert _file_info(rg, sg, rg.FMT_SG)[0] == rg.RG_EFORMAT
    os.remove(str(sg) + ".newids")
    binp = tmp_path / "g"
    hg.write(binp, rg.FMT_ADJ_BINARY)
    assert _file_info(rg, binp, rg.FMT_ADJ_BINARY) == (0, (hg.n, hg.m, -1, 0))
    with open(sg, "r+b") as f:
        f.truncate(40)
    assert _file_info(rg, sg, rg.FMT_SG)[0] == rg.RG_EFORMAT
    assert _file_info(rg, tmp_path / "missing.sg", rg.FMT_SG)[0] == rg.RG_EIO
    assert _file_info(rg, sg, rg.FMT_ADJ)[0] == rg.RG_EUNSUPPORTED
