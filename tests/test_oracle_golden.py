"""Pins the oracle (and the product's host-side generator / CSR builder /
source pickers) against golden fixtures written by the reference build
(tests/golden/make_golden.py).  CPU only.

The fixtures were captured with one OpenMP thread, so FP64 arrays are
compared bit-for-bit: the oracle visits vertices in the order such a run
does.  BFS parents from the reference are a valid tree; the levels they
imply are compared exactly.
"""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.fixture(scope="module")
def meta():
    with open(os.path.join(GOLD, "golden.json")) as f:
        return json.load(f)


def load(name):
    return dict(np.load(os.path.join(GOLD, f"ligra_{name}.npz")))


def levels_from_parents(parents, src):
    """Levels implied by a parent array (root = own parent)."""
    n = len(parents)
    lv = np.full(n, -1, np.int64)
    lv[src] = 0
    none = np.uint32(0xFFFFFFFF)
    for v in range(n):
        if lv[v] >= 0 or parents[v] == none:
            continue
        path = []
        u = v
        while lv[u] < 0:
            path.append(u)
            u = int(parents[u])
        d = lv[u]
        for w in reversed(path):
            d += 1
            lv[w] = d
    return lv


# ---- generator, builder, pickers -------------------------------------------

@pytest.mark.parametrize("scale,deg,uni", [(10, 16, 0), (10, 16, 1), (12, 8, 0)])
def test_generator_matches_gap(rg, meta, scale, deg, uni):
    s, d = rg.generate_edges(scale, deg, bool(uni))
    el = np.stack([s.astype(np.int32), d.astype(np.int32)], axis=1)
    g = meta["gap"][f"el_{scale}_{deg}_{uni}"]
    assert len(el) == g["count"]
    assert el[:8].tolist() == g["head"]
    assert sha(el) == g["sha"]


@pytest.mark.parametrize("scale,deg,uni", [(10, 16, 0), (10, 16, 1), (12, 8, 0)])
def test_csr_builder_matches_gap(rg, meta, scale, deg, uni):
    hg = rg.HostGraph.generate(scale, deg, uniform=bool(uni), symmetric=True)
    off, nb = hg.csr()
    g = meta["gap"][f"csr_{scale}_{deg}_{uni}"]
    assert (hg.n, hg.m) == (g["n"], g["m"])
    assert sha(off, nb) == g["sha"]


@pytest.mark.parametrize("scale,deg,uni", [(10, 16, 0), (10, 16, 1), (12, 8, 0)])
def test_oracle_relabel_matches_gap(rg, orc, meta, scale, deg, uni):
    # GAP RelabelByDegree (builder.h:246-282): (degree, id) descending, sorted lists.
    hg = rg.HostGraph.generate(scale, deg, uniform=bool(uni), symmetric=True)
    off, nb = hg.csr()
    ids = orc.degree_order(np.diff(off).astype(np.uint32))
    noff, nnb = orc.relabel_csr(off, nb, ids, strict=True)
    assert sha(noff, nnb) == meta["gap"][f"relabel_{scale}_{deg}_{uni}"]["sha"]


@pytest.mark.parametrize("scale,deg,uni", [(10, 16, 0), (10, 16, 1), (12, 8, 0)])
def test_source_picker_matches_gap(rg, meta, scale, deg, uni):
    hg = rg.HostGraph.generate(scale, deg, uniform=bool(uni), symmetric=True)
    assert hg.pick_sources(64).astype(np.int64).tolist() == meta["gap"][f"sources_{scale}_{deg}_{uni}"]


def test_hash_int_matches(rg, orc):
    # hashInt(ulong) first values at n = 2^16 (Radii's sources).
    assert [rg.hash_int(i) % 65536 for i in range(4)] == [orc.hash_int(i) % 65536 for i in range(4)]
    assert [rg.hash_int(i) % 65536 for i in range(4)] == [23697, 43106, 14515, 36764]


# ---- Ligra apps ---------------------------------------------------------------

GRAPHS = [("rmat10s", True), ("rmat10d", False)]


def _graph(rg, name):
    sym = name.endswith("s")
    return rg.HostGraph.generate(10, 16, symmetric=sym)


def _relabelled(orc, hg, ids):
    off, nb = hg.csr()
    noff, nnb = orc.relabel_csr(off, nb, ids, strict=False)
    if hg.symmetric:
        return noff, nnb, noff, nnb
    ioff, inb = hg.csr(in_edges=True)
    nioff, ninb = orc.relabel_csr(ioff, inb, ids, strict=False)
    return noff, nnb, nioff, ninb


@pytest.mark.parametrize("name,sym", GRAPHS)
def test_graph_fixture_is_same_graph(rg, name, sym):
    hg = _graph(rg, name)
    off, nb = hg.csr()
    assert bytes(load(name)["csr_sha"]).hex() == sha(off, nb)


@pytest.mark.parametrize("name,sym", GRAPHS)
def test_pagerank_bitexact(rg, orc, meta, name, sym):
    hg = _graph(rg, name)
    off, nb = hg.csr()
    fx = load(name)
    r, it, l1 = orc.pagerank(off, nb, 20)
    assert np.array_equal(r, fx["pno_pr_ranks"])
    s = meta["ligra"][name]["pno_pr"]
    assert it == s["iters"]
    assert abs(l1 - s["l1"]) <= 1e-5 * abs(s["l1"])   # printed with 6 digits


@pytest.mark.parametrize("name,sym", GRAPHS)
def test_pagerank_delta_bitexact(rg, orc, meta, name, sym):
    # PageRankDelta.C run by the reference (one thread): final p, printed
    # Num Iters and L1_Norm; and on the degree-sorted graph (degree-sorting
    # tree, -preprocess 1).
    hg = _graph(rg, name)
    off, nb = hg.csr()
    fx = load(name)
    r, it, l1 = orc.pagerank_delta(off, nb, 20)
    assert np.array_equal(r, fx["pno_prd_ranks"])
    s = meta["ligra"][name]["pno_prd"]
    assert it == s["iters"] and abs(l1 - s["l1"]) <= 1e-5 * abs(s["l1"])
    off2, nb2, _, _ = _relabelled(orc, hg, fx["ds_new_ids"])
    r2, it2, _ = orc.pagerank_delta(off2, nb2, 20)
    assert np.array_equal(r2, fx["ds_prd_ranks"])
    assert it2 == meta["ligra"][name]["ds_prd"]["iters"]


def test_pagerank_delta_closed_forms(orc):
    # Directed cycle: round 1 gives every vertex nghSum = 1/n, so p = 1/n
    # and Delta = 0.85/n + 0.15/n - 1/n (zero up to rounding): converged in
    # one round.  A directed star 0 -> {1..7} does not converge in two
    # rounds, so the count is maxiters + 1 (PageRankDelta.C:131), and after
    # round 1 the leaves hold 0.85/(8*7) + 0.15/8, the centre 0.15/8.
    n = 8
    off = np.arange(n + 1, dtype=np.uint64)
    nb = ((np.arange(n) + 1) % n).astype(np.uint32)
    r, it, l1 = orc.pagerank_delta(off, nb, 20)
    assert np.allclose(r, 1.0 / n, rtol=0, atol=1e-15) and it == 1 and l1 < 1e-7
    off = np.array([0] + [7] * 8, np.uint64)
    nb = np.arange(1, 8, dtype=np.uint32)
    r, it, l1 = orc.pagerank_delta(off, nb, 1)
    assert it == 2 and l1 > 1e-7
    assert np.isclose(r[0], 0.15 / 8, rtol=1e-15) and np.allclose(r[1:], 0.85 / 56 + 0.15 / 8, rtol=1e-15)


@pytest.mark.parametrize("name,sym", GRAPHS)
def test_pagerank_pull_bitexact(rg, orc, meta, name, sym):
    hg = _graph(rg, name)
    off, nb = hg.csr()
    ioff, inb = hg.csr(in_edges=True)
    fx = load(name)
    r, it, l1 = orc.pagerank_pull(off, ioff, inb, 20)
    assert np.array_equal(r, fx["pp_pr_ranks"])
    assert it == meta["ligra"][name]["pp_pr"]["iters"]


@pytest.mark.parametrize("name,sym", GRAPHS)
@pytest.mark.parametrize("var", ["pno", "pp"])
def test_bfs_levels(rg, orc, meta, name, sym, var):
    hg = _graph(rg, name)
    off, nb = hg.csr()
    fx = load(name)
    src = int(fx["bfs_sources"][0])
    par = fx[f"{var}_bfs_parents"]
    lv, ecc = orc.bfs_levels(off, nb, src)
    assert orc.check_parents(off, nb, src, lv, par) == 0
    assert np.array_equal(levels_from_parents(par, src), lv)
    s = meta["ligra"][name][f"{var}_bfs"]
    assert s["iters"] == ecc + 1
    assert s["reached"] == int((lv >= 0).sum())


@pytest.mark.parametrize("name,sym", GRAPHS)
@pytest.mark.parametrize("var", ["pno", "pp"])
def test_radii_exact(rg, orc, meta, name, sym, var):
    hg = _graph(rg, name)
    off, nb = hg.csr()
    rad, rounds = orc.radii(off, nb, rg.radii_starts(hg.n))
    assert np.array_equal(rad, load(name)[f"{var}_radii"])
    assert rounds == meta["ligra"][name][f"{var}_radii"]["rounds"]


def _same_fp(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


@pytest.mark.parametrize("name,sym", GRAPHS)
def test_bc_bitexact(rg, orc, meta, name, sym):
    hg = _graph(rg, name)
    off, nb = hg.csr()
    ioff, inb = hg.csr(in_edges=True)
    srcs = hg.pick_bc_sources(4)
    deps, comb = orc.bc(off, nb, ioff, inb, srcs)
    assert _same_fp(deps, load(name)["pno_bc_deps"])
    assert comb == meta["ligra"][name]["pno_bc"]["combined"]


# ---- DegreeSorting (degree-sorting tree, relaxed relabel) ------------------

@pytest.mark.parametrize("name,sym", GRAPHS)
def test_degree_sort_ids(rg, orc, name, sym):
    hg = _graph(rg, name)
    fx = load(name)
    deg = np.diff(hg.offsets(in_edges=True)).astype(np.uint32)  # -preprocess 1: in-degree
    assert np.array_equal(orc.degree_order(deg), fx["ds_new_ids"])


@pytest.mark.parametrize("name,sym", GRAPHS)
def test_degree_sorted_apps(rg, orc, meta, name, sym):
    hg = _graph(rg, name)
    fx = load(name)
    ids = fx["ds_new_ids"]
    off, nb, ioff, inb = _relabelled(orc, hg, ids)
    r, it, _ = orc.pagerank(off, nb, 20)
    assert np.array_equal(r, fx["ds_pr_ranks"])
    src = int(ids[int(fx["bfs_sources"][0])])
    lv, ecc = orc.bfs_levels(off, nb, src)
    assert orc.check_parents(off, nb, src, lv, fx["ds_bfs_parents"]) == 0
    starts = ids[rg.radii_starts(hg.n)]
    rad, _ = orc.radii(off, nb, starts)
    assert np.array_equal(rad, fx["ds_radii"])
    srcs = ids[hg.pick_bc_sources(4)]
    deps, comb = orc.bc(off, nb, ioff, inb, srcs)
    assert _same_fp(deps, fx["ds_bc_deps"])
    assert comb == meta["ligra"][name]["ds_bc"]["combined"]


# ---- triangle counting ------------------------------------------------------------

@pytest.mark.parametrize("name,scale", [("rmat10s", 10), ("rmat12s", 12)])
def test_tc_totals(rg, orc, meta, name, scale):
    hg = rg.HostGraph.generate(scale, 16, symmetric=True)
    off, nb = hg.csr()
    tot, pv = orc.tc(off, nb, per_vertex=True)
    assert tot == meta["tc"][name]["pno"] == meta["tc"][name]["ds"]
    assert tot == orc.tc_verify(off, nb)
    assert int(pv.sum()) == 3 * tot   # tc_local.cc:84


def test_check_parents_catches_planted_errors(orc):
    # Star 0 - {1..4} plus the path 4 - 5: every planted mistake (wrong
    # root, parent a level too high, parent without an arc, unreached vertex
    # with a parent) must be reported; the correct tree passes.
    src = np.array([0, 0, 0, 0, 4])
    dst = np.array([1, 2, 3, 4, 5])
    off, nbr = orc.csr_from_edges(7, src, dst, symmetric=True)
    lv, ecc = orc.bfs_levels(off, nbr, 0)
    assert list(lv) == [0, 1, 1, 1, 1, 2, -1] and ecc == 2
    good = np.array([0, 0, 0, 0, 0, 4, 0xFFFFFFFF], np.uint32)
    assert orc.check_parents(off, nbr, 0, lv, good) == 0
    for v, p in [(0, 1), (5, 0), (5, 3), (2, 1), (6, 5)]:
        bad = good.copy()
        bad[v] = p
        assert orc.check_parents(off, nbr, 0, lv, bad) == 1 + v, (v, p)


def test_identify_hubs_closed_forms(orc):
    # hubdup IO.h:759,779: hubs have in-degree above the num_hubs-th largest
    # (0-based), ties at that value excluded.
    deg = np.array([5, 5, 3, 3, 1, 0], np.uint32)
    hub, c = orc.identify_hubs(deg, 2)      # threshold 3
    assert c == 2 and hub.tolist() == [True, True, False, False, False, False]
    hub, c = orc.identify_hubs(deg, 1)      # threshold 5: the tie leaves no hub
    assert c == 0 and not hub.any()
    hub, c = orc.identify_hubs(deg, 4)      # threshold 1
    assert c == 4 and hub.tolist() == [True] * 4 + [False] * 2
    hub, c = orc.identify_hubs(deg, 0)      # threshold 5
    assert c == 0
    # num_hubs >= n (out of range in the reference): every vertex of
    # positive degree, as the library's selection does.
    hub, c = orc.identify_hubs(deg, 6)
    assert c == 5 and hub.tolist() == [True] * 5 + [False]


@pytest.mark.parametrize("seed", range(6))
def test_identify_hubs_characterisation(orc, seed):
    # For every k the hubs H are the unique set with: H an upper set of the
    # degrees (every hub above every non-hub), |H| <= k, and the largest
    # non-hub degree t held by more than k vertices counting those above
    # it (so no larger upper set within k exists).
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    deg = rng.zipf(1.8, size=n).clip(0, 5000).astype(np.uint32) * rng.integers(0, 2, size=n).astype(np.uint32)
    for k in sorted({0, 1, 2, n // 3, n - 1, n, n + 5, int(rng.integers(0, n + 1))}):
        hub, c = orc.identify_hubs(deg, k)
        assert c == int(hub.sum())
        if k >= n:
            assert np.array_equal(hub, deg > 0)
            continue
        assert c <= k
        if c:
            assert deg[hub].min() > (deg[~hub].max() if (~hub).any() else -1)
        if (~hub).any():
            t = deg[~hub].max()
            assert int((deg >= t).sum()) > k
