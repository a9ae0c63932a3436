"""GPU parity: every app through the C ABI against the oracle on seeded
graphs.  Integers (BFS levels, Radii) bit-exact; BFS parents a valid
shortest-path tree; FP64 (PageRank, BC) within 1e-10 relative L-inf, the
only difference being atomic summation order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10   # relative L-inf bound for FP64 results
VARIANTS = ["push", "hubdup", "radar"]


def rel_linf(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    assert np.array_equal(nan_a, nan_b), "NaN positions differ"
    a, b = a[~nan_a], b[~nan_b]
    if a.size == 0:
        return 0.0
    assert np.array_equal(np.isinf(a), np.isinf(b))
    fin = np.isfinite(a)
    assert np.array_equal(a[~fin], b[~fin])
    a, b = a[fin], b[fin]
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12s"])
@pytest.mark.parametrize("variant", VARIANTS + ["pull", "blocked"])
def test_pagerank(rg, orc, ctx, graphs, gname, variant):
    hg = graphs(gname)
    off, nbr = hg.csr()
    print(f"pagerank {gname} {variant}", flush=True)
    dg = rg.DeviceGraph(ctx, hg)
    res = dg.pagerank(20, rg.opts(variant, hubs=256))
    if variant == "pull":
        ioff, inbr = hg.csr(in_edges=True)
        ref, it, l1 = orc.pagerank_pull(off, ioff, inbr, 20)
    else:
        ref, it, l1 = orc.pagerank(off, nbr, 20)
    assert res.iters == it
    assert rel_linf(res.ranks, ref) < FP_TOL
    # L1 sums |p_curr - p_next|, differences of nearly equal numbers: its
    # error is absolute, bounded by the ranks' own error summed over n.
    assert abs(res.l1 - l1) <= FP_TOL * float(np.sum(ref))


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12d"])
@pytest.mark.parametrize("variant", VARIANTS + ["blocked", "pull"])
def test_bfs(rg, orc, ctx, graphs, gname, variant):
    hg = graphs(gname)
    off, nbr = hg.csr()
    srcs = hg.pick_sources(4)
    print(f"bfs {gname} {variant} {srcs}", flush=True)
    dg = rg.DeviceGraph(ctx, hg)
    for s in srcs:
        r = dg.bfs(int(s), rg.opts(variant, hubs=128))
        lv, ecc = orc.bfs_levels(off, nbr, int(s))
        assert orc.check_parents(off, nbr, int(s), lv, r.parents) == 0
        assert r.reached == int((lv >= 0).sum())
        assert r.iters == ecc + 1
        assert r.edges == int(np.diff(off)[lv >= 0].sum())


_RMAT20 = {}


@pytest.mark.parametrize("variant,threshold", [("pull", None), ("pull", 1 << 20), ("pull", 1 << 30), ("push", None),
                                               ("radar", None), ("blocked", None)])
def test_bfs_rmat20_levels_and_query_summaries(rg, orc, ctx, variant, threshold):
    # A million vertices: the bottom-up kernels run a thousand workgroups and
    # their next-frontier lists carry millions of vertices per level.  Every
    # parent is an in-neighbour one level up (levels exact) and every query
    # summary (iterations, reached, edges, level checksum) equals the
    # oracle's.
    if "g" not in _RMAT20:
        _RMAT20["g"] = rg.HostGraph.generate(20, 16, symmetric=True)
    hg = _RMAT20["g"]
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    o = rg.opts(variant, heavy_degree=1024) if threshold is None else \
        rg.opts(variant, threshold=threshold, heavy_degree=1024)
    for s in hg.pick_sources(4):
        print("bfs rmat20", variant, threshold, s, flush=True)
        lv, ecc = orc.bfs_levels(off, nbr, int(s))
        r = dg.bfs(int(s), o)
        assert orc.check_parents(off, nbr, int(s), lv, r.parents) == 0, s
        for _ in range(2):
            assert dg.bfs_query(int(s), o)[0].astuple() == orc.bfs_summary(off, lv), s


@pytest.mark.parametrize("threshold", [0, -1])
def test_pull_bfs_follows_transposes(rg, orc, ctx, graphs, threshold):
    # The pull engine's first-probe array belongs to one orientation: after
    # each transpose of a directed graph it is rebuilt from the new in-edges
    # (threshold 0: every level bottom-up).
    hg = graphs("rmat12d")
    csr = {False: hg.csr(), True: hg.csr(in_edges=True)}
    srcs = hg.pick_sources(3)
    dg = rg.DeviceGraph(ctx, hg)
    o = rg.opts("pull", threshold=threshold)
    for transposed in (False, True, False, True):
        off, nbr = csr[transposed]
        for s in srcs:
            r = dg.bfs(int(s), o)
            lv, ecc = orc.bfs_levels(off, nbr, int(s))
            assert orc.check_parents(off, nbr, int(s), lv, r.parents) == 0, (transposed, s)
            assert r.iters == ecc + 1 and r.reached == int((lv >= 0).sum())
        dg.transpose()


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12d", "rmat16s"])
@pytest.mark.parametrize("variant", VARIANTS + ["blocked"])
def test_radii(rg, orc, ctx, graphs, gname, variant):
    hg = graphs(gname)
    off, nbr = hg.csr()
    starts = rg.radii_starts(hg.n)
    print(f"radii {gname} {variant}", flush=True)
    dg = rg.DeviceGraph(ctx, hg)
    rad, rounds, _ = dg.radii(starts, rg.opts(variant, hubs=128))
    ref, rr = orc.radii(off, nbr, starts)
    assert rounds == rr
    assert np.array_equal(rad, ref)


@pytest.mark.parametrize("gname", ["rmat10s", "rmat12d"])
@pytest.mark.parametrize("variant", VARIANTS + ["blocked"])
def test_bc(rg, orc, ctx, graphs, gname, variant):
    hg = graphs(gname)
    off, nbr = hg.csr()
    ioff, inbr = hg.csr(in_edges=True)
    srcs = hg.pick_bc_sources(4)
    print(f"bc {gname} {variant} {srcs}", flush=True)
    dg = rg.DeviceGraph(ctx, hg)
    deps, comb, _ = dg.bc(srcs, rg.opts(variant, hubs=128))
    ref, rc = orc.bc(off, nbr, ioff, inbr, srcs)
    assert comb == rc
    assert rel_linf(deps, ref) < FP_TOL


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12d", "rmat16s"])
@pytest.mark.parametrize("threshold", [0, -1])
def test_bc_blocked_dense_levels(rg, orc, ctx, graphs, gname, threshold):
    # threshold 0: every forward and backward level on the blocked kernels
    # (directed graphs use the layouts of both orientations).
    hg = graphs(gname)
    off, nbr = hg.csr()
    ioff, inbr = hg.csr(in_edges=True)
    srcs = hg.pick_bc_sources(3)
    print(f"bc blocked {gname} {threshold} {srcs}", flush=True)
    dg = rg.DeviceGraph(ctx, hg)
    deps, comb, _ = dg.bc(srcs, rg.opts("blocked", threshold=threshold))
    ref, rc = orc.bc(off, nbr, ioff, inbr, srcs)
    assert comb == rc
    assert rel_linf(deps, ref) < FP_TOL


@pytest.mark.parametrize("gname", ["rmat10s", "rmat10d", "rmat12s", "rmat12d"])
def test_bc_dense_levels_repeat_exactly_enough(rg, orc, ctx, graphs, gname):
    # Every level dense (threshold 0), four sources with the carry-over, run
    # three times: each run within 1e-10 of the oracle and of the first run
    # (the only run-to-run difference the LDS atomic order).  A 2^10-vertex
    # graph is one workgroup per level; 2^12 several.
    hg = rg.HostGraph.generate(10, 16, symmetric=gname.endswith("s")) if gname.startswith("rmat10") else graphs(gname)
    off, nbr = hg.csr()
    ioff, inbr = hg.csr(in_edges=True)
    srcs = hg.pick_bc_sources(4)
    ref, rc = orc.bc(off, nbr, ioff, inbr, srcs)
    dg = rg.DeviceGraph(ctx, hg)
    first = None
    for rep in range(3):
        print(f"bc dense repeat {gname} {rep}", flush=True)
        for variant in ("blocked", "radar"):
            deps, comb, _ = dg.bc(srcs, rg.opts(variant, threshold=0))
            if not hg.symmetric and len(srcs) % 2:
                dg.transpose()   # an odd number of sources leaves a directed graph transposed
            assert comb == rc
            assert rel_linf(deps, ref) < FP_TOL, (rep, variant)
            if first is None:
                first = deps
            assert rel_linf(deps, first) < FP_TOL


def test_bc_blocked_split_hub_block(rg, orc, ctx):
    # A hub with 2^20 in-edges: its destination block is split over units
    # whose partial sums the merge kernel adds.
    leaves = 1 << 20
    src = np.concatenate([np.zeros(leaves, np.uint32), np.arange(1, 1001, dtype=np.uint32)])
    dst = np.concatenate([np.arange(1, leaves + 1, dtype=np.uint32), np.arange(1001, 2001, dtype=np.uint32)])
    hg = rg.HostGraph.from_edges(leaves + 1, src, dst, symmetric=True)
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    for s in ([5], [0], [1500, 3]):
        deps, comb, _ = dg.bc(np.array(s, np.uint32), rg.opts("blocked", threshold=0))
        ref, rc = orc.bc(off, nbr, off, nbr, np.array(s, np.uint32))
        assert comb == rc
        assert rel_linf(deps, ref) < FP_TOL


def check_bfs(rg, orc, dg, off, nbr, srcs, o):
    for s in srcs:
        r = dg.bfs(int(s), o)
        lv, ecc = orc.bfs_levels(off, nbr, int(s))
        assert orc.check_parents(off, nbr, int(s), lv, r.parents) == 0, f"source {s}"
        assert (r.iters, r.reached) == (ecc + 1, int((lv >= 0).sum()))
        q, _ = dg.bfs_query(int(s), o)
        assert q.astuple() == orc.bfs_summary(off, lv)


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12d", "rmat16s"])
@pytest.mark.parametrize("threshold", [0, -1])
@pytest.mark.parametrize("variant", ["blocked", "pull"])
def test_bfs_blocked_dense_levels(rg, orc, ctx, graphs, gname, threshold, variant):
    hg = graphs(gname)
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    print(f"bfs {variant} {gname} {threshold}", flush=True)
    check_bfs(rg, orc, dg, off, nbr, hg.pick_sources(4), rg.opts(variant, threshold=threshold, heavy_degree=64))


def test_bfs_pull_takes_first_frontier_in_neighbour(rg, orc, ctx, graphs):
    # With every level dense, the bottom-up scan (decodeInNghBreakEarly)
    # picks the first in-neighbour in the frontier, in in-list order, for
    # light and heavy vertices alike.
    hg = graphs("rmat12d")
    off, nbr = hg.csr()
    ioff, inb = hg.csr(in_edges=True)
    dg = rg.DeviceGraph(ctx, hg)
    for s in hg.pick_sources(3):
        r = dg.bfs(int(s), rg.opts("pull", threshold=0, heavy_degree=32))
        lv, _ = orc.bfs_levels(off, nbr, int(s))
        for v in np.nonzero(lv > 0)[0]:
            ins = inb[ioff[v]:ioff[v + 1]]
            first = ins[lv[ins] == lv[v] - 1][0]
            assert r.parents[v] == first, f"vertex {v}"


def test_bfs_blocked_split_hub_block(rg, orc, ctx):
    leaves = 1 << 20
    src = np.concatenate([np.zeros(leaves, np.uint32), np.arange(1, 1001, dtype=np.uint32)])
    dst = np.concatenate([np.arange(1, leaves + 1, dtype=np.uint32), np.arange(1001, 2001, dtype=np.uint32)])
    hg = rg.HostGraph.from_edges(leaves + 1, src, dst, symmetric=True)
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    check_bfs(rg, orc, dg, off, nbr, [5, 0, 1500], rg.opts("blocked", threshold=0))
    check_bfs(rg, orc, dg, off, nbr, [5, 0, 1500], rg.opts("pull", threshold=0))


def test_bfs_radar_on_degree_sorted(rg, orc, ctx, graphs):
    hg = graphs("rmat16s")
    dg = rg.DeviceGraph(ctx, hg)
    sg, ids = dg.degree_sort()
    off, nbr = sg.download()
    check_bfs(rg, orc, sg, off, nbr, ids[hg.pick_sources(4)], rg.opts("radar"))


@pytest.mark.parametrize("threshold", [0, -1, 1 << 60])
def test_radii_batch_edge_cases(rg, orc, ctx, graphs, threshold):
    # Repeated starts, fewer than 64 starts and no starts, on the blocked
    # (bit-parallel batch) engine, all levels dense / default / all sparse.
    hg = graphs("rmat12d")
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    o = rg.opts("blocked", threshold=threshold)
    base = rg.radii_starts(hg.n)
    for starts in (np.concatenate([base[:10], base[:3]]), base[:1], base):
        rad, rounds, _ = dg.radii(starts, o)
        ref, rr = orc.radii(off, nbr, starts)
        assert rounds == rr and np.array_equal(rad, ref)
    rad, rounds, _ = dg.radii(np.zeros(0, np.uint32), o)
    assert rounds == 0 and (rad == -1).all()


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12s", "rmat16s"])
@pytest.mark.parametrize("variant", VARIANTS + ["blocked"])
def test_pagerank_delta(rg, orc, ctx, graphs, gname, variant):
    hg = graphs(gname)
    off, nbr = hg.csr()
    print(f"pagerank_delta {gname} {variant}", flush=True)
    dg = rg.DeviceGraph(ctx, hg)
    for iters in (20, 3):
        res = dg.pagerank_delta(iters, rg.opts(variant, hubs=256))
        ref, it, l1 = orc.pagerank_delta(off, nbr, iters)
        assert res.iters == it
        assert rel_linf(res.ranks, ref) < FP_TOL
        # Round 1's Delta is 0.85 nghSum + 0.15/n - 1/n, a difference of
        # nearly equal numbers: the L1 error is absolute, bounded by the
        # ranks' own error summed over n (as for PageRank's L1).
        assert abs(res.l1 - l1) <= FP_TOL * float(np.sum(ref))


@pytest.mark.parametrize("gname", ["rmat14s", "rmat12d"])
@pytest.mark.parametrize("wg", [512, 1024])
@pytest.mark.parametrize("unroll", [1, 4])
def test_blocked_shapes(rg, orc, ctx, graphs, gname, wg, unroll):
    # Every workgroup shape / arcs-in-flight of the blocked engine (the
    # autotuner's candidates; the default is 1024 threads, 4 arcs), for
    # PageRank and for PageRankDelta's dense rounds.
    hg = graphs(gname)
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    o = rg.opts("blocked", wg_size=wg, unroll=unroll)
    res = dg.pagerank(20, o)
    ref, it, _ = orc.pagerank(off, nbr, 20)
    assert res.iters == it and rel_linf(res.ranks, ref) < FP_TOL
    res = dg.pagerank_delta(20, o)
    ref, it, _ = orc.pagerank_delta(off, nbr, 20)
    assert res.iters == it and rel_linf(res.ranks, ref) < FP_TOL


def _split_star(rg):
    # A hub with 2^20 in-edges (its block split over units) plus a path.
    leaves = 1 << 20
    src = np.concatenate([np.zeros(leaves, np.uint32), np.arange(1, 1001, dtype=np.uint32)])
    dst = np.concatenate([np.arange(1, leaves + 1, dtype=np.uint32), np.arange(1001, 2001, dtype=np.uint32)])
    return rg.HostGraph.from_edges(leaves + 1, src, dst, symmetric=True)


@pytest.mark.parametrize("gname", ["star", "rmat16s"])
@pytest.mark.parametrize("unit_order", [0, 1, 2])
@pytest.mark.parametrize("unit_queue", [0, 1])
def test_blocked_unit_schedules(rg, orc, ctx, graphs, gname, unit_order, unit_queue):
    # Every unit schedule of the blocked engine -- one work queue by first
    # source (default), per-XCD queues, the layout's order; persistent
    # workgroups or one per unit -- computes the oracle's PageRank and
    # PageRankDelta, on graphs whose hub blocks are split over several units
    # (the star's hub; the degree-sorted R-MAT's first 1024-vertex blocks).
    hg = _split_star(rg) if gname == "star" else graphs(gname)
    print(f"blocked schedule {gname} order={unit_order} queue={unit_queue}", flush=True)
    dg = rg.DeviceGraph(ctx, hg)
    if gname != "star":
        dg, _ = dg.degree_sort(sort_neighbors=True)
    off, nbr = dg.download()
    o = rg.opts("blocked", dst_block_bits=10, unit_order=unit_order, unit_queue=unit_queue)
    for _ in range(2):   # the second call replays the captured loop
        res = dg.pagerank(20, o)
        ref, it, _ = orc.pagerank(off, nbr, 20)
        assert res.iters == it and rel_linf(res.ranks, ref) < FP_TOL
    res = dg.pagerank_delta(20, o)
    ref, it, _ = orc.pagerank_delta(off, nbr, 20)
    assert res.iters == it and rel_linf(res.ranks, ref) < FP_TOL


@pytest.mark.parametrize("gname", ["rmat14s", "rmat12d"])
def test_autotune_choice_gives_oracle_answers(rg, orc, ctx, graphs, gname):
    # Whatever configuration the autotuner picks (engine, workgroup shape,
    # hub count, threshold), it computes the same PageRank and BFS levels.
    hg = graphs(gname)
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    sg, ids = dg.degree_sort()
    best, ms = sg.autotune("pagerank", base=rg.opts("radar"))
    assert ms > 0 and best.variant in (rg.BLOCKED, rg.RADAR)
    soff, snbr = sg.download()
    r = sg.pagerank(20, best)
    ref, it, _ = orc.pagerank(soff, snbr, 20)
    assert r.iters == it and rel_linf(r.ranks, ref) < FP_TOL
    srcs = ids[hg.pick_sources(4)]
    bt, _ = sg.autotune("bfs", sources=srcs, base=rg.opts("push"))
    assert bt.variant in (rg.PUSH, rg.RADAR, rg.PULL)
    if bt.variant == rg.RADAR:
        assert bt.hubs in (1024, 4096)
    for s in srcs:
        b = sg.bfs(int(s), bt)
        lv, ecc = orc.bfs_levels(soff, snbr, int(s))
        assert orc.check_parents(soff, snbr, int(s), lv, b.parents) == 0
        assert b.iters == ecc + 1


@pytest.mark.parametrize("n", [5, 1000, 1025, 3000])
def test_blocked_pagerank_small_graphs(rg, orc, ctx, n):
    # The packed arcs of the blocked PageRank layout on small vertex counts:
    # below 1024 vertices every source is its own segment (no low bits),
    # just above, one low bit.  A directed cycle (every rank 1/n, a closed
    # form) plus random chords, against the oracle.
    rng = np.random.default_rng(n)
    k = 4 * n
    src = np.concatenate([np.arange(n), rng.integers(0, n, k)]).astype(np.uint32)
    dst = np.concatenate([(np.arange(n) + 1) % n, rng.integers(0, n, k)]).astype(np.uint32)
    for sym in (False, True):
        hg = rg.HostGraph.from_edges(n, src, dst, symmetric=sym)
        off, nbr = hg.csr()
        dg = rg.DeviceGraph(ctx, hg)
        for vb in (0, 10):
            res = dg.pagerank(20, rg.opts("blocked", dst_block_bits=vb))
            ref, it, _ = orc.pagerank(off, nbr, 20)
            assert res.iters == it and rel_linf(res.ranks, ref) < FP_TOL, (sym, vb)
    cyc = rg.HostGraph.from_edges(n, np.arange(n, dtype=np.uint32), ((np.arange(n) + 1) % n).astype(np.uint32),
                                  symmetric=False)
    r = rg.DeviceGraph(ctx, cyc).pagerank(20, rg.opts("blocked"))
    assert np.allclose(r.ranks, 1.0 / n, rtol=1e-12, atol=0)
