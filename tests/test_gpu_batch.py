"""GPU tests of the batched (bit-parallel multi-source) BFS: every source's
level array equals the oracle's BFS levels bit-for-bit, and every batched
summary equals the single-source engine's answer for the same query
(rg_bfs_query) and the summary computed from the oracle levels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BIG = 1 << 60


def pick(n, k, seed):
    r = np.random.default_rng(seed)
    return r.integers(0, n, size=k).astype(np.uint32)


def check_batch(rg, orc, hg, dg, srcs, o=None):
    print("batch", dg.n, dg.m, len(srcs), flush=True)
    off, nb = hg.csr()
    out, lv, _ = dg.bfs_batch(srcs, o, want_levels=True)
    for i, s in enumerate(srcs):
        ref, _ = orc.bfs_levels(off, nb, int(s))
        assert np.array_equal(lv[i], ref), f"source {s} (slot {i})"
        assert out[i].astuple() == orc.bfs_summary(off, ref), f"source {s}"
    return out


@pytest.mark.parametrize("name", ["rmat12s", "rmat12d", "er12s", "er12d"])
@pytest.mark.parametrize("k", [1, 7, 64, 65, 130, 256])
def test_batch_levels_match_oracle(rg, orc, ctx, graphs, name, k):
    hg = graphs(name)
    dg = rg.DeviceGraph(ctx, hg)
    check_batch(rg, orc, hg, dg, pick(hg.n, k, 100 + k))


@pytest.mark.parametrize("threshold", [0, BIG])
@pytest.mark.parametrize("k", [64, 200])
def test_batch_all_dense_or_all_sparse(rg, orc, ctx, graphs, threshold, k):
    # threshold 0 forces every level onto the blocked dense kernel, a huge
    # one every level onto the push engine.
    hg = graphs("rmat12d")
    dg = rg.DeviceGraph(ctx, hg)
    check_batch(rg, orc, hg, dg, pick(hg.n, k, 7), rg.opts(threshold=threshold))


def test_batch_duplicates_and_isolated(rg, orc, ctx, graphs):
    # Repeated sources share vertices in the first frontier; isolated
    # sources finish at level 0.
    hg = graphs("rmat12s")
    dg = rg.DeviceGraph(ctx, hg)
    deg = np.diff(hg.offsets())
    iso = np.nonzero(deg == 0)[0][:3].astype(np.uint32)
    srcs = np.concatenate([np.full(5, 17, np.uint32), iso, pick(hg.n, 40, 3), np.full(3, 17, np.uint32)])
    out = check_batch(rg, orc, hg, dg, srcs)
    for i in range(len(iso)):
        assert out[5 + i].iters == 1 and out[5 + i].reached == 1


def test_batch_equals_single_queries(rg, orc, ctx, graphs):
    hg = graphs("rmat14d")
    dg = rg.DeviceGraph(ctx, hg)
    srcs = hg.pick_sources(96)
    out, _, _ = dg.bfs_batch(srcs)
    for i, s in enumerate(srcs):
        single, _ = dg.bfs_query(int(s))
        assert out[i].astuple() == single.astuple(), f"source {s}"


def test_batch_split_hub_block(rg, orc, ctx):
    # A star whose hub has 2^20 in-edges: its destination block is split
    # over several workgroups whose partial words are merged.
    leaves = 1 << 20
    src = np.zeros(leaves, np.uint32)
    dst = np.arange(1, leaves + 1, dtype=np.uint32)
    # a few extra edges between leaves so levels go past 2
    src = np.concatenate([src, dst[:1000]])
    dst = np.concatenate([dst, dst[1000:2000]])
    hg = rg.HostGraph.from_edges(leaves + 1, src, dst, symmetric=True)
    dg = rg.DeviceGraph(ctx, hg)
    srcs = np.array([0, 1, 5, 1500, 999, 77777, leaves], np.uint32)
    check_batch(rg, orc, hg, dg, srcs, rg.opts(threshold=0))
    check_batch(rg, orc, hg, dg, srcs)


def test_batch_degree_sorted(rg, orc, ctx, graphs):
    # RADAR layout: hubs are ids 0.., the first destination blocks are split.
    hg = graphs("rmat16s")
    dg = rg.DeviceGraph(ctx, hg)
    sg, ids = dg.degree_sort()
    srcs = hg.pick_sources(128)
    out_s, _, _ = sg.bfs_batch(ids[srcs])
    out_u, _, _ = dg.bfs_batch(srcs)
    off, nb = hg.csr()
    for i, s in enumerate(srcs[:16]):
        ref, _ = orc.bfs_levels(off, nb, int(s))
        it, reached, edges, _ = orc.bfs_summary(off, ref)
        assert out_u[i].astuple() == orc.bfs_summary(off, ref)
        # relabelling permutes vertex ids, so the checksum differs; the
        # rest of the summary is invariant
        assert (out_s[i].iters, out_s[i].reached, out_s[i].edges) == (it, reached, edges)


def test_batch_repeat_reuses_state(rg, orc, ctx, graphs):
    # Consecutive batches of different widths on one context: the state
    # left by a batch is reused, so answers must not depend on the order.
    hg = graphs("rmat12d")
    dg = rg.DeviceGraph(ctx, hg)
    a = pick(hg.n, 150, 11)
    b = pick(hg.n, 20, 12)
    first, _, _ = dg.bfs_batch(a)
    dg.bfs_batch(b)
    dg.bfs_batch(pick(hg.n, 256, 13))
    again, _, _ = dg.bfs_batch(a)
    assert [x.astuple() for x in first] == [x.astuple() for x in again]


def test_batch_rejects_bad_input(rg, ctx, graphs):
    hg = graphs("rmat10s")
    dg = rg.DeviceGraph(ctx, hg)
    with pytest.raises(rg.RGError):
        dg.bfs_batch(np.zeros(257, np.uint32))
    with pytest.raises(rg.RGError):
        dg.bfs_batch(np.array([hg.n], np.uint32))


# One persistent launch for the whole batch (rg_opts.batch_launch = 2): the
# same answers as the oracle and as the per-level kernels (batch_launch = 1).

PERSIST, LEVELS = 2, 1


@pytest.mark.parametrize("name", ["rmat12s", "rmat12d", "er12s"])
@pytest.mark.parametrize("k", [1, 7, 64, 65, 130, 256])
def test_persistent_batch_levels_match_oracle(rg, orc, ctx, graphs, name, k):
    hg = graphs(name)
    dg = rg.DeviceGraph(ctx, hg)
    check_batch(rg, orc, hg, dg, pick(hg.n, k, 300 + k), rg.opts(batch_launch=PERSIST))


@pytest.mark.parametrize("threshold", [0, BIG])
@pytest.mark.parametrize("k", [64, 200])
def test_persistent_batch_all_dense_or_all_sparse(rg, orc, ctx, graphs, threshold, k):
    hg = graphs("rmat12d")
    dg = rg.DeviceGraph(ctx, hg)
    check_batch(rg, orc, hg, dg, pick(hg.n, k, 9), rg.opts(threshold=threshold, batch_launch=PERSIST))


def test_persistent_batch_split_hub_block(rg, orc, ctx):
    # The star of test_batch_split_hub_block: split blocks merged between
    # two barriers of the one launch; its hub's out-list is a heavy item on
    # sparse levels.
    leaves = 1 << 20
    src = np.zeros(leaves, np.uint32)
    dst = np.arange(1, leaves + 1, dtype=np.uint32)
    src = np.concatenate([src, dst[:1000]])
    dst = np.concatenate([dst, dst[1000:2000]])
    hg = rg.HostGraph.from_edges(leaves + 1, src, dst, symmetric=True)
    dg = rg.DeviceGraph(ctx, hg)
    srcs = np.array([0, 1, 5, 1500, 999, 77777, leaves], np.uint32)
    for threshold in (0, -1, BIG):
        check_batch(rg, orc, hg, dg, srcs, rg.opts(threshold=threshold, batch_launch=PERSIST))


@pytest.mark.parametrize("launch", [LEVELS, PERSIST])
def test_batch_hub_block_over_many_merge_chunks(rg, orc, ctx, launch):
    # A star whose hub has 2^24 in-edges: its destination block is split
    # over 33 units of 2^19 arcs, merged in chunks of 16 units (two full
    # chunks and one of a single unit) before its vertices are finalised.
    leaves = 1 << 24
    src = np.zeros(leaves, np.uint32)
    dst = np.arange(1, leaves + 1, dtype=np.uint32)
    src = np.concatenate([src, dst[:1000]])
    dst = np.concatenate([dst, dst[1000:2000]])
    hg = rg.HostGraph.from_edges(leaves + 1, src, dst, symmetric=True)
    dg = rg.DeviceGraph(ctx, hg)
    srcs = np.array([0, 1, 5, 1500, 999, 77777, leaves], np.uint32)
    for threshold in (0, None):
        o = rg.opts(batch_launch=launch) if threshold is None else rg.opts(threshold=threshold, batch_launch=launch)
        out = check_batch(rg, orc, hg, dg, srcs, o)
        alone, _, _ = dg.bfs_batch(srcs, o)   # summaries without level arrays
        assert [x.astuple() for x in alone] == [x.astuple() for x in out]


def test_persistent_equals_per_level_launches(rg, orc, ctx, graphs):
    # Summaries and level arrays of both launch forms, alternated on one
    # context (each form leaves the state the other starts from).
    hg = graphs("rmat14d")
    dg = rg.DeviceGraph(ctx, hg)
    for k, seed in ((96, 21), (256, 22), (3, 23), (130, 24)):
        srcs = pick(hg.n, k, seed)
        a, la, _ = dg.bfs_batch(srcs, rg.opts(batch_launch=LEVELS), want_levels=True)
        b, lb, _ = dg.bfs_batch(srcs, rg.opts(batch_launch=PERSIST), want_levels=True)
        assert [x.astuple() for x in a] == [x.astuple() for x in b], k
        assert np.array_equal(la, lb), k


def test_persistent_batch_degree_sorted(rg, orc, ctx, graphs):
    hg = graphs("rmat16s")
    dg = rg.DeviceGraph(ctx, hg)
    sg, ids = dg.degree_sort(sort_neighbors=True)
    srcs = ids[hg.pick_sources(200)]
    a, _, _ = sg.bfs_batch(srcs, rg.opts(batch_launch=LEVELS))
    b, _, _ = sg.bfs_batch(srcs, rg.opts(batch_launch=PERSIST))
    assert [x.astuple() for x in a] == [x.astuple() for x in b]
    for i in (0, 77, 199):
        single, _ = sg.bfs_query(int(srcs[i]))
        assert b[i].astuple() == single.astuple()


@pytest.mark.parametrize("threshold", [0, -1, BIG])
def test_persistent_radii(rg, orc, ctx, graphs, threshold):
    # Radii as one bit-parallel batch in one launch: exact radii and rounds.
    for name in ("rmat12d", "rmat12s"):
        hg = graphs(name)
        off, nbr = hg.csr()
        dg = rg.DeviceGraph(ctx, hg)
        starts = rg.radii_starts(hg.n)
        rad, rounds, _ = dg.radii(starts, rg.opts("blocked", threshold=threshold, batch_launch=PERSIST))
        ref, rr = orc.radii(off, nbr, starts)
        assert rounds == rr and np.array_equal(rad, ref), name


def test_persistent_batches_from_two_contexts(rg, orc, ctx, graphs):
    # Two host threads, each with its own context, run persistent batches of
    # different widths at the same time: the launches take turns, and every
    # answer equals the per-level form's.
    import threading
    hg = graphs("rmat14d")
    ctxs = [rg.Context(0), rg.Context(0)]
    try:
        dgs = [rg.DeviceGraph(c, hg) for c in ctxs]
        srcs = [pick(hg.n, 200, 51), pick(hg.n, 40, 52)]
        want = [[x.astuple() for x in dgs[i].bfs_batch(srcs[i], rg.opts(batch_launch=LEVELS))[0]] for i in range(2)]
        got = [[], []]
        errs = []

        def run(i):
            try:
                for _ in range(4):
                    out, _, _ = dgs[i].bfs_batch(srcs[i], rg.opts(batch_launch=PERSIST))
                    got[i].append([x.astuple() for x in out])
            except Exception as e:   # reported below
                errs.append(e)

        ts = [threading.Thread(target=run, args=(i,)) for i in range(2)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert not errs, errs
        for i in range(2):
            assert all(g == want[i] for g in got[i]) and len(got[i]) == 4
    finally:
        for c in ctxs:
            c.close()


# Summaries alone (no level arrays): the per-source statistics share one
# 64-bit word for count and edge sum and walk the words with more than half
# of their sources set by their complement; RG_BATCH_WIDE_STATS forces the
# separate counters that graphs with 2^32 arcs or more take.  Every summary
# equals the oracle's, for widths of one, two and four words.
@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("launch", [LEVELS, PERSIST])
def test_batch_summaries_without_levels(rg, orc, ctx, graphs, monkeypatch, wide, launch):
    if wide:
        monkeypatch.setenv("RG_BATCH_WIDE_STATS", "1")
    for name, k, threshold in (("rmat14s", 256, None), ("rmat14d", 100, 0), ("rmat12s", 40, None),
                               ("er12s", 200, BIG)):
        hg = graphs(name)
        dg = rg.DeviceGraph(ctx, hg)
        off, nb = hg.csr()
        srcs = hg.pick_sources(k)
        o = rg.opts(batch_launch=launch) if threshold is None else rg.opts(threshold=threshold, batch_launch=launch)
        print("summaries", name, k, threshold, launch, wide, flush=True)
        out, _, _ = dg.bfs_batch(srcs, o)
        for i, s in enumerate(srcs):
            ref, _ = orc.bfs_levels(off, nb, int(s))
            assert out[i].astuple() == orc.bfs_summary(off, ref), f"{name} source {s} (slot {i})"
