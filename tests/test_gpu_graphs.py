"""GPU tests of the HIP-graph replay of iteration loops: the blocked
PageRank loop captured once per context and replayed.  Every replay, with
other eager work and synchronisations between the replays, agrees with the
eager launches (the only difference being the LDS atomic order: ranks
within 1e-10 relative L-inf, L1 within 1e-10 of the rank sum, iterations
exact) and with the oracle; a changed graph or
iteration count is captured again."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_pagerank_graph_replays_match_eager(rg, orc, ctx, graphs):
    hg = graphs("rmat14s")
    off, nbr = hg.csr()
    c = rg.Context(0)
    try:
        dg = rg.DeviceGraph(c, hg)
        o = rg.opts("blocked")
        c.graphs(False)
        eager = dg.pagerank(20, o)
        ref, it, _ = orc.pagerank(off, nbr, 20)
        assert eager.iters == it and rel(eager.ranks, ref) < FP_TOL
        c.graphs(True)
        src = hg.pick_sources(3)
        for rep in range(5):
            print("replay", rep, flush=True)
            r = dg.pagerank(20, o)
            assert r.iters == eager.iters, rep
            assert rel(r.ranks, eager.ranks) < FP_TOL, rep
            # L1 sums |p_curr - p_next|, differences of nearly equal ranks
            # whose last bits follow the LDS atomic order, so eager and
            # replay agree absolutely, to the ranks' error summed over n,
            # not relative to L1 itself.
            assert abs(r.l1 - eager.l1) <= FP_TOL * float(np.sum(eager.ranks)), rep
            # eager work on the same context between the replays
            dg.bfs(int(src[rep % 3]))
            dg.bc_batch(src)
            c.sync()
    finally:
        c.close()


def test_pagerank_graph_recaptured_on_change(rg, orc, ctx, graphs):
    # Alternating graphs and iteration counts on one context: each call
    # captures again and still equals the oracle.
    a, b = graphs("rmat12s"), graphs("rmat12d")
    c = rg.Context(0)
    try:
        da, db = rg.DeviceGraph(c, a), rg.DeviceGraph(c, b)
        o = rg.opts("blocked")
        for hg, dg, iters in [(a, da, 20), (b, db, 20), (a, da, 7), (a, da, 7), (b, db, 20), (a, da, 20)]:
            off, nbr = hg.csr()
            ref, it, l1 = orc.pagerank(off, nbr, iters)
            r = dg.pagerank(iters, o)
            assert r.iters == it and rel(r.ranks, ref) < FP_TOL, (hg.n, iters)
    finally:
        c.close()


def test_pagerank_converged_loop_replay(rg, orc, ctx):
    # A graph that converges before maxiters: the replayed loop's later
    # iterations do nothing, and the printed iterations / L1 are the
    # converged ones.
    n = 64
    a = np.arange(n, dtype=np.uint32)
    hg = rg.HostGraph.from_edges(n, a, (a + 1) % n, symmetric=True)   # a cycle: converges at once
    off, nbr = hg.csr()
    c = rg.Context(0)
    try:
        dg = rg.DeviceGraph(c, hg)
        ref, it, l1 = orc.pagerank(off, nbr, 50)
        for _ in range(3):
            r = dg.pagerank(50, rg.opts("blocked"))
            assert r.iters == it and rel(r.ranks, ref) < FP_TOL
    finally:
        c.close()


@pytest.mark.parametrize("gname,k", [("rmat14s", 64), ("rmat12d", 100), ("rmat14s", 256), ("rmat12d", 1)])
def test_batch_level_chunk_replays_match_eager(rg, orc, graphs, gname, k):
    # The bit-parallel batch's chunk of levels is captured once per word
    # width and replayed: batches of other sources, other eager work and
    # synchronisations in between, every replayed batch equal bit-for-bit to
    # the same batch launched kernel by kernel, and to the oracle.
    hg = graphs(gname)
    off, nbr = hg.csr()
    c = rg.Context(0)
    try:
        dg = rg.DeviceGraph(c, hg)
        srcs = hg.pick_sources(k + 8)
        sets = [srcs[:k], srcs[8:8 + k], srcs[:k]]
        c.graphs(False)
        eager = [dg.bfs_batch(s, want_levels=True)[:2] for s in sets]
        c.graphs(True)
        for rep in range(6):
            i = rep % 3
            print("batch replay", gname, k, rep, flush=True)
            out, lv, _ = dg.bfs_batch(sets[i], want_levels=rep % 2 == 0)
            assert [x.astuple() for x in out] == [x.astuple() for x in eager[i][0]], rep
            if lv is not None:
                assert np.array_equal(lv, eager[i][1]), rep
            dg.bfs(int(srcs[rep]))              # eager work on the same context
            dg.radii(rg.radii_starts(dg.n), rg.opts("blocked"))   # a batch of another width
            c.sync()
        for j in (0, len(sets[0]) - 1):
            ref, _ = orc.bfs_levels(off, nbr, int(sets[0][j]))
            assert np.array_equal(eager[0][1][j], ref)
    finally:
        c.close()


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12d", "rmat16s"])
@pytest.mark.parametrize("threshold", [-1, 0, 1 << 40])
def test_bfs_device_levels_match_host_levels(rg, orc, ctx, graphs, gname, threshold):
    # With RG_GRAPHS_BFS_LEVELS the pull engine decides its level kinds on
    # the device and replays a captured chunk of levels (eagerly launched
    # with loops off); by default the host reads every level's counts and
    # decides.  Both give the oracle's levels (a
    # valid parent tree), the same rounds, reached vertices and out-degree
    # sum, and the same query checksum, over repeated replays with eager
    # work between them (threshold 0: every level bottom-up; 2^40: every
    # level top-down; also a source without out-arcs).
    hg = graphs(gname)
    off, nbr = hg.csr()
    srcs = list(hg.pick_sources(3))
    deg = np.diff(off)
    zero = np.nonzero(deg == 0)[0]
    if zero.size:
        srcs.append(int(zero[0]))
    c = rg.Context(0)
    try:
        dg = rg.DeviceGraph(c, hg)
        o = rg.opts("pull", threshold=threshold, heavy_degree=64)
        host = {s: (dg.bfs(int(s), o), dg.bfs_query(int(s), o)[0].astuple()) for s in srcs}
        for rep in range(4):
            c.graphs(rep != 3, bfs_device_levels=True)
            for s in srcs:
                print(f"bfs device levels {gname} thr={threshold} src={s} rep={rep}", flush=True)
                r = dg.bfs(int(s), o)
                lv, ecc = orc.bfs_levels(off, nbr, int(s))
                assert orc.check_parents(off, nbr, int(s), lv, r.parents) == 0
                h, hq = host[s]
                assert (r.iters, r.reached, r.edges) == (h.iters, h.reached, h.edges) == (
                    ecc + 1, int((lv >= 0).sum()), int(deg[lv >= 0].sum()))
                q, _ = dg.bfs_query(int(s), o)
                assert q.astuple() == hq == orc.bfs_summary(off, lv)
            dg.pagerank(3, rg.opts("blocked"))   # eager work between the replays
            c.sync()
    finally:
        c.close()
