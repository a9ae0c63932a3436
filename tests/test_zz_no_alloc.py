"""GPU test of rg_bfs_noalloc, the single-source BFS that allocates
nothing (its scratch is one fixed device array per GPU): its answers
through the C ABI equal rg_bfs's with the push engine -- the same rounds,
reached count and traversed-edge count, the same level array (checksum, and
levels recomputed from both parent arrays), parents a valid shortest-path
tree (checked by the oracle) -- on two calls in a row, the second on a
smaller graph, and graphs beyond the fixed size are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UINT_MAX = 0xFFFFFFFF


def bfs_levels_from_parents(parents, src):
    n = len(parents)
    lv = np.full(n, -1, np.int64)
    lv[src] = 0
    for v in range(n):
        if parents[v] == UINT_MAX or lv[v] >= 0:
            continue
        path, u = [], v
        while lv[u] < 0:
            path.append(u)
            u = int(parents[u])
        for w in reversed(path):
            lv[w] = lv[u] + 1
            u = w
    return lv


def check_equal_to_rg_bfs(rg, orc, ctx, hg, src):
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    try:
        print(f"noalloc n={dg.n} m={dg.m} src={src}", flush=True)
        par = rg.DeviceArray(ctx, dg.n, np.uint32)
        par.fill(7)   # stale contents: the call overwrites every entry
        s = dg.bfs_noalloc(src, par)
        p_na = par.read()
        ref = dg.bfs(src, rg.opts("push"))
        q, _ = dg.bfs_query(src, rg.opts("push"))
        assert (s.iters, s.reached, s.edges) == (ref.iters, ref.reached, ref.edges)
        assert s.astuple() == q.astuple()
        lv, ecc = orc.bfs_levels(off, nbr, src)
        assert orc.check_parents(off, nbr, src, lv, p_na) == 0
        assert np.array_equal(bfs_levels_from_parents(p_na, src), bfs_levels_from_parents(ref.parents, src))
        assert s.astuple() == orc.bfs_summary(off, lv)
        par.close()
    finally:
        dg.close()


def test_noalloc_bfs_matches_rg_bfs_twice_smaller_second(rg, orc, ctx, graphs):
    # Two calls in a row on one GPU, the second on a smaller graph (the
    # fixed array keeps the first call's contents).
    big, small = graphs("rmat14s"), graphs("rmat12d")
    check_equal_to_rg_bfs(rg, orc, ctx, big, int(big.pick_sources(1)[0]))
    check_equal_to_rg_bfs(rg, orc, ctx, small, int(small.pick_sources(1)[0]))


def test_noalloc_bfs_dense_rounds_and_repeats(rg, orc, ctx, graphs):
    # A symmetric R-MAT crosses the m/20 switch (dense rounds over the byte
    # map); repeated sources on one graph give the same summaries.
    hg = graphs("rmat16s")
    for src in hg.pick_sources(3):
        check_equal_to_rg_bfs(rg, orc, ctx, hg, int(src))


def test_noalloc_bfs_refuses_graphs_beyond_the_fixed_array(rg, ctx):
    # 2^23 vertices need 9 * 2^23 bytes > 64 MiB: refused before any launch.
    n = 1 << 23
    a = np.arange(16, dtype=np.uint32)
    hg = rg.HostGraph.from_edges(n, a, a + 1, symmetric=True)
    dg = rg.DeviceGraph(ctx, hg)
    try:
        par = rg.DeviceArray(ctx, n, np.uint32)
        with pytest.raises(rg.RGError, match="bytes of scratch"):
            dg.bfs_noalloc(0, par)
        with pytest.raises(rg.RGError):
            dg.bfs_noalloc(n, par)   # source out of range
        par.close()
    finally:
        dg.close()
