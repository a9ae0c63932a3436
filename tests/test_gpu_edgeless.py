"""GPU tests on graphs without edges (and a single vertex): every app gives
the oracle's answer -- PageRank converges to 0.15/n, every source reaches
only itself, BC dependencies are NaN except at the source, no triangles."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10


@pytest.mark.parametrize("n,symmetric", [(1, True), (5, True), (5, False)])
def test_edgeless_graph_every_app(rg, orc, ctx, n, symmetric):
    e = np.array([], np.uint32)
    hg = rg.HostGraph.from_edges(n, e, e, symmetric=symmetric)
    off, nbr = hg.csr()
    ioff, inbr = hg.csr(in_edges=True)
    dg = rg.DeviceGraph(ctx, hg)
    ref, it, _ = orc.pagerank(off, nbr, 20)
    for v in ("push", "hubdup", "radar", "pull", "blocked"):
        print("pagerank", n, symmetric, v, flush=True)
        r = dg.pagerank(20, rg.opts(v))
        assert r.iters == it and np.max(np.abs(r.ranks - ref)) <= FP_TOL * np.max(np.abs(ref)), v
    for v in ("push", "hubdup", "radar", "pull", "blocked"):
        print("bfs", n, symmetric, v, flush=True)
        b = dg.bfs(n - 1, rg.opts(v))
        assert (b.iters, b.reached, b.edges) == (1, 1, 0), v
        assert b.parents[n - 1] == n - 1 and np.all(np.delete(b.parents, n - 1) == 0xFFFFFFFF)
    srcs = np.arange(n, dtype=np.uint32)
    for launch in (1, 2):   # per-level kernels, one persistent launch
        print("bfs batch", launch, flush=True)
        out, lv, _ = dg.bfs_batch(srcs, rg.opts(batch_launch=launch), want_levels=True)
        for i, s in enumerate(srcs):
            ref_lv, _ = orc.bfs_levels(off, nbr, int(s))
            assert np.array_equal(lv[i], ref_lv) and out[i].astuple() == orc.bfs_summary(off, ref_lv)
    print("radii", flush=True)
    rad, rounds, _ = dg.radii(srcs, rg.opts("radar"))
    ref_rad, ref_rounds = orc.radii(off, nbr, srcs)
    assert np.array_equal(rad, ref_rad) and rounds == ref_rounds
    print("bc", flush=True)
    for v in ("push", "radar"):
        deps, comb, _ = dg.bc([0], rg.opts(v))
        if not symmetric:
            dg.transpose()
        ref_d, ref_c = orc.bc(off, nbr, ioff, inbr, [0])
        assert comb == ref_c and np.array_equal(np.isnan(deps), np.isnan(ref_d))
    for launch in (0, 1, 2):   # library's choice, per-level kernels, one persistent launch
        print("bc batch", launch, flush=True)
        bs, deps, _ = dg.bc_batch(srcs, rg.opts(batch_launch=launch), want_deps=True)
        for i, s in enumerate(srcs):
            ref_d, ref_c = orc.bc(off, nbr, ioff, inbr, [int(s)])
            assert bs[i].rounds == ref_c and bs[i].reached == 1
            assert np.array_equal(np.isnan(deps[i]), np.isnan(ref_d))
    if symmetric:
        print("tc", flush=True)
        assert dg.tc()[0] == 0
