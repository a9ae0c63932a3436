"""GPU DegreeSorting and triangle counting against the oracle (bit-exact:
integers only), plus RADAR on a degree-sorted graph against push-only."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12d"])
@pytest.mark.parametrize("by_outdeg", [0, 1])
@pytest.mark.parametrize("strict", [0, 1])
def test_degree_sort(rg, orc, ctx, graphs, gname, by_outdeg, strict):
    hg = graphs(gname)
    print(f"degree_sort {gname} out={by_outdeg} strict={strict}", flush=True)
    dg = rg.DeviceGraph(ctx, hg)
    ng, ids = dg.degree_sort(by_outdeg=bool(by_outdeg), sort_neighbors=bool(strict))
    off, nb = hg.csr()
    key = np.diff(hg.offsets(in_edges=not (by_outdeg or hg.symmetric))).astype(np.uint32)
    ref_ids = orc.degree_order(key)
    assert np.array_equal(ids, ref_ids)
    noff, nnb = orc.relabel_csr(off, nb, ref_ids, strict=bool(strict))
    goff, gnb = ng.download()
    assert np.array_equal(goff, noff) and np.array_equal(gnb, nnb)
    if not hg.symmetric:
        ioff, inb = hg.csr(in_edges=True)
        nioff, ninb = orc.relabel_csr(ioff, inb, ref_ids, strict=bool(strict))
        gio, gin = ng.download(in_edges=True)
        assert np.array_equal(gio, nioff) and np.array_equal(gin, ninb)


@pytest.mark.parametrize("gname", ["rmat10s", "rmat12s", "er12s"])
@pytest.mark.parametrize("relabel", [False, True])
def test_tc(rg, orc, ctx, graphs, gname, relabel):
    hg = graphs(gname)
    print(f"tc {gname} relabel={relabel}", flush=True)
    dg = rg.DeviceGraph(ctx, hg)
    if relabel:
        dg, _ = dg.degree_sort(by_outdeg=True, sort_neighbors=True)
    off, nb = dg.download()
    tot, pv, _ = dg.tc(per_vertex=True)
    rtot, rpv = orc.tc(off, nb, per_vertex=True)
    assert tot == rtot
    assert np.array_equal(pv, rpv)
    assert int(pv.sum()) == 3 * tot


def test_tc_rejects_directed(rg, ctx, graphs):
    dg = rg.DeviceGraph(ctx, graphs("rmat12d"))
    with pytest.raises(rg.RGError):
        dg.tc()


@pytest.mark.parametrize("gname", ["rmat14s", "rmat14d"])
def test_radar_on_sorted_graph_matches_push(rg, orc, ctx, graphs, gname):
    """RADAR (hubs = ids 0..k-1 of the degree-sorted graph) must equal
    push-only on the same sorted graph: PageRank within 1e-10, BFS levels and
    Radii exactly."""
    hg = graphs(gname)
    print(f"radar-vs-push {gname}", flush=True)
    dg, ids = rg.DeviceGraph(ctx, hg).degree_sort(by_outdeg=False, sort_neighbors=False)
    off, nb = dg.download()
    a = dg.pagerank(20, rg.opts("push"))
    b = dg.pagerank(20, rg.opts("radar", hubs=4096))
    assert a.iters == b.iters
    assert np.max(np.abs(a.ranks - b.ranks)) / np.max(np.abs(a.ranks)) < 1e-10
    src = int(ids[hg.pick_sources(1)[0]])
    lv, ecc = orc.bfs_levels(off, nb, src)
    r = dg.bfs(src, rg.opts("radar", hubs=8192))
    assert orc.check_parents(off, nb, src, lv, r.parents) == 0 and r.iters == ecc + 1
    starts = ids[rg.radii_starts(hg.n)]
    ra, _, _ = dg.radii(starts, rg.opts("push"))
    rb, _, _ = dg.radii(starts, rg.opts("radar", hubs=4096))
    assert np.array_equal(ra, rb)
