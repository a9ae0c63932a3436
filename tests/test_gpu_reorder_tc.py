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


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d"])
@pytest.mark.parametrize("chunk", [1, 700, 5000])
def test_degree_sort_chunked_lists(rg, orc, ctx, graphs, gname, chunk, monkeypatch):
    """The chunked list sort that graphs of 2^31 arcs or more take, forced
    on a small graph: chunks of at most `chunk` arcs (hub lists longer than
    that are chunks of their own; chunk 1 makes every list one) give the
    same strict CSR as the oracle."""
    hg = graphs(gname)
    print(f"degree_sort chunked {gname} chunk={chunk}", flush=True)
    monkeypatch.setenv("RG_DEGREE_SORT_CHUNK", str(chunk))
    ng, ids = rg.DeviceGraph(ctx, hg).degree_sort(by_outdeg=False, sort_neighbors=True)
    off, nb = hg.csr()
    key = np.diff(hg.offsets(in_edges=not hg.symmetric)).astype(np.uint32)
    ref_ids = orc.degree_order(key)
    assert np.array_equal(ids, ref_ids)
    noff, nnb = orc.relabel_csr(off, nb, ref_ids, strict=True)
    goff, gnb = ng.download()
    assert np.array_equal(goff, noff) and np.array_equal(gnb, nnb)
    if not hg.symmetric:
        ioff, inb = hg.csr(in_edges=True)
        nioff, ninb = orc.relabel_csr(ioff, inb, ref_ids, strict=True)
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


@pytest.mark.parametrize("gname", ["rmat10s", "rmat12s", "er12s", "rmat14s"])
@pytest.mark.parametrize("relabel", [False, True])
def test_tc_total_only(rg, orc, ctx, graphs, gname, relabel):
    # The totals-only kernels (LDS bitmaps of N+(v) for v < 2^19, then the
    # flattened per-u kernel) against the oracle.
    hg = graphs(gname)
    dg = rg.DeviceGraph(ctx, hg)
    if relabel:
        dg, _ = dg.degree_sort(by_outdeg=True, sort_neighbors=True)
    off, nb = dg.download()
    tot, _ = dg.tc()
    assert tot == orc.tc(off, nb)


def test_tc_cliques_across_the_bitmap_split(rg, orc, ctx):
    # 2^18 disjoint 4-cliques on randomly permuted ids in [0, 2^20).  With
    # heavy_degree 1 every vertex below 2^19 (the bitmap's range) is counted
    # by the bitmap kernel, so cliques straddle the boundary between the two
    # totals kernels; with the default degree split (512) the flat kernel
    # counts them all.  Each clique holds 4 triangles.
    n = 1 << 20
    perm = np.random.default_rng(5).permutation(n).astype(np.uint32).reshape(-1, 4)
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    src = np.concatenate([perm[:, a] for a, _ in pairs])
    dst = np.concatenate([perm[:, b] for _, b in pairs])
    hg = rg.HostGraph.from_edges(n, src, dst, symmetric=True)
    dg = rg.DeviceGraph(ctx, hg)
    for o in (rg.opts(heavy_degree=1), None):
        tot, _ = dg.tc(o)
        assert tot == 4 * perm.shape[0]
    ptot, pv, _ = dg.tc(per_vertex=True)
    assert ptot == 4 * perm.shape[0] and np.all(pv == 3)


@pytest.mark.parametrize("dmin", [0, 2, 300, 600, 5000])
def test_tc_degree_split_with_hubs(rg, orc, ctx, dmin):
    # Hubs 0..7 of degree ~1000-1700 joined to each other and to random
    # light vertices, which are joined among themselves: triangles hub-hub-
    # light, hub-light-light and light-light-light, so the split (first
    # vertex of degree < dmin; 0 = the default 512) cuts through them.
    rng = np.random.default_rng(17)
    n, hubs = 20000, 8
    hs, hd = np.triu_indices(hubs, 1)
    spokes_s = np.repeat(np.arange(hubs), 1000 + 100 * np.arange(hubs))
    spokes_d = rng.integers(hubs, n, size=spokes_s.size)
    light_s = rng.integers(hubs, n, size=60000)
    light_d = rng.integers(hubs, n, size=60000)
    src = np.concatenate([hs, spokes_s, light_s]).astype(np.uint32)
    dst = np.concatenate([hd, spokes_d, light_d]).astype(np.uint32)
    hg = rg.HostGraph.from_edges(n, src, dst, symmetric=True)
    dg = rg.DeviceGraph(ctx, hg)
    dg, _ = dg.degree_sort(by_outdeg=True, sort_neighbors=True)
    off, nb = dg.download()
    want = orc.tc(off, nb)
    assert want > 0
    tot, _ = dg.tc(rg.opts(heavy_degree=dmin))
    assert tot == want


def test_tc_total_matches_per_vertex_path_rmat20(rg, ctx):
    # Both totals kernels on a degree-sorted R-MAT with 2^20 vertices, against
    # the per-vertex merge kernel (itself checked against the oracle above).
    hg = rg.HostGraph.generate(20, 16, symmetric=True)
    dg = rg.DeviceGraph(ctx, hg)
    sg, _ = dg.degree_sort(by_outdeg=True, sort_neighbors=True)
    tot, _ = sg.tc()
    ptot, pv, _ = sg.tc(per_vertex=True)
    assert tot == ptot and int(pv.sum()) == 3 * tot
