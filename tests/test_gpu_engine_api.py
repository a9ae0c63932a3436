"""GPU tests of the engine-level C ABI -- Ligra's vertexSubset, edgeMap,
vertexMap and vertexFilter (PNO Ligra/ligra/ligra.h:264-302,336-340,
447-509; vertexSubset.h:113-229) -- driven through ctypes as the
reference's apps drive the C++ templates, and checked against the oracle:
BFS (BFS.C:73-117), PageRank (PageRank.C:92-141), PageRankDelta
(PageRankDelta.C:90-155), Radii (Radii.C:98-157) and one BC source
(BC-manyiter-v2.C:124-241), each app loop written out here with the
reference's phases.  Integers bit-exact, parents a valid shortest-path tree,
FP64 within 1e-10 relative L-inf (L1 norms within 1e-10 of the rank sum)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10
UINT_MAX = 0xFFFFFFFF


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    fa, fb = np.isfinite(a), np.isfinite(b)
    assert np.array_equal(fa, fb)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    if not fa.any():
        return 0.0
    return float(np.max(np.abs(a[fa] - b[fb])) / max(np.max(np.abs(b[fb])), 1e-300))


def levels_of(off, nbr, src, parents):
    """Levels implied by a parent array (walks to the root)."""
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


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12d"])
@pytest.mark.parametrize("variant", ["push", "hubdup", "radar"])
def test_bfs_through_edgemap(rg, orc, ctx, graphs, gname, variant):
    hg = graphs(gname)
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    o = rg.opts(variant, hubs=128)
    for src in hg.pick_sources(3):
        src = int(src)
        print(f"bfs edgemap {gname} {variant} src {src}", flush=True)
        parents = rg.DeviceArray(ctx, dg.n, np.uint32)
        parents.fill(UINT_MAX)
        parents.write([src], first=src)
        frontier = rg.Subset(dg, [src])
        it = 0
        while frontier.size() > 0:   # BFS.C:93-100
            nxt = dg.edgemap(rg.OP_BFS, frontier, [parents], flags=rg.DENSE_FORWARD, o=o)
            frontier.close()
            frontier = nxt
            it += 1
        p = parents.read()
        lv, ecc = orc.bfs_levels(off, nbr, src)
        assert it == ecc + 1
        assert orc.check_parents(off, nbr, src, lv, p) == 0
        assert int(np.sum(p != UINT_MAX)) == int(np.sum(lv >= 0))
        frontier.close()
        parents.close()


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12s"])
@pytest.mark.parametrize("variant", ["push", "hubdup", "radar"])
def test_pagerank_through_edgemap(rg, orc, ctx, graphs, gname, variant):
    hg = graphs(gname)
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    n = dg.n
    o = rg.opts(variant, hubs=256)
    maxiters, damping, eps = 20, 0.85, 1e-7
    print(f"pagerank edgemap {gname} {variant}", flush=True)
    p_curr, p_next, contrib = (rg.DeviceArray(ctx, n) for _ in range(3))
    p_curr.fill(1.0 / n)
    p_next.fill(0.0)
    frontier = rg.Subset(dg, all_vertices=True)
    it, l1 = 0, 0.0
    while True:   # PageRank.C:120-133
        it += 1
        if it > maxiters:
            break
        dg.vertexmap(rg.VOP_PR_CONTRIB, frontier, [p_curr, contrib])
        dg.edgemap(rg.OP_PAGERANK, frontier, [contrib, p_next], flags=rg.NO_OUTPUT | rg.DENSE_FORWARD, o=o)
        dg.vertexmap(rg.VOP_PR_UPDATE, frontier, [p_next], damping, (1 - damping) * (1 / n))
        dg.vertexmap(rg.VOP_ABS_DIFF, frontier, [p_curr, p_next])
        l1 = p_curr.sum()
        if l1 < eps:
            break
        p_curr.fill(0.0)
        p_curr, p_next = p_next, p_curr
    ranks = (p_curr if it > maxiters else p_next).read()
    ref, ref_it, ref_l1 = orc.pagerank(off, nbr, maxiters)
    assert it == ref_it
    assert rel(ranks, ref) < FP_TOL
    assert abs(l1 - ref_l1) <= FP_TOL * float(np.sum(ref))


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12s", "rmat14s"])
@pytest.mark.parametrize("maxiters", [20, 3])
def test_pagerank_delta_through_vertexfilter(rg, orc, ctx, graphs, gname, maxiters):
    hg = graphs(gname)
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    n, m = dg.n, dg.m
    damping, eps, eps2 = 0.85, 1e-7, 0.01
    print(f"pagerank_delta vertexfilter {gname} {maxiters}", flush=True)
    p, delta, ngh, contrib = (rg.DeviceArray(ctx, n) for _ in range(4))
    p.fill(0.0)
    delta.fill(1.0 / n)
    ngh.fill(0.0)
    frontier = rg.Subset(dg, all_vertices=True)
    every = rg.Subset(dg, all_vertices=True)
    rnd, l1, sizes = 0, 0.0, []
    while True:   # PageRankDelta.C:128-144
        rnd += 1
        if rnd > maxiters:
            break
        dg.vertexmap(rg.VOP_PR_CONTRIB, every, [delta, contrib])
        dg.edgemap(rg.OP_PAGERANK, frontier, [contrib, ngh], threshold=m // 20,
                   flags=rg.NO_OUTPUT | rg.DENSE_FORWARD)
        if rnd == 1:
            active = dg.vertexfilter(rg.FOP_PRD_FIRST, every, [p, delta, ngh], (damping, 1.0 / n, eps2))
        else:
            active = dg.vertexfilter(rg.FOP_PRD, every, [p, delta, ngh], (damping, eps2))
        ids = active.ids()
        assert np.all(np.diff(ids.astype(np.int64)) > 0), "vertexFilter output not ascending"
        sizes.append(len(ids))
        dg.vertexmap(rg.VOP_ABS, every, [ngh, delta])
        l1 = ngh.sum()
        if l1 < eps:
            break
        ngh.fill(0.0)
        frontier.close()
        frontier = active
    ref, ref_it, ref_l1 = orc.pagerank_delta(off, nbr, maxiters)
    assert rnd == ref_it
    assert rel(p.read(), ref) < FP_TOL
    assert abs(l1 - ref_l1) <= FP_TOL * float(np.sum(ref))
    # The frontier shrinks: the sparse rounds (|F| + degsum <= m/20) ran too.
    if maxiters == 20:
        assert min(sizes) < n


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d"])
def test_radii_through_edgemap(rg, orc, ctx, graphs, gname):
    hg = graphs(gname)
    off, nbr = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    n = dg.n
    starts = rg.radii_starts(n)
    print(f"radii edgemap {gname}", flush=True)
    vis, nvis = rg.DeviceArray(ctx, n, np.uint64), rg.DeviceArray(ctx, n, np.uint64)
    radii = rg.DeviceArray(ctx, n, np.uint32)
    vis.fill(0)
    nvis.fill(0)
    radii.fill(UINT_MAX)   # -1
    for i, v in enumerate(starts):   # Radii.C:120-127
        radii.write([0], first=int(v))
        nvis.write(np.array([1 << i], np.uint64), first=int(v))
    frontier = rg.Subset(dg, starts)
    rnd = 0
    while frontier.size() > 0:   # Radii.C:132-140
        rnd += 1
        dg.vertexmap(rg.VOP_RADII_SYNC, frontier, [vis, nvis])
        nxt = dg.edgemap(rg.OP_RADII, frontier, [vis, nvis, radii], round_=rnd, flags=rg.DENSE_FORWARD)
        frontier.close()
        frontier = nxt
    ref, ref_rounds = orc.radii(off, nbr, starts)
    assert rnd == ref_rounds
    assert np.array_equal(radii.read().view(np.int32), ref)


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d"])
def test_bc_through_edgemap(rg, orc, ctx, graphs, gname):
    hg = graphs(gname)
    off, nbr = hg.csr()
    ioff, inbr = hg.csr(in_edges=True)
    dg = rg.DeviceGraph(ctx, hg)
    n = dg.n
    src = int(hg.pick_bc_sources(1)[0])
    print(f"bc edgemap {gname} src {src}", flush=True)
    paths, deps = rg.DeviceArray(ctx, n), rg.DeviceArray(ctx, n)
    visited = rg.DeviceArray(ctx, n, np.uint8)
    deps.fill(0.0)   # BC-manyiter-v2.C:145
    paths.fill(0.0)
    paths.write([1.0], first=src)
    visited.fill(0)
    visited.write(np.array([1], np.uint8), first=src)
    every = rg.Subset(dg, all_vertices=True)
    frontier = rg.Subset(dg, [src])
    levels = [frontier]
    rnd = 0
    while frontier.size() > 0:   # BC-manyiter-v2.C:190-198
        rnd += 1
        out = dg.edgemap(rg.OP_BC_FWD, frontier, [paths, visited], flags=rg.DENSE_FORWARD)
        dg.vertexmap(rg.VOP_MARK, out, [visited])
        levels.append(out)
        frontier = out
    dg.vertexmap(rg.VOP_INVERT, every, [paths])   # inverseNumPaths, 206-208
    visited.fill(0)
    frontier = levels[rnd - 1]
    dg.vertexmap(rg.VOP_BC_BACK, frontier, [visited, deps, paths])
    dg.transpose()   # 217 (never undone)
    for r in range(rnd - 2, -1, -1):   # 218-225
        dg.edgemap(rg.OP_BC_BACK, frontier, [deps, visited], flags=rg.NO_OUTPUT | rg.DENSE_FORWARD)
        frontier = levels[r]
        dg.vertexmap(rg.VOP_BC_BACK, frontier, [visited, deps, paths])
    dg.vertexmap(rg.VOP_BC_FINAL, every, [deps, paths])   # 230-232
    ref, comb = orc.bc(off, nbr, ioff, inbr, [src])
    assert rnd == comb
    assert rel(deps.read(), ref) < FP_TOL


def test_vertexfilter_and_subsets(rg, ctx, graphs):
    hg = graphs("rmat12d")
    dg = rg.DeviceGraph(ctx, hg)
    n = dg.n
    rs = np.random.RandomState(7)
    x = rs.randint(0, 3, n).astype(np.uint32)
    xa = rg.DeviceArray(ctx, n, np.uint32)
    xa.write(x)
    every = rg.Subset(dg, all_vertices=True)
    assert every.size() == n and np.array_equal(every.ids(), np.arange(n, dtype=np.uint32))
    out = dg.vertexfilter(rg.FOP_U32_EQ, every, [xa], (1,))
    assert np.array_equal(out.ids(), np.flatnonzero(x == 1).astype(np.uint32))
    # A listed input in any order: the output is the kept ones, ascending.
    lst = rs.permutation(n)[: n // 3].astype(np.uint32)
    sub = rg.Subset(dg, lst)
    assert sub.size() == len(lst) and np.array_equal(sub.ids(), lst)
    out2 = dg.vertexfilter(rg.FOP_U32_EQ, sub, [xa], (2,))
    assert np.array_equal(out2.ids(), np.sort(lst[x[lst] == 2]))
    flags = (rs.rand(n) < 0.01).astype(np.uint8)
    fa = rg.DeviceArray(ctx, n, np.uint8)
    fa.write(flags)
    out3 = dg.vertexfilter(rg.FOP_U8_NONZERO, every, [fa])
    assert np.array_equal(out3.ids(), np.flatnonzero(flags).astype(np.uint32))
    # Nothing kept, and an empty input.
    out4 = dg.vertexfilter(rg.FOP_U32_EQ, every, [xa], (7,))
    assert out4.size() == 0
    empty = rg.Subset(dg, [])
    assert dg.vertexfilter(rg.FOP_U32_EQ, empty, [xa], (1,)).size() == 0
    # Flags left clear: the same filter again gives the same answer.
    assert np.array_equal(dg.vertexfilter(rg.FOP_U32_EQ, every, [xa], (1,)).ids(), out.ids())
    # Bad arguments are errors, not aborts.
    with pytest.raises(rg.RGError):
        dg.vertexfilter(rg.FOP_U32_EQ, every, [fa], (1,))   # wrong element size
    with pytest.raises(rg.RGError):
        dg.vertexfilter(99, every, [xa], (1,))
    # plusReduce in a fixed order: the same bits on every call.
    vals = rg.DeviceArray(ctx, n)
    v = rs.rand(n)
    vals.write(v)
    s1, s2 = vals.sum(), vals.sum()
    assert s1 == s2 and abs(s1 - float(np.sum(v))) <= 1e-12 * float(np.sum(v))
