"""GPU parity sweep on a larger graph than the per-app tests: GAP R-MAT
scale 18 (symmetric and directed), every engine, against the oracle --
PageRank and PageRankDelta within 1e-10 relative L-inf with the oracle's
iteration counts, BFS parents a valid shortest-path tree with the oracle's
levels, Radii exact, BC within 1e-10 (NaN positions exact), and on the
degree-sorted graph the bit-parallel batch's levels and the triangle count
exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10
ENGINES = ["push", "hubdup", "radar", "pull", "blocked"]
_CACHE = {}


def rel(a, b):
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb), "NaN positions differ"
    a, b = a[~na], b[~nb]
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)) if a.size else 0.0


def refs(rg, orc, sym):
    if sym not in _CACHE:
        hg = rg.HostGraph.generate(18, 16, symmetric=sym)
        off, nbr = hg.csr()
        ioff, inb = hg.csr(in_edges=True)
        srcs = hg.pick_sources(3)
        bsrc = hg.pick_bc_sources(3)
        starts = rg.radii_starts(hg.n)
        _CACHE[sym] = dict(
            hg=hg, off=off, nbr=nbr, srcs=srcs, bsrc=bsrc, starts=starts,
            pr=orc.pagerank(off, nbr, 20), prd=orc.pagerank_delta(off, nbr, 20),
            lv=[orc.bfs_levels(off, nbr, int(s)) for s in srcs],
            radii=orc.radii(off, nbr, starts), bc=orc.bc(off, nbr, ioff, inb, bsrc))
    return _CACHE[sym]


@pytest.mark.parametrize("sym", [True, False])
@pytest.mark.parametrize("engine", ENGINES)
def test_engine_matches_oracle(rg, orc, ctx, sym, engine):
    R = refs(rg, orc, sym)
    print(f"parity rmat18 sym={sym} {engine}", flush=True)
    dg = rg.DeviceGraph(ctx, R["hg"])
    o = rg.opts(engine, hubs=4096)
    ref, it, _ = R["pr"]
    r = dg.pagerank(20, o)
    assert r.iters == it and rel(r.ranks, ref) < FP_TOL
    if engine != "pull":
        ref, it, _ = R["prd"]
        r = dg.pagerank_delta(20, o)
        assert r.iters == it and rel(r.ranks, ref) < FP_TOL
    for s, (lv, ecc) in zip(R["srcs"], R["lv"]):
        r = dg.bfs(int(s), o)
        assert orc.check_parents(R["off"], R["nbr"], int(s), lv, r.parents) == 0 and r.iters == ecc + 1
    rad, rounds, _ = dg.radii(R["starts"], o)
    assert (rounds, rad.tolist()) == (R["radii"][1], R["radii"][0].tolist())
    if engine != "pull":
        deps, comb, _ = dg.bc(R["bsrc"], o)
        assert comb == R["bc"][1] and rel(deps, R["bc"][0]) < FP_TOL


@pytest.mark.parametrize("sym", [True, False])
def test_sorted_graph_batch_and_tc(rg, orc, ctx, sym):
    R = refs(rg, orc, sym)
    dg = rg.DeviceGraph(ctx, R["hg"])
    sg, ids = dg.degree_sort(sort_neighbors=True)
    soff, snb = sg.download()
    _, lev, _ = sg.bfs_batch(ids[R["srcs"]], rg.opts("blocked"), want_levels=True)
    for i, s in enumerate(ids[R["srcs"]]):
        assert np.array_equal(lev[i], orc.bfs_levels(soff, snb, int(s))[0])
    if sym:
        tot, _ = sg.tc()
        assert tot == orc.tc(soff, snb)


@pytest.mark.parametrize("sym", [True, False])
def test_batches_in_both_launch_forms(rg, orc, ctx, sym):
    # The bit-parallel batches on the degree-sorted graph (split hub blocks,
    # vertices above the 4096-arc chunk threshold) as per-level kernels and
    # as one persistent launch: BFS levels of 256 sources equal, and equal to
    # the oracle for a sample; Radii exact; BC rows of 64 sources equal
    # between the forms and the oracle's single-source BC within 1e-10 for a
    # sample.
    R = refs(rg, orc, sym)
    dg = rg.DeviceGraph(ctx, R["hg"])
    sg, ids = dg.degree_sort(sort_neighbors=True)
    soff, snb = sg.download()
    print(f"batches rmat18 sym={sym}", flush=True)
    srcs = ids[R["hg"].pick_sources(256)]
    a, la, _ = sg.bfs_batch(srcs, rg.opts(batch_launch=1), want_levels=True)
    b, lb, _ = sg.bfs_batch(srcs, rg.opts(batch_launch=2), want_levels=True)
    assert [x.astuple() for x in a] == [x.astuple() for x in b]
    assert np.array_equal(la, lb)
    for i in (0, 100, 255):
        assert np.array_equal(lb[i], orc.bfs_levels(soff, snb, int(srcs[i]))[0])
    starts = ids[R["starts"]]
    for launch in (1, 2):
        rad, rounds, _ = sg.radii(starts, rg.opts("blocked", batch_launch=launch))
        ref_rad, ref_rounds = orc.radii(soff, snb, starts)
        assert rounds == ref_rounds and np.array_equal(rad, ref_rad), launch
    bsrc = ids[R["hg"].pick_bc_sources(64)]
    sioff, sinb = (soff, snb) if sym else sg.download(in_edges=True)
    rows = {}
    for launch in (1, 2):
        out, deps, _ = sg.bc_batch(bsrc, rg.opts("radar", batch_launch=launch), want_deps=True)
        rows[launch] = (out, deps)
    (oa, da), (ob, db) = rows[1], rows[2]
    assert [(x.rounds, x.reached, x.nonfinite) for x in oa] == [(x.rounds, x.reached, x.nonfinite) for x in ob]
    for i in range(len(bsrc)):
        assert rel(db[i], da[i]) < FP_TOL, i
    for i in (0, 31, 63):
        ref, rounds = orc.bc(soff, snb, sioff, sinb, [int(bsrc[i])])
        assert rel(db[i], ref) < FP_TOL and ob[i].rounds == rounds, i
