"""GPU tests of the serving loop: continuous batching of BFS queries over
several slots, mixed with BC and PageRank queries.  Every answer in the mix
equals the same query run alone (BFS summaries bit-exact; BC / PageRank
result summaries within 1e-10 relative, non-finite counts exact), and the
alone answers are checked against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10


def weights(orc, n):
    v = np.arange(n, dtype=np.uint64)
    h = orc.level_hash(v, np.full(n, 0xFFFFFFFF, np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def fp_summary(orc, x):
    f = np.isfinite(x)
    return float(np.sum(x[f] * weights(orc, len(x))[f])), int((~f).sum())


def close(a, b, tol=FP_TOL):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("name", ["rmat14s", "rmat12d"])
def test_mix_equals_alone(rg, orc, ctx, graphs, name):
    hg = graphs(name)
    dg = rg.DeviceGraph(ctx, hg)
    off, nb = hg.csr()
    in_off, in_nb = hg.csr(in_edges=True) if not hg.symmetric else (off, nb)
    q = rg.poisson_workload(hg, 120, 20000.0, seed=5)
    print("serve", name, len(q), flush=True)
    res, st = dg.serve(q, rg.serve_config(slots=3))
    assert st.queries == len(q)
    sizes = [r.batch_size for r in res if r.kind == rg.Q_BFS]
    assert max(sizes) > 1, "no BFS queries were coalesced"
    pr_alone = None
    for i, (qq, r) in enumerate(zip(q, res)):
        assert r.kind == qq.kind
        assert r.latency_ms >= 0 and r.end_ms >= r.start_ms >= r.arrival_ms - 1e-6
        if qq.kind == rg.Q_BFS:
            alone, _ = dg.bfs_query(qq.source)
            assert r.bfs.astuple() == alone.astuple(), f"query {i}"
            lv, _ = orc.bfs_levels(off, nb, qq.source)
            assert alone.astuple() == orc.bfs_summary(off, lv)
        elif qq.kind == rg.Q_BC:
            deps, comb, _ = dg.bc([qq.source])
            if not hg.symmetric:
                dg.transpose()   # BC leaves a directed graph transposed (BC-manyiter-v2.C:217)
            ws, nf = fp_summary(orc, deps)
            assert r.nonfinite == nf and close(r.wsum, ws), f"query {i}"
            ref, _ = orc.bc(off, nb, in_off, in_nb, [qq.source])
            ws2, nf2 = fp_summary(orc, ref)
            assert nf == nf2 and close(ws, ws2)
        else:
            if pr_alone is None:
                pr_alone = dg.pagerank(qq.iters)
                ref, it, _ = orc.pagerank(off, nb, qq.iters)
                assert pr_alone.iters == it
                assert np.max(np.abs(pr_alone.ranks - ref)) <= FP_TOL * np.max(np.abs(ref))
            ws, nf = fp_summary(orc, pr_alone.ranks)
            assert r.iters == pr_alone.iters and r.nonfinite == nf and close(r.wsum, ws), f"query {i}"


def test_slots_overlap_and_batches_grow_under_load(rg, ctx, graphs):
    hg = graphs("rmat14s")
    dg = rg.DeviceGraph(ctx, hg)
    q = rg.poisson_workload(hg, 600, 0.0, mix=(1, 0, 0), seed=9)   # everything arrives at t = 0
    res, st = dg.serve(q, rg.serve_config(slots=3, max_batch=64))
    slots = {r.slot for r in res}
    assert st.slots == 3 and len(slots) > 1
    # batches of different slots run at the same time
    iv = sorted({(r.slot, r.start_ms, r.end_ms) for r in res}, key=lambda t: t[1])
    assert any(a[0] != b[0] and b[1] < a[2] for a, b in zip(iv, iv[1:]))
    assert st.mean_batch > 8 and max(r.batch_size for r in res) == 64
    # items start (on the trace's clock) in the order the scheduler took them
    starts = dict(sorted({(r.batch_id, r.start_ms) for r in res}))
    ordered = [starts[b] for b in sorted(starts)]
    assert all(a <= b for a, b in zip(ordered, ordered[1:]))


def test_one_at_a_time_same_answers(rg, ctx, graphs):
    hg = graphs("rmat12s")
    dg = rg.DeviceGraph(ctx, hg)
    q = rg.poisson_workload(hg, 80, 5000.0, seed=3)
    a, sa = dg.serve(q, rg.serve_config(slots=2))
    b, sb = dg.serve(q, rg.serve_config(one_at_a_time=True))
    assert sb.slots == 1 and max(r.batch_size for r in b) == 1
    for x, y in zip(a, b):
        assert x.kind == y.kind
        if x.kind == rg.Q_BFS:
            assert x.bfs.astuple() == y.bfs.astuple()
        else:
            assert x.nonfinite == y.nonfinite and close(x.wsum, y.wsum)


def test_admission_by_memory(rg, orc, ctx, graphs):
    # Items are admitted against the device memory the slots' query state
    # holds.  With room for little more than one slot's full state (reserved
    # on the first slot before the clock starts), the two BC batches of the
    # trace cannot both hold a BC batch's state: one waits for memory or runs
    # as narrower batches,
    # the measured peak stays within the budget, and every answer is the one
    # the query gets alone.  With room to spare nothing waits; with room for
    # less than one slot the configuration is refused.
    hg = graphs("rmat16s")
    off, nb = hg.csr()
    dg = rg.DeviceGraph(ctx, hg)
    q = rg.poisson_workload(hg, 100, 0.0, mix=(0.0, 1.0, 0.0), seed=4)   # two BC batches of 64 and 36
    dg.serve(q, rg.serve_config(slots=4))   # builds the layouts (their build's scratch is freed after)
    res, st = dg.serve(q, rg.serve_config(slots=4))
    free = st.hbm_budget + rg.serve_config().hbm_reserve
    print("admission, room to spare:", st.asdict(), flush=True)
    assert st.slots == 4 and st.mem_waits == 0 and st.peak_bytes <= st.hbm_budget
    tight = rg.serve_config(slots=4, hbm_reserve=int(free - 1.15 * st.slot_bytes))
    res1, st1 = dg.serve(q, tight)
    print("admission, tight:", st1.asdict(), flush=True)
    assert st1.slots == 4 and st1.mem_waits > 0
    assert st1.peak_bytes <= st1.hbm_budget
    for a, b in zip(res, res1):
        if a.kind == rg.Q_BC:
            assert a.nonfinite == b.nonfinite and close(a.wsum, b.wsum)
    with pytest.raises(rg.RGError):
        dg.serve(q, rg.serve_config(slots=4, hbm_reserve=int(free - st.slot_bytes // 2)))


def test_bc_queries_batched(rg, ctx, graphs):
    # BC queries that have arrived are answered in batches of up to 64; each
    # answer equals the same query answered alone.
    hg = graphs("rmat12d")
    dg = rg.DeviceGraph(ctx, hg)
    q = rg.poisson_workload(hg, 150, 0.0, mix=(0.2, 0.8, 0.0), seed=12)
    a, st = dg.serve(q, rg.serve_config(slots=2))
    b, st1 = dg.serve(q, rg.serve_config(one_at_a_time=True))
    sizes = [r.batch_size for r in a if r.kind == rg.Q_BC]
    assert max(sizes) > 1 and max(sizes) <= 64
    # the stats' mean widths: queries of a kind per work item of that kind
    for kind, mean, mean1 in ((rg.Q_BFS, st.mean_batch, st1.mean_batch), (rg.Q_BC, st.mean_bc_batch, st1.mean_bc_batch)):
        items = {r.batch_id for r in a if r.kind == kind}
        assert mean == pytest.approx(sum(r.kind == kind for r in a) / len(items)) and mean1 == 1.0
    assert st.mean_bc_batch > 1
    for i, (x, y) in enumerate(zip(a, b)):
        assert x.kind == y.kind
        if x.kind == rg.Q_BC:
            assert (x.iters, x.nonfinite) == (y.iters, y.nonfinite) and close(x.wsum, y.wsum), f"query {i}"
        else:
            assert x.bfs.astuple() == y.bfs.astuple()



@pytest.mark.parametrize("max_batch", [2, 4])
def test_small_batch_limit(rg, ctx, graphs, max_batch):
    # A batch limit below the calibrated widths (1, 8, 64, 256): the cost
    # curve is timed at the widths up to the limit, batches never exceed
    # it, and every answer equals the one-at-a-time answer.
    hg = graphs("rmat12s")
    dg = rg.DeviceGraph(ctx, hg)
    q = rg.poisson_workload(hg, 120, 0.0, mix=(0.6, 0.3, 0.1), seed=31)
    a, st = dg.serve(q, rg.serve_config(slots=2, max_batch=max_batch))
    b, _ = dg.serve(q, rg.serve_config(one_at_a_time=True))
    assert max(r.batch_size for r in a) <= max_batch
    for i, (x, y) in enumerate(zip(a, b)):
        assert x.kind == y.kind
        if x.kind == rg.Q_BFS:
            assert x.bfs.astuple() == y.bfs.astuple(), f"query {i}"
        else:
            assert x.nonfinite == y.nonfinite and close(x.wsum, y.wsum), f"query {i}"


@pytest.mark.parametrize("one_at_a_time", [True, False])
def test_hubdup_views_leave_base_metadata(rg, orc, graphs, one_at_a_time):
    # The serving slots' views share the base graph's hub metadata, heavy
    # list and layouts.  A HUBDUP BC query transposes a directed view, whose
    # hub metadata (built for the other orientation) is then rebuilt: the
    # view builds its own and leaves the base's arrays alone, so the base
    # still answers HUBDUP queries correctly afterwards and frees cleanly.
    hg = graphs("rmat12d")
    off, nb = hg.csr()
    in_off, in_nb = hg.csr(in_edges=True)
    c = rg.Context(0)
    try:
        dg = rg.DeviceGraph(c, hg)
        o = rg.opts("hubdup", hubs=256)
        ref, it, _ = orc.pagerank(off, nb, 20)
        before = dg.pagerank(20, o)   # base builds its hub metadata and heavy list
        assert before.iters == it
        q = rg.poisson_workload(hg, 40, 0.0, mix=(0.4, 0.4, 0.2), seed=11)
        print("serve hubdup views", one_at_a_time, flush=True)
        res, st = dg.serve(q, rg.serve_config(slots=2, one_at_a_time=one_at_a_time, o=o))
        assert st.queries == len(q)
        for qq, r in zip(q, res):
            if qq.kind == rg.Q_BC:
                ref_bc, _ = orc.bc(off, nb, in_off, in_nb, [qq.source])
                ws, nf = fp_summary(orc, ref_bc)
                assert r.nonfinite == nf and close(r.wsum, ws)
        after = dg.pagerank(20, o)
        assert after.iters == it
        assert np.max(np.abs(after.ranks - ref)) <= FP_TOL * np.max(np.abs(ref))
        src = int(hg.pick_sources(1)[0])
        b = dg.bfs(src, rg.opts("hubdup", hubs=128))
        lv, _ = orc.bfs_levels(off, nb, src)
        assert orc.check_parents(off, nb, src, lv, b.parents) == 0
        dg.close()
    finally:
        c.close()


@pytest.mark.parametrize("alone", [0, 1 << 2])
def test_pagerank_alone_policy(rg, graphs, alone):
    # alone_kinds: a PageRank query never runs beside another item (the other
    # slots drain first and start nothing until it ends); without it items
    # overlap it.  Answers are the same either way.
    hg = graphs("rmat14s")
    c = rg.Context(0)
    try:
        dg = rg.DeviceGraph(c, hg)
        q = rg.poisson_workload(hg, 120, 0.0, mix=(0.6, 0.2, 0.2), seed=21)
        res, st = dg.serve(q, rg.serve_config(slots=3, alone_kinds=alone))
        ref, _ = dg.serve(q, rg.serve_config(one_at_a_time=True))
        items = {(r.batch_id, r.kind, r.slot, r.start_ms, r.end_ms) for r in res}
        pr = [x for x in items if x[1] == rg.Q_PAGERANK]
        overlap = any(o[0] != p[0] and o[3] < p[4] and p[3] < o[4] for p in pr for o in items)
        print("alone", alone, "PageRank items", len(pr), "overlap", overlap, flush=True)
        assert len(pr) > 0
        if alone:
            assert not overlap
        for a, b in zip(res, ref):
            if a.kind == rg.Q_BFS:
                assert a.bfs.astuple() == b.bfs.astuple()
            else:
                assert a.nonfinite == b.nonfinite and close(a.wsum, b.wsum)
    finally:
        c.close()
