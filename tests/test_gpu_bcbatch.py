"""GPU tests of the batched single-source BC (rg_bc_batch, 1..256 sources
in batches of 64, a last batch under 12 sources one at a time unless
batch_launch is 1): every row of
the batch's dependencies equals the oracle's single-source BC for that
source within 1e-10 relative L-inf (NaN at exactly the unreached vertices),
and every summary equals the one computed from the oracle's array (rounds,
reached and non-finite counts exact, weighted sum within 1e-10)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP_TOL = 1e-10
BIG = 1 << 60


def rel_linf(a, b):
    nan_a, nan_b = np.isnan(a), np.isnan(b)
    assert np.array_equal(nan_a, nan_b), "NaN positions differ"
    a, b = a[~nan_a], b[~nan_b]
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(b))
    if a.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def close(a, b, tol=FP_TOL):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300)


def pick(n, k, seed):
    return np.random.default_rng(seed).integers(0, n, size=k).astype(np.uint32)


def check(rg, orc, hg, dg, srcs, o=None):
    print("bc batch", dg.n, dg.m, len(srcs), flush=True)
    off, nb = hg.csr()
    ioff, inb = hg.csr(in_edges=True)
    out, deps, _ = dg.bc_batch(srcs, o, want_deps=True)
    w = rg.summary_weights(hg.n)
    for i, s in enumerate(srcs):
        ref, rounds = orc.bc(off, nb, ioff, inb, [int(s)])
        assert rel_linf(deps[i], ref) < FP_TOL, f"source {s} (slot {i})"
        fin = np.isfinite(ref)
        lv, _ = orc.bfs_levels(off, nb, int(s))
        assert (out[i].rounds, out[i].reached, out[i].nonfinite) == (rounds, int((lv >= 0).sum()),
                                                                      int((~fin).sum())), f"source {s}"
        assert close(out[i].wsum, float(np.sum(ref[fin] * w[fin]))), f"source {s}"
    return out, deps


# launch 0: the library's choice (batches under 12 sources run one source
# at a time); launch 1: every batch on the bit-parallel engine as per-level
# kernels; launch 2: every batch as one persistent launch (forward levels,
# path counts and backward levels in one grid).
@pytest.mark.parametrize("launch", [0, 1, 2])
@pytest.mark.parametrize("name", ["rmat12s", "rmat12d", "er12s", "er12d"])
@pytest.mark.parametrize("k", [1, 5, 17, 64])
def test_bc_batch_matches_oracle(rg, orc, ctx, graphs, name, k, launch):
    hg = graphs(name)
    dg = rg.DeviceGraph(ctx, hg)
    check(rg, orc, hg, dg, pick(hg.n, k, 300 + k), rg.opts(batch_launch=launch))


@pytest.mark.parametrize("launch", [1, 2])
@pytest.mark.parametrize("threshold", [0, BIG])
def test_bc_batch_all_dense_or_all_sparse(rg, orc, ctx, graphs, threshold, launch):
    # The forward levels' kind (blocked dense or push) must not change the
    # answer: the path counts are summed by the same hook either way.
    hg = graphs("rmat12d")
    dg = rg.DeviceGraph(ctx, hg)
    check(rg, orc, hg, dg, pick(hg.n, 64, 11), rg.opts(threshold=threshold, batch_launch=launch))


def test_bc_batch_duplicates_and_isolated(rg, orc, ctx, graphs):
    hg = graphs("rmat12s")
    dg = rg.DeviceGraph(ctx, hg)
    deg = np.diff(hg.offsets())
    iso = np.nonzero(deg == 0)[0][:3].astype(np.uint32)
    srcs = np.concatenate([np.full(4, 9, np.uint32), iso, pick(hg.n, 30, 5), np.full(2, 9, np.uint32)])
    out, deps = check(rg, orc, hg, dg, srcs)
    for i in range(len(iso)):
        assert out[4 + i].rounds == 1 and out[4 + i].reached == 1
    assert np.array_equal(deps[0], deps[1], equal_nan=True)


@pytest.mark.parametrize("symmetric", [True, False])
def test_bc_batch_heavy_vertices(rg, orc, ctx, symmetric):
    # Hubs of in- and out-degree 20000 (above the 4096 chunk threshold):
    # their path counts and dependencies are summed over chunk items.
    leaves = 20000
    hub_out, hub_in = 0, leaves + 1
    lv = np.arange(1, leaves + 1, dtype=np.uint32)
    src = np.concatenate([np.full(leaves, hub_out, np.uint32), lv, lv[:3000]])
    dst = np.concatenate([lv, np.full(leaves, hub_in, np.uint32), lv[3000:6000]])
    hg = rg.HostGraph.from_edges(leaves + 2, src, dst, symmetric=symmetric)
    dg = rg.DeviceGraph(ctx, hg)
    srcs = np.array([hub_out, hub_in, 5, 4000, 2999, 17000], np.uint32)
    check(rg, orc, hg, dg, srcs, rg.opts(batch_launch=1))
    check(rg, orc, hg, dg, srcs, rg.opts(threshold=0, batch_launch=1))
    check(rg, orc, hg, dg, srcs, rg.opts(batch_launch=2))
    check(rg, orc, hg, dg, srcs, rg.opts(threshold=0, batch_launch=2))
    check(rg, orc, hg, dg, srcs)


@pytest.mark.parametrize("launch", [1, 2])
@pytest.mark.parametrize("length", [64, 65, 100])
def test_bc_batch_deep_levels(rg, orc, ctx, length, launch):
    # A path: from vertex 0 the last level is length - 1.  Up to level 63
    # the batch engine answers; deeper sources are answered by single-source
    # runs.
    a = np.arange(length - 1, dtype=np.uint32)
    hg = rg.HostGraph.from_edges(length, a, a + 1, symmetric=True)
    dg = rg.DeviceGraph(ctx, hg)
    check(rg, orc, hg, dg, np.array([0, length // 2, length - 1], np.uint32), rg.opts(batch_launch=launch))


@pytest.mark.parametrize("launch", [1, 2])
def test_bc_batch_deep_then_regrown_batch(rg, orc, ctx, launch):
    # 70 sources on a path of 80 vertices: the first batch of 64 holds a
    # source 79 levels deep (single-source runs write its rows), the second
    # batch 41 levels deep (level words grown past the first 32, run again).
    length = 80
    a = np.arange(length - 1, dtype=np.uint32)
    hg = rg.HostGraph.from_edges(length, a, a + 1, symmetric=True)
    dg = rg.DeviceGraph(ctx, hg)
    srcs = np.concatenate([np.array([0], np.uint32), np.arange(10, 73, dtype=np.uint32),
                           np.full(6, 40, np.uint32)])
    check(rg, orc, hg, dg, srcs, rg.opts(batch_launch=launch))


@pytest.mark.parametrize("launch", [0, 1, 2])
@pytest.mark.parametrize("k", [65, 130, 256])
def test_bc_batch_uneven_batches(rg, orc, ctx, graphs, k, launch):
    # More than 64 sources: batches of 64 in order, the last one uneven
    # (with launch 0, a last batch of 1 or 2 sources runs one at a time).
    hg = graphs("rmat10s")
    dg = rg.DeviceGraph(ctx, hg)
    check(rg, orc, hg, dg, pick(hg.n, k, 40 + k), rg.opts(batch_launch=launch))


def test_bc_batch_bigger_graph(rg, orc, ctx, graphs):
    hg = graphs("rmat16s")
    dg = rg.DeviceGraph(ctx, hg)
    check(rg, orc, hg, dg, hg.pick_bc_sources(64))


def test_bc_batch_repeated_calls(rg, orc, ctx, graphs):
    # The per-context state is reused: a batch after a wider one and after a
    # different graph gives the same answers.
    big = graphs("rmat14s")
    small = graphs("rmat12d")
    d1 = rg.DeviceGraph(ctx, big)
    d2 = rg.DeviceGraph(ctx, small)
    o = rg.opts(batch_launch=1)
    check(rg, orc, big, d1, pick(big.n, 64, 1), o)
    check(rg, orc, small, d2, pick(small.n, 5, 2), o)
    check(rg, orc, big, d1, pick(big.n, 3, 3), o)


def test_bc_batch_errors(rg, ctx, graphs):
    hg = graphs("rmat10s")
    dg = rg.DeviceGraph(ctx, hg)
    with pytest.raises(rg.RGError):
        dg.bc_batch(np.arange(257, dtype=np.uint32))
    with pytest.raises(rg.RGError):
        dg.bc_batch(np.array([hg.n], np.uint32))
    with pytest.raises(rg.RGError):
        dg.bc_batch(np.array([], np.uint32))


def test_bc_batch_persistent_bigger_graph_and_repeats(rg, orc, ctx, graphs):
    # One persistent launch per batch on a larger graph, alternated with the
    # per-level form on the same context: both equal the oracle.
    hg = graphs("rmat16s")
    dg = rg.DeviceGraph(ctx, hg)
    srcs = hg.pick_bc_sources(64)
    a, da = check(rg, orc, hg, dg, srcs, rg.opts(batch_launch=2))
    b, db = check(rg, orc, hg, dg, srcs, rg.opts(batch_launch=1))
    assert [(x.rounds, x.reached, x.nonfinite) for x in a] == [(x.rounds, x.reached, x.nonfinite) for x in b]
    check(rg, orc, hg, dg, srcs[:20], rg.opts(batch_launch=2))
