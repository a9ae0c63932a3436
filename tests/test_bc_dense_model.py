"""CPU model of the dense BC levels (csrc/bcdense.hip: k_bc_dense and
k_bc_merge, driven by bc_push's level loop in csrc/apps.hip) for the
reference's BC (BC-manyiter-v2.C:124-241).

Each dense launch is restated at workgroup granularity on the destination-
blocked layout: a unit (workgroup) sums, for every destination of its
block that is not yet visited, the values of its in-neighbours on the
frontier bitmap; a block held by one unit writes its destinations after a
workgroup barrier, a split block writes per-unit partials that a second
kernel adds in unit order.  The values are read from the same array the
destinations are written to (path counts forward, dependencies backward),
so the launch is free of cross-workgroup hazards only if no workgroup reads
a value that another writes.  The model asserts, on every dense level of
several sources of directed and symmetric R-MAT graphs with threshold 0
(every level dense) and small units (hub blocks split):

* every frontier source is visited, so no value read is a destination
  written (reads are frontier sources, writes unvisited destinations);
* the frontier bitmap read and the next-level bitmap written are distinct
  buffers, and the next bitmap equals the next level exactly;
* per-unit partials are written by their own unit only;

and that with the arcs of each unit taken in random orders (the LDS
atomic order) the dependencies match the oracle to FP64 rounding."""
import numpy as np
import pytest

VB = 12            # query-layout destinations per block (kQueryVB)


def layout(in_off, in_nbr, n, unit_cap):
    """In-arcs grouped by destination block, ordered by source id inside a
    block, cut into units of at most unit_cap arcs: a list of units, each
    (block, src array, dst array)."""
    deg = np.diff(in_off.astype(np.int64))
    dst = np.repeat(np.arange(n, dtype=np.int64), deg)
    src = in_nbr.astype(np.int64)
    order = np.lexsort((src, dst >> VB))
    src, dst = src[order], dst[order]
    units = []
    blocks = dst >> VB
    for b in np.unique(blocks):
        sel = np.nonzero(blocks == b)[0]
        for lo in range(0, len(sel), unit_cap):
            idx = sel[lo:lo + unit_cap]
            units.append((int(b), src[idx], dst[idx]))
    return units


def dense_launch(units, n, active, visited, val, out, fwd, rng, next_bits_buf, active_buf):
    """One k_bc_dense (+ k_bc_merge) launch; returns the joining vertices
    (forward).  Asserts the cross-workgroup read/write sets are disjoint."""
    assert next_bits_buf is not active_buf
    per_block = {}
    for u, (b, s, d) in enumerate(units):
        per_block.setdefault(b, []).append(u)
    reads = np.zeros(n, bool)        # values any workgroup reads
    writes = np.zeros(n, bool)       # destinations any workgroup (or the merge) writes
    partial_owner = {}
    sums = {}
    for u, (b, s, d) in enumerate(units):
        keep = (~visited[d]) & active[s]
        s_k, d_k = s[keep], d[keep]
        reads[s_k] = True
        acc = np.zeros(1 << VB)
        for i in rng.permutation(len(s_k)):          # LDS atomic order
            acc[d_k[i] - (b << VB)] += val[s_k[i]]
        sums[u] = acc
        if len(per_block[b]) > 1:
            partial_owner[u] = u                     # part[u << VB ...] written by u alone
    assert not (reads & ~visited).any(), "a frontier source is not visited"
    joined = []
    new_out = out.copy()
    for b, us in per_block.items():
        x = sums[us[0]].copy()
        for u in us[1:]:                             # k_bc_merge: unit order
            x += sums[u]
        for l in range(1 << VB):
            dd = (b << VB) + l
            if dd >= n or visited[dd] or x[l] == 0.0:
                continue
            writes[dd] = True
            if fwd:
                new_out[dd] = x[l]
                joined.append(dd)
            else:
                new_out[dd] += x[l]
    assert not (reads & writes).any(), "a value read is a destination written in the same launch"
    assert set(partial_owner) == {u for us in per_block.values() if len(us) > 1 for u in us}
    out[:] = new_out
    return np.array(sorted(joined), np.int64)


def bc_model(off, nbr, in_off, in_nbr, sources, unit_cap, rng):
    """bc_push's loop with every level dense (threshold 0)."""
    n = len(off) - 1
    fwd_units = layout(in_off, in_nbr, n, unit_cap)
    bwd_units = layout(off, nbr, n, unit_cap)      # in-arcs of the transpose = out-arcs
    deps = np.zeros(n)
    transposed = False
    for src in sources:
        f_units, b_units = (fwd_units, bwd_units) if not transposed else (bwd_units, fwd_units)
        np_ = np.zeros(n)
        np_[src] = 1.0
        visited = np.zeros(n, bool)
        visited[src] = True
        levels = [np.array([src])]
        bits = [np.zeros(n, bool)]
        bits[0][src] = True
        while len(levels[-1]):
            nxt_bits = np.zeros(n, bool)
            joined = dense_launch(f_units, n, bits[-1], visited, np_, np_, True, rng, nxt_bits, bits[-1])
            nxt_bits[joined] = True
            visited[joined] = True                  # k_mark
            levels.append(joined)
            bits.append(nxt_bits)
        rounds = len(levels) - 1
        with np.errstate(divide="ignore"):
            inv = 1.0 / np_
        visited[:] = False
        lv = levels[rounds - 1]
        visited[lv] = True
        deps[lv] += inv[lv]
        for r in range(rounds - 2, -1, -1):
            dense_launch(b_units, n, bits[r + 1], visited, deps, deps, False, rng, None, bits[r + 1])
            lv = levels[r]
            visited[lv] = True
            deps[lv] += inv[lv]
        with np.errstate(invalid="ignore"):
            deps = (deps - inv) / inv
        transposed = not transposed
    return deps


@pytest.mark.parametrize("gname,sym", [("rmat10s", True), ("rmat10d", False)])
@pytest.mark.parametrize("unit_cap", [1 << 19, 512])
def test_dense_bc_launches_have_disjoint_read_and_write_sets(rg, orc, gname, sym, unit_cap):
    hg = rg.HostGraph.generate(10, 16, symmetric=sym)
    off, nbr = hg.csr()
    in_off, in_nbr = hg.csr(in_edges=True)
    srcs = [int(s) for s in hg.pick_bc_sources(4)]
    rng = np.random.RandomState(3)
    got = bc_model(off, nbr, in_off, in_nbr, srcs, unit_cap, rng)
    ref, _ = orc.bc(off, nbr, in_off, in_nbr, srcs)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    f = np.isfinite(ref)
    assert np.max(np.abs(got[f] - ref[f])) <= 1e-12 * np.max(np.abs(ref[f]))
