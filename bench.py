"""Benchmark: GTEPS of push-only vs RADAR PageRank and BFS, batched BFS
queries/s and the serving loop's p50/p99, on one MI355X.

Graphs come from the GAP R-MAT generator (a/b/c = 0.57/0.19/0.19, seed
27491095, edge factor 16, self-loops and duplicates removed; int32 ids,
uint64 offsets, FP64 ranks).

* Headline (symmetric R-MAT scale 26, n = 67M, m = 2.1B arcs): one "step"
  is one PageRank run of 20 iterations with RADAR -- the degree-sorted
  graph on the engine the autotuner picks (the blocked engine, whose hub
  blocks are split over private LDS copies); value = 20 m / step time.
  Beside it on the same graph: push-only PageRank, BFS from 16 GAP-picked
  sources (push-only vs autotuned RADAR), batched BFS (256 sources as one
  bit-parallel batch vs one query at a time) and the CPU oracle baseline.
* Variants table (directed R-MAT scale 24): PageRank and BFS for push-only,
  HUBDUP, DegreeSorting, RADAR, pull and the blocked engine.
* Radii (64 starts) and BC (one source) on symmetric R-MAT scale 25 and an
  Erdos-Renyi control of the same size: push-only, HUBDUP, RADAR.
* Triangle counting on the degree-sorted symmetric R-MAT scale 22.
* Serving (symmetric R-MAT scale 22, degree-sorted): BFS-only and 80/15/5
  BFS/BC/PageRank-10 traces, saturated and as Poisson arrivals at 80% of
  the one-at-a-time capacity, batched (continuous batching) vs one at a
  time.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale S]
Prints one JSON line.
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "radar-graph_amd"), os.path.join(ROOT, "oracle")]

PEAK_HBM_GBS = 8000.0   # MI355X HBM3E peak, GB/s
ENGINES = {0: "push", 1: "pull", 2: "hubdup", 3: "radar", 4: "blocked"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def median_pr(g, it, o, reps=3):
    g.pagerank(it, o, want_ranks=False)
    return sorted(g.pagerank(it, o, want_ranks=False).ms for _ in range(reps))[reps // 2]


def bfs_gteps(g, srcs, o):
    """GTEPS over the sources (scanned arcs m_scan / time) and the fraction
    of the HBM roofline at BFS's algorithmic bytes, 8 m_scan + 24 n_reach
    per source (4 B column + 4 B state per scanned arc; offset, parent and
    queue entry per reached vertex)."""
    g.bfs(int(srcs[0]), o, want_parents=False)
    te, tr, tm = 0, 0, 0.0
    for s in srcs:
        r = g.bfs(int(s), o, want_parents=False)
        te += r.edges
        tr += r.reached
        tm += r.ms
    gbs = (8 * te + 24 * tr) / (tm * 1e-3) / 1e9
    return {"gteps": round(te / tm / 1e6, 2), "ms_per_source": round(tm / len(srcs), 3),
            "alg_gbs": round(gbs, 1), "roofline_frac": round(gbs / PEAK_HBM_GBS, 4)}


def host_cpu():
    """The host CPU the baseline runs on: the cores this process may use
    (sched_getaffinity), the OpenMP thread cap if one is set (the GPU host
    sets OMP_NUM_THREADS to its CPU share), nproc and the lscpu model."""
    allowed = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    used = min(allowed, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else allowed
    model, sockets, per_core = None, None, None
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                model = v.strip()
            elif k.strip() == "Socket(s)":
                sockets = v.strip()
            elif k.strip() == "Thread(s) per core":
                per_core = v.strip()
    except Exception:
        pass
    return {"threads_used": used, "affinity_cpus": allowed, "nproc": os.cpu_count(),
            "omp_num_threads": cap, "model": model, "sockets": sockets, "threads_per_core": per_core}


def bc_bytes(summary):
    """BC's algorithmic bytes for one source: forward and backward 24 B per
    scanned arc each (4 B column + 4 B level + 16 B FP64 read-modify-write)
    plus 28 B per reached vertex."""
    return 48 * summary.edges + 28 * summary.reached


def variants_table(rg, args):
    """Every variant on directed R-MAT (the study's second configuration)."""
    t0 = time.time()
    hg = rg.HostGraph.generate(args.variants_scale, 16, symmetric=False)
    n, m, _ = hg.info()
    ctx = rg.Context(0)
    dg = rg.DeviceGraph(ctx, hg)
    sg, ids = dg.degree_sort()
    log(f"variants graph: rmat{args.variants_scale} directed n={n} m={m} ({time.time() - t0:.1f}s)")
    it = args.iters
    out = {"graph": f"rmat{args.variants_scale} directed", "n": n, "m": m, "pagerank": {}, "bfs": {}}
    # The blocked engine on the original ids: the better of one and four
    # arcs in flight per lane (one wins on this unsorted graph, four on the
    # degree-sorted ones; the autotuner makes that choice per graph).
    ms_u = {u: median_pr(dg, it, rg.opts("blocked", unroll=u), reps=1) for u in (1, 4)}
    best_u = min(ms_u, key=ms_u.get)
    pr_runs = [("push", dg, rg.opts("push")), ("hubdup", dg, rg.opts("hubdup", hubs=args.hubs)),
               ("degree_sorting", sg, rg.opts("push")), ("radar_atomic", sg, rg.opts("radar", hubs=args.hubs)),
               ("pull", dg, rg.opts("pull")), ("blocked_unsorted", dg, rg.opts("blocked", unroll=best_u))]
    tuned, _ = sg.autotune("pagerank", base=rg.opts("radar", hubs=args.hubs))
    pr_runs.append(("radar_tuned", sg, tuned))
    for name, g, o in pr_runs:
        ms = median_pr(g, it, o)
        out["pagerank"][name] = {"ms": round(ms, 3), "gteps": round(it * m / ms / 1e6, 2)}
        if name == "blocked_unsorted":
            out["pagerank"][name]["unroll"] = best_u
        log(f"  pagerank {name:16s} {ms:8.2f} ms {it * m / ms / 1e6:8.2f} GTEPS")
    # Hub count k swept from the LDS budget to the L2 budget: LDS copies of
    # k hubs per workgroup alone, then 16K in LDS plus k2 per-XCD copies.
    sweep = []
    for variant, g in (("hubdup", dg), ("radar", sg)):
        for k, k2 in ((1024, 1), (4096, 1), (16384, 1), (16384, 65536), (16384, 262144), (16384, 1048576)):
            ms = median_pr(g, it, rg.opts(variant, hubs=k, l2_hubs=k2), reps=1)
            sweep.append({"variant": variant, "lds_hubs": k, "l2_hubs": max(k, k2), "ms": round(ms, 3),
                          "gteps": round(it * m / ms / 1e6, 2)})
            log(f"  pagerank {variant} k={k} k2={max(k, k2)}: {it * m / ms / 1e6:8.2f} GTEPS")
    out["hub_sweep"] = sweep
    srcs = hg.pick_sources(args.variants_bfs_sources)
    bt, _ = sg.autotune("bfs", sources=ids[srcs[:4]], base=rg.opts("push"))
    # hubdup_full_lds: the hub copy sized to the whole LDS budget and used on
    # every launch (one or two workgroups per CU), as before the default copy
    # left room for four workgroups and small launches went without copies.
    full = dict(hubs=1 << 30, hub_min_arcs=-1)
    bfs_runs = [("push", dg, srcs, rg.opts("push")), ("hubdup", dg, srcs, rg.opts("hubdup", hubs=args.hubs)),
                ("hubdup_full_lds", dg, srcs, rg.opts("hubdup", **full)),
                ("degree_sorting", sg, ids[srcs], rg.opts("push")),
                ("radar", sg, ids[srcs], rg.opts("radar", hubs=args.hubs)),
                ("radar_full_lds", sg, ids[srcs], rg.opts("radar", **full)),
                ("push_pull", dg, srcs, rg.opts("pull")), ("radar_tuned", sg, ids[srcs], bt)]
    for name, g, ss, o in bfs_runs:
        out["bfs"][name] = bfs_gteps(g, ss, o)
        log(f"  bfs {name:16s} {out['bfs'][name]['ms_per_source']:8.3f} ms/src "
            f"{out['bfs'][name]['gteps']:8.2f} GTEPS")
    out["bfs"]["radar_tuned"]["engine"] = ENGINES[bt.variant]
    for g in (sg, dg):
        g.close()
    ctx.close()
    return out


def radii_bc_study(rg, args):
    """Radii (64 hashInt starts) and BC (one GAP-picked source) on symmetric
    R-MAT and on an Erdos-Renyi control with the same n and about the same
    m (uniform degrees: hub copies should gain nothing there)."""
    out = {}
    rm = None
    for kind in ("rmat", "er"):
        t0 = time.time()
        if kind == "rmat":
            hg = rg.HostGraph.generate(args.radii_scale, 16, symmetric=True)
        else:
            deg = max(1, int(round(rm / 2 / (1 << args.radii_scale))))
            hg = rg.HostGraph.generate(args.radii_scale, deg, uniform=True, symmetric=True)
        n, m, _ = hg.info()
        rm = rm or m
        ctx = rg.Context(0)
        dg = rg.DeviceGraph(ctx, hg)
        sg, ids = dg.degree_sort()
        log(f"{kind}{args.radii_scale} symmetric n={n} m={m} ({time.time() - t0:.1f}s)")
        starts = rg.radii_starts(n)
        row = {"n": n, "m": m, "radii_ms": {}, "bc_ms": {}}
        full = dict(hubs=1 << 30, hub_min_arcs=-1)   # as in the variants table
        for name, g, st, o in [("push", dg, starts, rg.opts("push")),
                               ("hubdup", dg, starts, rg.opts("hubdup", hubs=args.hubs)),
                               ("hubdup_full_lds", dg, starts, rg.opts("hubdup", **full)),
                               ("radar", sg, ids[starts], rg.opts("radar", hubs=args.hubs))]:
            g.radii(st, o)
            _, rounds, ms = g.radii(st, o)
            row["radii_ms"][name] = round(ms, 3)
            row["radii_rounds"] = rounds
        src = hg.pick_bc_sources(1)
        for name, g, ss, o in [("push", dg, src, rg.opts("push")),
                               ("hubdup", dg, src, rg.opts("hubdup", hubs=args.hubs)),
                               ("hubdup_full_lds", dg, src, rg.opts("hubdup", **full)),
                               ("radar", sg, ids[src], rg.opts("radar", hubs=args.hubs))]:
            g.bc(ss, o)
            _, _, ms = g.bc(ss, o)
            row["bc_ms"][name] = round(ms, 3)
            if name == "push":
                one, _ = dg.bfs_query(int(ss[0]))
                row["bc_roofline_frac"] = {}
            row["bc_roofline_frac"][name] = round(bc_bytes(one) / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
        # 64 BC queries: one batch (rg_bc_batch) vs one single-source RADAR
        # run per source on the same degree-sorted graph.
        src64 = ids[hg.pick_bc_sources(64)]
        o = rg.opts("radar", hubs=args.hubs)
        sg.bc_batch(src64, o)
        _, _, bms = sg.bc_batch(src64, o)
        single, eq = 0.0, 0
        for s in src64:
            single += sg.bc([s], o)[2]
            eq += bc_bytes(sg.bfs_query(int(s))[0])
        # Equivalent bandwidth: the single-source algorithmic bytes of the 64
        # queries over the batch time (the batch shares its arc scans, so
        # this can exceed the HBM peak).
        row["bc_64_queries"] = {"batch_ms": round(bms, 3), "single_ms": round(single, 3),
                                "speedup": round(single / bms, 2),
                                "equivalent_gbs": round(eq / (bms * 1e-3) / 1e9, 1)}
        log(f"  radii ms {row['radii_ms']}  bc ms {row['bc_ms']}  64 BC queries: batch {bms:.1f} ms vs "
            f"{single:.1f} ms one at a time")
        if kind == "rmat":
            # Uneven source counts (1..256): rg_bc_batch runs them as batches
            # of 64 with the last one partial; one at a time = the mean
            # single-source time above times the count.
            allsrc = ids[hg.pick_bc_sources(256)]
            per_single = single / len(src64)
            row["bc_uneven_batches"] = {}
            for k in (1, 37, 100, 256):
                sg.bc_batch(allsrc[:k], o)
                _, _, kms = sg.bc_batch(allsrc[:k], o)
                row["bc_uneven_batches"][str(k)] = {"ms": round(kms, 2), "queries_per_s": round(k / kms * 1e3, 1),
                                                    "vs_one_at_a_time": round(per_single * k / kms, 2)}
            log(f"  BC uneven batches: {row['bc_uneven_batches']}")
        out[f"{kind}{args.radii_scale}"] = row
        for x in (sg, dg):
            x.close()
        hg.close()
        ctx.close()
    return out


def tc_study(rg, args):
    """Triangle counting on the degree-sorted symmetric R-MAT (oriented by
    id after the relabel: low degree -> high degree), reorder time reported
    apart from count time."""
    hg = rg.HostGraph.generate(args.tc_scale, 16, symmetric=True)
    n, m, _ = hg.info()
    ctx = rg.Context(0)
    dg = rg.DeviceGraph(ctx, hg)
    t0 = time.time()
    sg, _ = dg.degree_sort(sort_neighbors=True)
    ctx.sync()
    reorder = (time.time() - t0) * 1e3
    sg.tc()
    total, ms = sg.tc()
    # TC is latency-bound (list probes), not a stream: it is reported by its
    # compulsory bytes (offsets and the oriented lists read once, 8 (n + 1)
    # + 4 m+) and by the set-intersection probes its kernels make.  The
    # totals kernels split at the first vertex of degree below 512 (at most
    # 2^19): an arc (u, v) with v below the split tests the neighbours of u
    # below v against an LDS bitmap of N+(v) (its position of v in N+(u)
    # probes), one above it binary-searches each w of N+(v) in N+(u)
    # (d+(v) probes).  The merge model of SURVEY's TC row (8 (n + 1) + 4 m+ +
    # 4 sum of d+(u) + d+(v) over oriented arcs) is kept beside them.
    off, nbr = sg.download()
    deg = np.diff(off.astype(np.int64))
    u = np.repeat(np.arange(n, dtype=np.uint32), deg)
    low = nbr < u
    dplus = np.bincount(u[low], minlength=n).astype(np.int64)
    mplus = int(low.sum())
    scan = int(dplus[u[low]].sum() + dplus[nbr[low]].sum())
    below = np.nonzero(deg < 512)[0]
    split = int(min(below[0] if len(below) else n, 1 << 19))
    pos = (np.arange(len(nbr), dtype=np.int64) - off[:-1].astype(np.int64)[u])[low]   # v's index in N(u)
    vv = nbr[low].astype(np.int64)
    probes = int(np.where(vv < split, pos, dplus[vv]).sum())
    del off, nbr, u, low, pos, vv
    merge_bytes = 8 * (n + 1) + 4 * mplus + 4 * scan
    comp_bytes = 8 * (n + 1) + 4 * mplus
    out = {"graph": f"rmat{args.tc_scale} symmetric, degree-sorted", "n": n, "m": m, "triangles": total,
           "count_ms": round(ms, 3), "reorder_ms": round(reorder, 2),
           "compulsory_bytes": comp_bytes,
           "compulsory_gbs": round(comp_bytes / (ms * 1e-3) / 1e9, 1),
           "compulsory_frac": round(comp_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
           "probes": probes, "giga_probes_per_s": round(probes / (ms * 1e-3) / 1e9, 2), "bitmap_split": split,
           "merge_model_bytes": merge_bytes,
           "merge_model_frac": round(merge_bytes / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    log(f"tc rmat{args.tc_scale}: {total} triangles in {ms:.1f} ms (reorder {reorder:.1f} ms)")
    for x in (sg, dg):
        x.close()
    hg.close()
    ctx.close()
    return out


def serve_bench(rg, args, resident=None):
    """Serving over a resident degree-sorted symmetric R-MAT graph: seeded
    query traces (BFS only; BC only; 80/15/5 BFS/BC/PageRank-10) run saturated (all
    queries at t = 0) and as Poisson arrivals at 80% of the one-at-a-time
    capacity, batched (continuous batching, 2 slots, admission by device
    memory, PageRank queries run alone) vs one at a time.
    resident = (host graph, degree-sorted device graph, new ids, scale) of a
    graph already on the device (the headline graph), else one of scale
    --serve-scale is built."""
    t0 = time.time()
    if resident:
        hg, g, ids, scale = resident
        ctx = None
        bc_count = 64
    else:
        scale = args.serve_scale
        hg = rg.HostGraph.generate(scale, 16, symmetric=True)
        ctx = rg.Context(0)
        g0 = rg.DeviceGraph(ctx, hg)
        g, ids = g0.degree_sort()
        g0.close()
        bc_count = 128
    n, m, _ = hg.info()
    log(f"serving graph: rmat{scale} symmetric n={n} m={m} ({time.time() - t0:.1f}s)")

    def trace(count, rate, mix, seed):
        q = rg.poisson_workload(hg, count, rate, mix=mix, seed=seed)
        for x in q:
            x.source = int(ids[x.source])
        return q

    def row(st):
        return {"queries_per_s": round(st.qps, 1), "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3),
                "mean_batch": round(st.mean_batch, 1), "mean_bc_batch": round(st.mean_bc_batch, 1),
                "slots": st.slots,
                "peak_gb": round(st.peak_bytes / 1e9, 2), "hbm_budget_gb": round(st.hbm_budget / 1e9, 1),
                "mem_waits": st.mem_waits}

    out = {"graph": f"rmat{scale} symmetric, degree-sorted", "n": n, "m": m}
    # (name, mix, saturated count, Poisson count): the Poisson traces are
    # long enough that their p99 rests on about ten queries.
    for name, mix, count, pcount in [("bfs", (1, 0, 0), args.serve_queries, max(args.serve_queries, 1024)),
                                     ("bc", (0, 1, 0), bc_count, 256),
                                     ("mix_80_15_5", (0.8, 0.15, 0.05), 256, 1024)]:
        q = trace(count, 0.0, mix, 27491095)
        _, one = g.serve(q, rg.serve_config(one_at_a_time=True))
        _, bat = g.serve(q, rg.serve_config(slots=2))
        rate = 0.8 * one.qps
        q2 = trace(pcount, rate, mix, 1)
        _, one_p = g.serve(q2, rg.serve_config(one_at_a_time=True))
        _, bat_p = g.serve(q2, rg.serve_config(slots=2))
        out[name] = {"saturated": {"queries": count, "one_at_a_time": row(one), "batched": row(bat),
                                   "speedup": round(bat.qps / one.qps, 2)},
                     f"poisson_{rate:.0f}qps": {"queries": pcount, "one_at_a_time": row(one_p),
                                                "batched": row(bat_p)}}
        log(f"serving rmat{scale} {name}: saturated {bat.qps:.0f} vs {one.qps:.0f} q/s one at a time; at "
            f"{rate:.0f} q/s p50/p99 {bat_p.p50_ms:.2f}/{bat_p.p99_ms:.2f} ms vs "
            f"{one_p.p50_ms:.2f}/{one_p.p99_ms:.2f} ms")
    if not resident:
        g.close()
        ctx.close()
    return out


def profile_run(rg, args):
    """One app on one engine, repeated, for counter collection under
    rocprofv3: --profile app:variant:scale[d|s][:key=value,...] (app bfs,
    query (BFS summary only), batch (--batch sources, degree-sorted graph),
    bcbatch (--batch BC queries, degree-sorted graph), pagerank, prd
    (PageRankDelta), radii or bc; variant push, hubdup, radar, blocked or pull; d =
    directed, s = symmetric; radar, blocked, batch and query run on the
    degree-sorted graph (sorted=1 / sorted=0 forces the degree-sorted / original
    graph for any variant); graphs=0 turns HIP-graph replay off; other options
    are rg.opts keywords, e.g. hubs=1073741824, hub_min_arcs=-1)."""
    parts = args.profile.split(":")
    app, variant, gspec = parts[0], parts[1], parts[2]
    kw = {}
    if len(parts) > 3 and parts[3]:
        for item in parts[3].split(","):
            k, v = item.split("=")
            kw[k] = int(v)
    scale, sym = int(gspec[:-1]), gspec[-1] == "s"
    graphs_on = kw.pop("graphs", 1)   # graphs=0: launch every kernel (no HIP-graph replay)
    sort = kw.pop("sorted", -1)       # 1 / 0: run on the degree-sorted / original graph whatever the variant
    hg = rg.HostGraph.generate(scale, 16, symmetric=sym)
    ctx = rg.Context(0)
    ctx.graphs(bool(graphs_on))
    dg = rg.DeviceGraph(ctx, hg)
    ids = None
    if sort == 1 or (sort == -1 and (variant in ("radar", "blocked") or app in ("batch", "query", "bcbatch"))):
        g, ids = dg.degree_sort(sort_neighbors=True)
    else:
        g = dg
    o = rg.opts(variant, **kw)
    srcs = hg.pick_sources(max(args.profile_reps, args.batch))
    if ids is not None:
        srcs = ids[srcs]

    def run(i):
        if app == "bfs":
            return g.bfs(int(srcs[i]), o, want_parents=False).ms
        if app == "query":
            return g.bfs_query(int(srcs[i]), o)[1]
        if app == "batch":   # --batch sources as one bit-parallel batch (the variant is unused)
            return g.bfs_batch(srcs[:args.batch], o)[2]
        if app == "bcbatch":   # --batch BC queries (rg_bc_batch) with the variant's single-source engine
            return g.bc_batch(srcs[:args.batch], o)[2]
        if app == "pagerank":
            return g.pagerank(args.iters, o, want_ranks=False).ms
        if app == "prd":
            return g.pagerank_delta(args.iters, o, want_ranks=False).ms
        if app == "radii":
            st = rg.radii_starts(g.n)
            return g.radii(st if ids is None else ids[st], o)[2]
        return g.bc(srcs[i:i + 1], o)[2]

    run(0)   # set-up outside the timed runs (hub metadata, layouts)
    ms = sum(run(i) for i in range(args.profile_reps)) / args.profile_reps
    ctx.sync()
    log(f"profile {args.profile}: {ms:.3f} ms per run (device time)")
    return {"profile": args.profile, "ms_per_run": round(ms, 3), "reps": args.profile_reps}


def graph_cache_trip(rg, ctx, sg, new_ids, args):
    """The restart path: the degree-sorted graph saved as GAP .sg with its
    new_ids sidecar (rg_graph_save), then loaded back (rg_graph_load, mmap ->
    pinned staging -> device) in place of generating, building and sorting
    it.  The reload is checked (sizes, new_ids, one PageRank's iteration
    count and L1 equal to the original graph's).  Skipped when the cache
    directory lacks room.  The file was just written, so the load reads it
    from the page cache: a cold restart adds the disk's read time."""
    d = args.graph_cache or os.path.join(tempfile.gettempdir(), "rg_graph_cache")
    os.makedirs(d, exist_ok=True)
    n, m = sg.n, sg.m
    nbytes = 17 + (n + 1) * 8 + m * 4
    if shutil.disk_usage(d).free < nbytes + n * 4 + (2 << 30):
        log(f"graph cache: skipped, {d} has {shutil.disk_usage(d).free / 1e9:.1f} GB free")
        return {"skipped": "not enough free disk"}
    path = os.path.join(d, f"rmat{args.scale}s_degree_sorted.sg")
    try:
        t0 = time.time()
        sg.save(path, rg.FMT_SG, new_ids=new_ids)
        t_save = time.time() - t0
        t0 = time.time()
        lg, lids = rg.DeviceGraph.load(ctx, path, rg.FMT_SG)
        ctx.sync()
        t_load = time.time() - t0
        o = rg.opts("blocked")
        a, b = sg.pagerank(3, o, want_ranks=False), lg.pagerank(3, o, want_ranks=False)
        ok = (lids is not None and np.array_equal(lids, new_ids) and (lg.n, lg.m) == (n, m)
              and a.iters == b.iters and abs(a.l1 - b.l1) <= 1e-10)
        lg.close()
    finally:
        for f in (path, path + ".newids"):
            if os.path.exists(f):
                os.remove(f)
    log(f"graph cache: save {t_save * 1e3:.0f} ms, load {t_load * 1e3:.0f} ms "
        f"({nbytes / t_load / 1e9:.1f} GB/s), reload equal: {ok}")
    return {"file_bytes": nbytes, "save_ms": round(t_save * 1e3, 1), "load_ms": round(t_load * 1e3, 1),
            "load_gbs": round(nbytes / t_load / 1e9, 2), "reload_equal": bool(ok), "format": "GAP .sg + new_ids"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=26, help="R-MAT scale of the headline graph (symmetric)")
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--bfs-sources", type=int, default=16)
    ap.add_argument("--batch", type=int, default=256, help="sources per batched BFS")
    ap.add_argument("--hubs", type=int, default=0, help="LDS hub slots (0 = LDS budget)")
    ap.add_argument("--variants-scale", type=int, default=24, help="directed R-MAT of the variants table (0: skip)")
    ap.add_argument("--variants-bfs-sources", type=int, default=64)
    ap.add_argument("--radii-scale", type=int, default=25, help="Radii/BC study: R-MAT and ER scale (0: skip)")
    ap.add_argument("--tc-scale", type=int, default=22, help="triangle counting R-MAT scale (0: skip)")
    ap.add_argument("--serve-scale", type=int, default=22, help="R-MAT scale of the serving graph (symmetric)")
    ap.add_argument("--serve-queries", type=int, default=512)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-resident", action="store_true",
                    help="also serve the resident headline graph (off by default: its mixed-query Poisson trace, "
                         "batched on two slots, has ended in a GPU memory fault in two of three runs; cause not found)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--sections", default="headline,variants,radii,tc,serve",
                    help="comma list of sections to run (headline,variants,radii,tc,serve)")
    ap.add_argument("--profile", default="", help="section 'profile': app:variant:scale[d|s][:opts] (see profile_run)")
    ap.add_argument("--profile-reps", type=int, default=8)
    ap.add_argument("--graph-cache", default="",
                    help="directory for the save/reload of the degree-sorted headline graph "
                         "(default: <tmp>/rg_graph_cache; 'off' skips it)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run the library's plain 1 GiB stream copy once (k_stream_copy: the "
                         "known-size kernel that scales FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design, not measured here", file=sys.stderr)
        sys.exit(2)

    import radar_graph as rg

    sections = set(args.sections.split(","))
    if "headline" not in sections:
        # Sections on their own (e.g. for counter collection): no headline line.
        out = {}
        if "profile" in sections and args.profile:
            out["profile"] = profile_run(rg, args)
        if "variants" in sections and args.variants_scale > 0:
            out["variants"] = variants_table(rg, args)
        if "radii" in sections and args.radii_scale > 0:
            out["radii_bc"] = radii_bc_study(rg, args)
        if "tc" in sections and args.tc_scale > 0:
            out["tc"] = tc_study(rg, args)
        if "serve" in sections and not args.no_serve:
            out["serving"] = {f"rmat{args.serve_scale}": serve_bench(rg, args)}
        print(json.dumps(out), flush=True)
        return

    t0 = time.time()
    hg = rg.HostGraph.generate(args.scale, 16, uniform=False, symmetric=True)
    n, m, _ = hg.info()
    t_gen = time.time() - t0
    log(f"graph: rmat{args.scale} symmetric n={n} m={m} (host build {t_gen:.1f}s)")
    ctx = rg.Context(0)
    if args.calibrate_copy:
        gb, _ = ctx.copy_bandwidth(1 << 30, reps=1)
        log(f"calibration copy (k_stream_copy): 1 GiB read + 1 GiB write, {gb:.0f} GB/s")
        gr, _, gms = ctx.read_bandwidth(1 << 32, reps=1, gathers=1 << 27)
        log(f"calibration read (k_stream_read): 4 GiB read once, {gr:.0f} GB/s; "
            f"k_gather_read: 2^27 random 8-byte loads over it, {gms:.3f} ms ({(1 << 27) / gms / 1e6:.2f} G loads/s)")
    stream_gbs, _ = ctx.copy_bandwidth(1 << 30, reps=5)
    log(f"measured stream copy: {stream_gbs:.0f} GB/s (read + write)")
    dg = rg.DeviceGraph(ctx, hg)
    t0 = time.time()
    # Strict DegreeSorting (neighbour lists sorted by new id, hubs first):
    # a bottom-up BFS level finds a frontier parent in fewer probes.
    sg, new_ids = dg.degree_sort(by_outdeg=False, sort_neighbors=True)
    ctx.sync()
    t_sort = time.time() - t0
    log(f"degree sort (GPU, sorted lists): {t_sort * 1e3:.1f} ms")
    cache = graph_cache_trip(rg, ctx, sg, new_ids, args) if args.graph_cache != "off" and not args.quick else None
    it = args.iters

    # RADAR PageRank on the engine the autotuner picks for this graph.
    t0 = time.time()
    tuned, _ = sg.autotune("pagerank", base=rg.opts("radar", hubs=args.hubs))
    t_tune = time.time() - t0
    engine = ENGINES[tuned.variant]
    log(f"autotune pagerank: engine={engine} wg={tuned.wg_size} unroll={tuned.unroll} ({t_tune:.1f}s)")
    g, o = sg, tuned
    for _ in range(args.warmup):
        g.pagerank(it, o, want_ranks=False)
    ctx.sync()
    ctx.timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.pagerank(it, o, want_ranks=False)
    ctx.sync()
    wall = time.perf_counter() - t0
    launches, kms = ctx.kernel_time()
    ctx.timing(False)
    ms_step = wall * 1e3 / args.steps
    value = args.steps * it * m / wall / 1e9
    # Algorithmic bytes of one timed launch: the blocked engine runs a whole
    # iteration in it (k_blk_pr_qx, plus k_blk_pr_merge for split hub
    # blocks, timed together), the per-iteration 20 m + 48 n of the push
    # model: per arc 4 B column + 16 B accumulator read-modify-write, per
    # vertex 8 B offset + 8 B contribution + 8 B accumulator read + 8 B old
    # rank + 8 B new rank + 8 B accumulator reset.  The atomic engine's
    # timed launch is the edge phase only: 20 m + 16 n.
    bytes_launch = 20 * m + (48 if engine == "blocked" else 16) * n
    avg_launch_ms = kms / max(launches, 1)
    achieved = bytes_launch / (avg_launch_ms * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_pagerank_radar_rmat{args.scale}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    # The blocked layout's own compulsory bytes per iteration: per arc its
    # streamed bytes -- 4 B when the library packs the arcs (the low source
    # bits and the 14-bit local destination in one word, which it does when
    # ceil(log2 n) - 10 + 14 <= 32), else a 4 B source id + a 2 B local
    # destination -- and 40 B per vertex (offset, contribution read once,
    # old rank, new rank, next contribution); the gathers of contrib[s]
    # beyond one read per vertex are what the L2 does not keep.
    packed = max(0, (n - 1).bit_length() - 10) + 14 <= 32
    compulsory = (4 if packed else 6) * m + 40 * n
    # Whole iterations (edge and vertex phases) at 20 m + 48 n bytes each.
    pr_frac = it * (20 * m + 48 * n) / (ms_step * 1e-3) / 1e9 / PEAK_HBM_GBS
    pr = {"radar_tuned": {"ms": round(ms_step, 3), "gteps": round(value, 2), "engine": engine,
                          "wg_size": tuned.wg_size, "unroll": tuned.unroll,
                          "iteration_roofline_frac": round(pr_frac, 4)}}
    bfs, batched = {}, None
    if not args.quick:
        ms = median_pr(dg, it, rg.opts("push"), reps=1)
        pr["push"] = {"ms": round(ms, 3), "gteps": round(it * m / ms / 1e6, 2)}
        pr["radar_vs_push"] = round(pr["radar_tuned"]["gteps"] / pr["push"]["gteps"], 2)
        # RADAR in the reference's own form: the degree-sorted graph with LDS
        # copies of hubs 0..k-1 merged by atomics (RD PageRank.C:42-70), beside
        # the blocked engine the autotuner picks.
        ms = median_pr(sg, it, rg.opts("radar", hubs=args.hubs), reps=1)
        pr["radar_atomic"] = {"ms": round(ms, 3), "gteps": round(it * m / ms / 1e6, 2), "engine": "radar",
                              "vs_push": round(pr["push"]["ms"] / ms, 2)}
        log(f"pagerank push {pr['push']['ms']:.1f} ms {pr['push']['gteps']} GTEPS; tuned {pr['radar_vs_push']}x; "
            f"atomic RADAR {pr['radar_atomic']['gteps']} GTEPS ({pr['radar_atomic']['vs_push']}x)")
        # PageRankDelta, 20 rounds: push-only on the original ids vs the
        # blocked engine on the degree-sorted graph (dense rounds blocked,
        # sparse rounds pushed from the active list).
        prd = {}
        for name, gg, oo in (("push", dg, rg.opts("push")), ("blocked", sg, rg.opts("blocked"))):
            gg.pagerank_delta(it, oo, want_ranks=False)
            runs = sorted((gg.pagerank_delta(it, oo, want_ranks=False) for _ in range(3)), key=lambda r: r.ms)
            prd[name] = {"ms": round(runs[1].ms, 3), "rounds": runs[1].iters}
        prd["blocked_vs_push"] = round(prd["push"]["ms"] / prd["blocked"]["ms"], 2)
        pr["pagerank_delta"] = prd
        log(f"pagerank_delta push {prd['push']['ms']:.1f} ms, blocked {prd['blocked']['ms']:.1f} ms "
            f"({prd['blocked_vs_push']}x, {prd['blocked']['rounds']} rounds)")
        # BFS from GAP-picked sources: push-only vs autotuned RADAR.
        srcs = hg.pick_sources(args.bfs_sources)
        bt, _ = sg.autotune("bfs", sources=new_ids[srcs[:4]], base=rg.opts("push"))
        bfs["push"] = bfs_gteps(dg, srcs, rg.opts("push"))
        bfs["radar_tuned"] = bfs_gteps(sg, new_ids[srcs], bt)
        bfs["radar_tuned"].update({"engine": ENGINES[bt.variant], "threshold": bt.threshold,
                                   "heavy_degree": bt.heavy_degree})
        bfs["radar_vs_push"] = round(bfs["radar_tuned"]["gteps"] / bfs["push"]["gteps"], 2)
        bfs["roofline_note"] = ("roofline_frac counts the push model's bytes, 8 B per arc of every reached vertex and "
                                "24 B per reached vertex; bottom-up levels stop at an unvisited vertex's first "
                                "frontier in-neighbour and never read most arcs, so the pull engine's figure can "
                                "pass 1 (profiles/k_bfs_pull_light_bfs_rmat26.md has its counter bytes)")
        # RADAR in the reference's own form (RD BFS.C:43-94): degree-sorted
        # ids, LDS copies of hubs 0..k-1, dense levels on the blocked layout;
        # and the push-pull engine alone on the same graph.
        bfs["radar"] = bfs_gteps(sg, new_ids[srcs], rg.opts("radar", hubs=args.hubs))
        bfs["radar"]["vs_push"] = round(bfs["radar"]["gteps"] / bfs["push"]["gteps"], 2)
        bfs["pull_degree_sorted"] = bfs_gteps(sg, new_ids[srcs], rg.opts("pull"))
        log(f"bfs push {bfs['push']['gteps']} GTEPS, radar_tuned ({ENGINES[bt.variant]}) "
            f"{bfs['radar_tuned']['gteps']} GTEPS ({bfs['radar_vs_push']}x), RADAR (hub copies) "
            f"{bfs['radar']['gteps']} GTEPS ({bfs['radar']['vs_push']}x), pull {bfs['pull_degree_sorted']['gteps']}")
        # Batched BFS: 256 sources as one bit-parallel batch vs the same
        # queries one at a time on the best single-source engine.
        bsrc = new_ids[hg.pick_sources(args.batch)]
        sg.bfs_query(int(bsrc[0]), bt)
        t0 = time.perf_counter()
        single = [sg.bfs_query(int(s), bt)[0] for s in bsrc[:64]]
        t_single = (time.perf_counter() - t0) / 64
        reps = 3

        def time_batch(o):
            sg.bfs_batch(bsrc, o)
            t0 = time.perf_counter()
            for _ in range(reps):
                res, _, _ = sg.bfs_batch(bsrc, o)
            return (time.perf_counter() - t0) / reps, res

        # The library's choice of launch form, then each form forced: per-level
        # kernels (HIP-graph chunks of four levels) and one persistent launch.
        t_batch, out = time_batch(rg.opts())
        t_levels, out_levels = time_batch(rg.opts(batch_launch=1))
        t_persist, out_persist = time_batch(rg.opts(batch_launch=2))
        same = (all(a.astuple() == b.astuple() for a, b in zip(out[:64], single))
                and [x.astuple() for x in out] == [x.astuple() for x in out_levels]
                and [x.astuple() for x in out] == [x.astuple() for x in out_persist])
        batched = {"sources": len(bsrc), "batch_ms": round(t_batch * 1e3, 3),
                   "launch": "one persistent launch" if m >= (1 << 27) else "per-level kernels (HIP graph chunks)",
                   "per_level_launches_ms": round(t_levels * 1e3, 3),
                   "persistent_launch_ms": round(t_persist * 1e3, 3),
                   "queries_per_s": round(len(bsrc) / t_batch, 1),
                   "one_at_a_time_queries_per_s": round(1.0 / t_single, 1),
                   "one_at_a_time_engine": ENGINES[bt.variant],
                   "speedup": round(t_single * len(bsrc) / t_batch, 2),
                   "equivalent_gteps": round(sum(x.edges for x in out) / t_batch / 1e9, 1),
                   "answers_equal_single": same}
        log(f"batched bfs: {batched['queries_per_s']} q/s vs {batched['one_at_a_time_queries_per_s']} q/s "
            f"one at a time ({batched['speedup']}x), equal answers: {same}")

    cpu = None
    if not args.no_cpu:
        import pyoracle
        off, nbr = hg.csr()
        host = host_cpu()
        cores = host["threads_used"]
        t0 = time.perf_counter()
        pyoracle.pagerank(off, nbr, 1, nthreads=cores)
        per_iter = time.perf_counter() - t0
        cpu_iters = int(max(2, min(60, round(12.0 / max(per_iter, 1e-3)))))
        t0 = time.perf_counter()
        _, done_iters, _ = pyoracle.pagerank(off, nbr, cpu_iters, nthreads=cores)
        tc = time.perf_counter() - t0
        cpu_iters = min(done_iters, cpu_iters)
        cpu = {"value": round(cpu_iters * m / tc / 1e9, 4), "unit": "GTEPS", "cores": cores, "kind": "port",
               "sample": f"oracle push PageRank (OpenMP, CAS-loop FP64 adds as the reference), "
                         f"{cpu_iters} iterations on the same rmat{args.scale} graph, {tc:.1f} s",
               "host": host}
        log(f"cpu baseline: {cpu['value']} GTEPS on {cores} threads")
        del off, nbr
    dg.close()
    serving = {}
    if not args.quick and "serve" in sections and not args.no_serve and args.serve_resident:
        # Serving over the headline graph itself, resident and reordered (off
        # by default: its mixed-query Poisson trace, batched on two slots, has
        # ended in a GPU memory fault in two of three runs; cause not found).
        serving[f"rmat{args.scale}"] = serve_bench(rg, args, resident=(hg, sg, new_ids, args.scale))
    sg.close()
    hg.close()
    ctx.close()

    variants = None
    if not args.quick and "variants" in sections and args.variants_scale > 0:
        variants = variants_table(rg, args)
    study = None
    if not args.quick and "radii" in sections and args.radii_scale > 0:
        study = radii_bc_study(rg, args)
    tc = None
    if not args.quick and "tc" in sections and args.tc_scale > 0:
        tc = tc_study(rg, args)
    if not args.quick and "serve" in sections and not args.no_serve:
        serving[f"rmat{args.serve_scale}"] = serve_bench(rg, args)

    out = {
        "metric": "GTEPS per app on RMAT-26 + batched BFS queries/s, p50/p99 vs one-at-a-time",
        "value": round(value, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (GAP R-MAT generator, seed 27491095)",
        "config": {"workload": f"PageRank {it} iterations, RADAR (degree-sorted graph, autotuned engine: {engine}), "
                               f"symmetric rmat{args.scale} edge factor 16",
                   "n": n, "m": m, "iterations": it, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "note": ("achieved counts the push model's 20 m + 48 n bytes per iteration (4 B column and a "
                              "16 B accumulator read-modify-write in HBM per arc); the blocked engine accumulates "
                              "in LDS and streams 4-6 B per arc, so it moves fewer bytes than the model and frac "
                              "can pass 1: its own bytes are compulsory_* and traffic") if engine == "blocked" else None,
                     "measured_stream_copy_gbs": round(stream_gbs, 1),
                     "frac_of_measured_copy": round(achieved / stream_gbs, 4),
                     "compulsory_bytes_per_launch": compulsory if engine == "blocked" else None,
                     "compulsory_gbs": round(compulsory / (avg_launch_ms * 1e-3) / 1e9, 1) if engine == "blocked" else None,
                     "compulsory_frac_of_measured_copy": (round(compulsory / (avg_launch_ms * 1e-3) / 1e9 / stream_gbs, 4)
                                                          if engine == "blocked" else None),
                     "traffic_frac_of_measured_copy": (round(traffic / (avg_launch_ms * 1e-3) / 1e9 / stream_gbs, 4)
                                                       if traffic else None),
                     "kernel": "k_blk_pr_qx + k_blk_pr_merge" if engine == "blocked" else "edge_light_kernel<PrOp,SRC_ALL,HUB_PREFIX>",
                     "launches": launches, "avg_launch_ms": round(avg_launch_ms, 4),
                     "bytes_per_launch": bytes_launch},
        "cpu_baseline": cpu,
        "pagerank": pr,
        "bfs": bfs,
        "batched_bfs": batched,
        "variants": variants,
        "radii_bc": study,
        "tc": tc,
        "serving": serving or None,
        "graph_cache": cache,
        "preprocess_ms": {"host_graph_build": round(t_gen * 1e3, 1), "gpu_degree_sort": round(t_sort * 1e3, 2),
                          "autotune": round(t_tune * 1e3, 1)},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
