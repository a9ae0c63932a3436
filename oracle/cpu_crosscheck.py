"""TEST INFRASTRUCTURE: times the oracle's push PageRank against the
reference's own PageRank build (oracle/_ref/pno_PageRank, compiled from
push-new-orig/Ligra by `make -C oracle ref`) on the same symmetric GAP R-MAT
graph with the same OpenMP thread count, so that the oracle's time -- the
CPU baseline bench.py reports on the GPU host, where the reference cannot
run -- can be read as the reference's.

The reference times its own Compute (PageRank.C:92-93,136-137, "Run
Time(sec)"); `-rounds R` runs Compute R + 1 times after loading
(ligra.h:615-620): the first run is taken as warm-up and the median of the
other R is reported.  Both sides run -maxiters 20 (PageRank.C:97-131) and
stop at the same iteration (their printed counts are compared).

    python oracle/cpu_crosscheck.py --scale 20 --threads 8
"""
import argparse
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "radar-graph_amd"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--rounds", type=int, default=4)
    args = ap.parse_args()
    exe = os.path.join(HERE, "_ref", "pno_PageRank")
    if not os.path.exists(exe):
        sys.exit("oracle/_ref/pno_PageRank is missing: make -C oracle ref")
    import numpy as np
    import pyoracle
    import radar_graph as rg

    hg = rg.HostGraph.generate(args.scale, 16, symmetric=True)
    n, m, _ = hg.info()
    off, nbr = hg.csr()
    env = dict(os.environ, OMP_NUM_THREADS=str(args.threads), OMP_PROC_BIND="close")
    with tempfile.TemporaryDirectory() as d:
        base = os.path.join(d, f"rmat{args.scale}")
        hg.write(base, rg.FMT_ADJ_BINARY)
        cmd = [exe, "-s", "-b", "-maxiters", "20", "-rounds", str(args.rounds), base]
        out = subprocess.run(cmd, env=env, check=True, capture_output=True, text=True).stdout
    runs = [float(l.split(":")[1]) for l in out.splitlines() if l.startswith("Run Time(sec)")]
    ref_s = float(np.median(runs[1:]))
    iters = [int(l.split("=")[1]) for l in out.splitlines() if "Num Iters" in l][-1]
    # The reference prints maxiters + 1 when it stops at the limit (PageRank.C:139).
    edge_passes = min(iters, 20)
    times = []
    for _ in range(3):
        t = time.perf_counter()
        _, it, _ = pyoracle.pagerank(off, nbr, 20, nthreads=args.threads)
        times.append(time.perf_counter() - t)
    orc_s = float(np.median(times))
    print(f"rmat{args.scale} symmetric n={n} m={m}, {args.threads} threads, 20 iterations "
          f"(reference printed {iters}, oracle {it})")
    print(f"reference pno_PageRank Compute: {ref_s:.3f} s = {edge_passes * m / ref_s / 1e9:.3f} GTEPS")
    print(f"oracle orc_pagerank:            {orc_s:.3f} s = {20 * m / orc_s / 1e9:.3f} GTEPS")
    print(f"oracle / reference time: {orc_s / ref_s:.2f}")


if __name__ == "__main__":
    main()
