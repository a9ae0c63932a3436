"""Writes the golden fixtures in tests/golden/ from the reference build.

Runs the reference's own, unmodified programs (built into oracle/_ref by
`make -C oracle ref` from the reference sources) on small seeded graphs and
stores what they produce:

  * GAP generator / builder / SourcePicker output (oracle/ref/gap_graph_dump)
    as SHA-256 digests plus a short prefix;
  * per-vertex arrays of the Ligra apps, captured from the reference's
    Compute with one OpenMP thread (so FP64 results are deterministic), for
    push-new-orig, push-pull and degree-sorting; plus their printed summary
    lines;
  * GAP tc_local triangle totals.

The graphs are produced by the product's generator (radar_graph) and written
in Ligra's AdjacencyGraph format / GAP's .el format for the reference to
read; the generator itself is checked against the reference's by the digests.

Usage: python tests/golden/make_golden.py   (needs /root/reference built into
oracle/_ref; the fixtures are committed, and the GPU tests only read them).
"""
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "radar-graph_amd"), os.path.join(ROOT, "oracle")]
import radar_graph as rg  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def run(cmd, env_extra=None):
    env = dict(os.environ, OMP_NUM_THREADS="1")
    if env_extra:
        env.update(env_extra)
    out = subprocess.run(cmd, env=env, capture_output=True, text=True, check=True).stdout
    return out


def parse(out, key):
    m = re.search(re.escape(key) + r"\s*=\s*([-+0-9.eE]+)", out)
    return m.group(1) if m else None


def capture(binary, args, sizes, n, dtypes, tmp):
    prefix = os.path.join(tmp, "cap")
    out = run([os.path.join(REF, binary)] + args,
              {"RG_DUMP": prefix, "RG_DUMP_SIZES": ",".join(str(s) for s in sizes)})
    arrs = [np.fromfile(f"{prefix}.arr{i}.bin", dtypes[i]) for i in range(len(sizes))]
    ids = np.fromfile(f"{prefix}.ids.bin", np.uint32)
    return out, arrs, ids


def gap_fixtures(tmp):
    fx = {}
    for scale, deg, uni in [(10, 16, 0), (10, 16, 1), (12, 8, 0)]:
        p = os.path.join(tmp, "el.bin")
        subprocess.run([os.path.join(REF, "gap_graph_dump"), "el", str(scale), str(deg), str(uni), p],
                       check=True, capture_output=True)
        el = np.fromfile(p, np.int32).reshape(-1, 2)
        fx[f"el_{scale}_{deg}_{uni}"] = {"sha": sha(el), "head": el[:8].tolist(), "count": len(el)}
        for mode in ("csr", "relabel"):
            subprocess.run([os.path.join(REF, "gap_graph_dump"), mode, str(scale), str(deg), str(uni), p],
                           check=True, capture_output=True)
            raw = open(p, "rb").read()
            n, m = np.frombuffer(raw[:16], np.int64)
            off = np.frombuffer(raw[16:16 + 8 * (n + 1)], np.int64)
            nb = np.frombuffer(raw[16 + 8 * (n + 1):], np.int32)[:m]
            fx[f"{mode}_{scale}_{deg}_{uni}"] = {"sha": sha(off.astype(np.uint64), nb.astype(np.uint32)),
                                                 "n": int(n), "m": int(m)}
        subprocess.run([os.path.join(REF, "gap_graph_dump"), "sources", str(scale), str(deg), str(uni),
                        "64", p], check=True, capture_output=True)
        fx[f"sources_{scale}_{deg}_{uni}"] = np.fromfile(p, np.int32).tolist()
    return fx


def ligra_fixtures(tmp, name, hg, sym):
    """Per-app captures on one graph.  Returns (arrays, summaries)."""
    arrays, summ = {}, {}
    path = os.path.join(tmp, name + ".adj")
    hg.write(path, rg.FMT_ADJ)
    n = hg.n
    flag = ["-s"] if sym else []
    srcs = hg.pick_sources(3)
    arrays["bfs_sources"] = srcs
    for var in ("pno", "pp", "ds"):
        pre = []   # the variant's own default (-preprocess 1 for DS)
        out, (par,), ids = capture(f"cap_{var}_BFS", flag + ["-rounds", "0", "-r", str(int(srcs[0])), path],
                                   [4], n, [np.uint32], tmp)
        arrays[f"{var}_bfs_parents"] = par
        arrays[f"{var}_new_ids"] = ids
        summ[f"{var}_bfs"] = {"iters": int(parse(out, "Num. Iters until convergence")),
                              "reached": int(parse(out, "No. of nodes in BFS-Tree")), "src": int(srcs[0])}
        out, (a0, a1), _ = capture(f"cap_{var}_PageRank", flag + ["-rounds", "0", "-maxiters", "20", path],
                                   [8, 8], n, [np.float64, np.float64], tmp)
        arrays[f"{var}_pr_ranks"] = a0 if a0.sum() > a1.sum() else a1
        summ[f"{var}_pr"] = {"iters": int(parse(out, "Num Iters")), "l1": float(parse(out, "L1_Norm"))}
        if var != "pp":
            # PageRankDelta.C frees p, Delta, nghSum last (153): slot 2 from the end is p.
            out, (_, _, prd), _ = capture(f"cap_{var}_PageRankDelta",
                                          flag + ["-rounds", "0", "-maxiters", "20", path],
                                          [8, 8, 8], n, [np.float64] * 3, tmp)
            arrays[f"{var}_prd_ranks"] = prd
            summ[f"{var}_prd"] = {"iters": int(parse(out, "Num Iters")), "l1": float(parse(out, "L1_Norm"))}
        out, (rad,), _ = capture(f"cap_{var}_Radii", flag + ["-rounds", "0", path], [4], n, [np.int32], tmp)
        arrays[f"{var}_radii"] = rad
        summ[f"{var}_radii"] = {"rounds": int(parse(out, "Iters until convergence"))}
        out, (dep,), _ = capture(f"cap_{var}_BC-manyiter-v2", flag + ["-rounds", "0", "-maxiters", "4", path],
                                 [8], n, [np.float64], tmp)
        arrays[f"{var}_bc_deps"] = dep
        summ[f"{var}_bc"] = {"combined": int(parse(out, "Combined Iters (in front traversal)"))}
        del pre
    return arrays, summ


def tc_fixtures(tmp, name, hg):
    path = os.path.join(tmp, name + ".el")
    hg.write(path, rg.FMT_EL)
    res = {}
    for var in ("pno", "ds"):
        out = run([os.path.join(REF, f"{var}_tc_local"), "-f", path, "-s", "-n", "1", "-a"])
        m = re.search(r"^(\d+) triangles", out, re.M)
        res[var] = int(m.group(1))
    return res


def main():
    if not os.path.exists(os.path.join(REF, "cap_pno_BFS")):
        sys.exit("oracle/_ref is not built: make -C oracle ref")
    with tempfile.TemporaryDirectory() as tmp:
        meta = {"gap": gap_fixtures(tmp), "ligra": {}, "tc": {}}
        for name, scale, sym in [("rmat10s", 10, True), ("rmat10d", 10, False)]:
            hg = rg.HostGraph.generate(scale, 16, symmetric=sym)
            arrays, summ = ligra_fixtures(tmp, name, hg, sym)
            off, nb = hg.csr()
            arrays["csr_sha"] = np.frombuffer(bytes.fromhex(sha(off, nb)), np.uint8)
            np.savez_compressed(os.path.join(HERE, f"ligra_{name}.npz"), **arrays)
            meta["ligra"][name] = summ
        for name, scale in [("rmat10s", 10), ("rmat12s", 12)]:
            hg = rg.HostGraph.generate(scale, 16, symmetric=True)
            meta["tc"][name] = tc_fixtures(tmp, name, hg)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
