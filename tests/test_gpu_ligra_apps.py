"""GPU tests of the drop-in apps (radar-graph_amd/bin/*, built on the
Ligra-style C++ API and the engine-level C ABI): the same command lines as
the reference's apps, checked against the reference's own outputs captured
in tests/golden (one-thread runs of push-new-orig and degree-sorting) and
against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "radar-graph_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")
FP_TOL = 1e-10


def run(app, args):
    print("app", app, " ".join(args), flush=True)
    p = subprocess.run([os.path.join(BIN, app)] + args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def val(out, key):
    return re.findall(re.escape(key) + r"\s*=\s*([-+0-9.eE]+)", out)[-1]


def same_fp(a, b, tol=FP_TOL):
    a, b = np.asarray(a), np.asarray(b)
    if not np.array_equal(np.isnan(a), np.isnan(b)):
        return False
    a, b = a[~np.isnan(a)], b[~np.isnan(b)]
    if not np.array_equal(np.isinf(a), np.isinf(b)):
        return False
    f = np.isfinite(a)
    return float(np.max(np.abs(a[f] - b[f]), initial=0.0)) <= tol * max(float(np.max(np.abs(b[f]), initial=0.0)), 1e-300)


@pytest.fixture(scope="module")
def golden():
    import json
    return json.load(open(os.path.join(GOLD, "golden.json")))


def graph_file(rg, tmp_path, name):
    sym = name.endswith("s")
    hg = rg.HostGraph.generate(10, 16, symmetric=sym)
    path = str(tmp_path / (name + ".adj"))
    hg.write(path, rg.FMT_ADJ)
    return hg, path, (["-s"] if sym else [])


@pytest.mark.parametrize("name", ["rmat10s", "rmat10d"])
@pytest.mark.parametrize("pre,var", [("-1", "pno"), ("1", "ds")])
@pytest.mark.parametrize("variant", ["push", "radar"])
def test_bfs_app(rg, orc, golden, tmp_path, name, pre, var, variant):
    hg, path, flag = graph_file(rg, tmp_path, name)
    fx = np.load(os.path.join(GOLD, f"ligra_{name}.npz"))
    src = int(fx["bfs_sources"][0])
    out = str(tmp_path / "parents.bin")
    o = run("BFS", flag + ["-rounds", "0", "-r", str(src), "-preprocess", pre, "-variant", variant,
                           "-k", "64", "-out", out, path])
    s = golden["ligra"][name][f"{var}_bfs"]
    assert int(val(o, "Num. Iters until convergence")) == s["iters"]
    assert int(val(o, "No. of nodes in BFS-Tree")) == s["reached"]
    par = np.fromfile(out, np.uint32)
    off, nb = hg.csr()
    if var == "ds":
        ids = fx["ds_new_ids"]
        off, nb = orc.relabel_csr(off, nb, ids, strict=False)
        src = int(ids[src])
    lv, _ = orc.bfs_levels(off, nb, src)
    assert orc.check_parents(off, nb, src, lv, par) == 0


@pytest.mark.parametrize("name", ["rmat10s", "rmat10d"])
@pytest.mark.parametrize("pre,var", [("-1", "pno"), ("1", "ds")])
@pytest.mark.parametrize("variant,fused", [("push", False), ("hubdup", False), ("radar", False), ("push", True),
                                           ("hubdup", True), ("radar", True), ("blocked", True)])
def test_pagerank_app(rg, golden, tmp_path, name, pre, var, variant, fused):
    # Without -fused every iteration runs the reference's edgeMap / vertexMap /
    # plusReduce sequence through the engine ABI; -fused the device loop.
    _, path, flag = graph_file(rg, tmp_path, name)
    fx = np.load(os.path.join(GOLD, f"ligra_{name}.npz"))
    out = str(tmp_path / "ranks.bin")
    o = run("PageRank", flag + ["-rounds", "0", "-maxiters", "20", "-preprocess", pre, "-variant", variant,
                                "-k", "128", "-out", out] + (["-fused"] if fused else []) + [path])
    gold = golden["ligra"][name][f"{var}_pr"]
    assert int(val(o, "Num Iters")) == gold["iters"]
    ranks = np.fromfile(out, np.float64)
    assert same_fp(ranks, fx[f"{var}_pr_ranks"])
    # The printed L1 (6 significant digits, as the reference prints it).
    assert abs(float(val(o, "L1_Norm")) - gold["l1"]) <= 1e-5 * gold["l1"] + FP_TOL * float(np.sum(ranks))


@pytest.mark.parametrize("name", ["rmat10s", "rmat10d"])
@pytest.mark.parametrize("pre,var", [("-1", "pno"), ("1", "ds")])
def test_radii_app(rg, golden, tmp_path, name, pre, var):
    _, path, flag = graph_file(rg, tmp_path, name)
    fx = np.load(os.path.join(GOLD, f"ligra_{name}.npz"))
    out = str(tmp_path / "radii.bin")
    o = run("Radii", flag + ["-rounds", "0", "-preprocess", pre, "-out", out, path])
    assert int(val(o, "Iters until convergence")) == golden["ligra"][name][f"{var}_radii"]["rounds"]
    assert np.array_equal(np.fromfile(out, np.int32), fx[f"{var}_radii"])


@pytest.mark.parametrize("name", ["rmat10s", "rmat10d"])
@pytest.mark.parametrize("pre,var", [("-1", "pno"), ("1", "ds")])
def test_bc_app(rg, golden, tmp_path, name, pre, var):
    _, path, flag = graph_file(rg, tmp_path, name)
    fx = np.load(os.path.join(GOLD, f"ligra_{name}.npz"))
    out = str(tmp_path / "deps.bin")
    o = run("BC", flag + ["-rounds", "0", "-maxiters", "4", "-preprocess", pre, "-out", out, path])
    assert int(val(o, "Combined Iters (in front traversal)")) == golden["ligra"][name][f"{var}_bc"]["combined"]
    assert same_fp(np.fromfile(out, np.float64), fx[f"{var}_bc_deps"])


def test_rounds_repeat_compute(rg, tmp_path):
    # -rounds R runs Compute R+1 times (ligra.h:607-612); every run prints
    # the same summary.
    _, path, flag = graph_file(rg, tmp_path, "rmat10d")
    o = run("BFS", ["-rounds", "2", "-r", "5", path])
    assert len(re.findall(r"Num. Iters until convergence = (\d+)", o)) == 3
    assert len(set(re.findall(r"No. of nodes in BFS-Tree = (\d+)", o))) == 1


@pytest.mark.parametrize("relabel", [False, True])
def test_tc_local_app(orc, golden, relabel):
    args = ["-g", "10", "-n", "1", "-v", "-a"] + (["-r"] if relabel else [])
    o = run("tc_local", args)
    assert "Verification:           PASS" in o or re.search(r"Verification:\s+PASS", o)
    total = int(re.search(r"^(\d+) triangles", o, re.M).group(1))
    assert total == golden["tc"]["rmat10s"]["pno"]


@pytest.mark.parametrize("name", ["rmat10s", "rmat10d"])
@pytest.mark.parametrize("pre,var", [("-1", "pno"), ("1", "ds")])
@pytest.mark.parametrize("variant,fused", [("push", False), ("radar", False), ("push", True), ("radar", True),
                                           ("blocked", True)])
def test_pagerank_delta_app(rg, golden, tmp_path, name, pre, var, variant, fused):
    # Without -fused every round runs edgeMap / vertexFilter / plusReduce
    # through the engine ABI; -fused the device loop.
    _, path, flag = graph_file(rg, tmp_path, name)
    fx = np.load(os.path.join(GOLD, f"ligra_{name}.npz"))
    out = str(tmp_path / "prd.bin")
    o = run("PageRankDelta", flag + ["-rounds", "0", "-maxiters", "20", "-preprocess", pre, "-variant", variant,
                                     "-k", "128", "-out", out] + (["-fused"] if fused else []) + [path])
    assert int(val(o, "Num Iters")) == golden["ligra"][name][f"{var}_prd"]["iters"]
    assert same_fp(np.fromfile(out, np.float64), fx[f"{var}_prd_ranks"])
