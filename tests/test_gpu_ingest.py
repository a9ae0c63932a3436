"""Fast ingest (rg_graph_load / rg_graph_save): binary graph files straight
to the device and back, against the host readers and writers (bit-exact CSR
and CSC), the new_ids sidecar of a degree-sorted graph, and malformed files
refused before any kernel uses them."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def assert_same_csr(dg, hg):
    off, nb = dg.download()
    hoff, hnb = hg.csr()
    assert np.array_equal(off, hoff) and np.array_equal(nb, hnb)
    if not hg.symmetric:
        ioff, inb = dg.download(in_edges=True)
        hioff, hinb = hg.csr(in_edges=True)
        assert np.array_equal(ioff, hioff) and np.array_equal(inb, hinb)


@pytest.mark.parametrize("gname", ["rmat12s", "rmat12d", "er12d"])
@pytest.mark.parametrize("fmt", ["sg", "bin"])
def test_load_matches_host_reader(rg, ctx, graphs, tmp_path, gname, fmt):
    hg = graphs(gname)
    f = rg.FMT_SG if fmt == "sg" else rg.FMT_ADJ_BINARY
    path = str(tmp_path / f"{gname}.{fmt}")
    print(f"ingest {gname} {fmt}", flush=True)
    hg.write(path, f)
    ref = rg.HostGraph.read(path, f, symmetric=hg.symmetric)
    dg, ids = rg.DeviceGraph.load(ctx, path, f, symmetric=hg.symmetric)
    assert ids is None and (dg.n, dg.m, dg.symmetric) == (ref.n, ref.m, ref.symmetric)
    assert_same_csr(dg, ref)
    assert_same_csr(dg, hg)


def test_load_streams_several_chunks(rg, orc, ctx, tmp_path):
    # 2^20 vertices, ~30M arcs: the neighbour array spans two 64 MiB staging
    # buffers; PageRank on the loaded graph equals the oracle's.
    hg = rg.HostGraph.generate(20, 16, symmetric=True)
    path = str(tmp_path / "rmat20s.sg")
    hg.write(path, rg.FMT_SG)
    dg, _ = rg.DeviceGraph.load(ctx, path, rg.FMT_SG)
    assert_same_csr(dg, hg)
    off, nbr = hg.csr()
    r = dg.pagerank(5, rg.opts("blocked"))
    ref, it, _ = orc.pagerank(off, nbr, 5)
    assert r.iters == it and float(np.max(np.abs(r.ranks - ref)) / np.max(ref)) < 1e-10


@pytest.mark.parametrize("gname", ["rmat14s", "rmat12d"])
@pytest.mark.parametrize("fmt", ["sg", "bin"])
def test_save_load_degree_sorted_with_new_ids(rg, ctx, graphs, tmp_path, gname, fmt):
    # The cache of a DegreeSorted graph: saved from the device with its
    # new_ids, loaded back equal, new_ids from the sidecar; saving without
    # new_ids removes a stale sidecar.
    hg = graphs(gname)
    f = rg.FMT_SG if fmt == "sg" else rg.FMT_ADJ_BINARY
    sg, ids = rg.DeviceGraph(ctx, hg).degree_sort(sort_neighbors=True)
    path = str(tmp_path / f"{gname}_ds.{fmt}")
    sg.save(path, f, new_ids=ids)
    lg, lids = rg.DeviceGraph.load(ctx, path, f, symmetric=hg.symmetric)
    assert lids is not None and np.array_equal(lids, ids)
    for in_edges in ((False, True) if not hg.symmetric else (False,)):
        a, b = sg.download(in_edges), lg.download(in_edges)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    sg.save(path, f)
    _, lids = rg.DeviceGraph.load(ctx, path, f, symmetric=hg.symmetric)
    assert lids is None


def test_save_keeps_original_orientation(rg, ctx, graphs, tmp_path):
    hg = graphs("rmat12d")
    dg = rg.DeviceGraph(ctx, hg)
    dg.transpose()
    path = str(tmp_path / "t.sg")
    dg.save(path)
    lg, _ = rg.DeviceGraph.load(ctx, path)
    assert_same_csr(lg, hg)


def _write_sg(path, n, off, nbr):
    with open(path, "wb") as f:
        f.write(bytes([0]))
        f.write(np.int64(len(nbr)).tobytes())
        f.write(np.int64(n).tobytes())
        f.write(np.asarray(off, np.uint64).tobytes())
        f.write(np.asarray(nbr, np.uint32).tobytes())


@pytest.mark.parametrize("case", ["nbr_out_of_range", "decreasing", "bad_last", "truncated"])
def test_malformed_files_refused(rg, ctx, tmp_path, case):
    path = str(tmp_path / f"{case}.sg")
    off, nbr = [0, 2, 3, 4], [1, 2, 0, 0]
    if case == "nbr_out_of_range":
        nbr = [1, 7, 0, 0]
    elif case == "decreasing":
        off = [0, 3, 2, 4]
    elif case == "bad_last":
        off = [0, 2, 3, 3]
    _write_sg(path, 3, off, nbr)
    if case == "truncated":
        with open(path, "r+b") as f:
            f.truncate(30)
    with pytest.raises(rg.RGError) as e:
        rg.DeviceGraph.load(ctx, path)
    assert e.value.code == rg.RG_EFORMAT
