"""Python binding of the radar_graph C ABI (include/radar_graph.h).

Loads ``librg.so`` from this directory with ctypes; importing fails loudly
when the library has not been built (there is no CPU fallback).  The classes
mirror the reference's objects: :class:`HostGraph` is a Ligra/GAP graph in
host memory, :class:`DeviceGraph` the device CSR that the apps (the
reference's ``Compute`` bodies: PageRank, BFS, Radii, BC, triangle
counting) run on.
"""
from __future__ import annotations

import ctypes as C
import weakref
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librg.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"radar_graph: {LIB_PATH} is missing; build it first "
        "(python -c 'import __graft_entry__ as g; g.build()' or make -C radar-graph_amd)")

lib = C.CDLL(LIB_PATH)

RG_OK, RG_EINVAL, RG_ENOMEM, RG_EDEVICE, RG_EIO, RG_EFORMAT, RG_EUNSUPPORTED = range(7)
FMT_ADJ, FMT_ADJ_BINARY, FMT_EL, FMT_SG, FMT_WEL, FMT_GR, FMT_METIS, FMT_MTX = range(8)
PUSH, PULL, HUBDUP, RADAR, BLOCKED = range(5)
HUBS_LDS, HUBS_L2 = range(2)
VARIANTS = {"push": PUSH, "pull": PULL, "hubdup": HUBDUP, "radar": RADAR, "blocked": BLOCKED}

_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p


class RGError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"radar_graph error {code}: {msg}")
        self.code = code


class Opts(C.Structure):
    """rg_opts: variant, hub placement, hub count k, sparse/dense threshold,
    workgroup size and the wave-per-vertex cut-over degree."""
    _fields_ = [("variant", C.c_int), ("hub_placement", C.c_int), ("hubs", C.c_int64),
                ("threshold", C.c_int64), ("wg_size", C.c_int), ("heavy_degree", C.c_int),
                ("l2_hubs", C.c_int64), ("dst_block_bits", C.c_int), ("src_segment_bits", C.c_int),
                ("unroll", C.c_int), ("hub_min_arcs", C.c_int64), ("unit_order", C.c_int),
                ("unit_queue", C.c_int), ("batch_launch", C.c_int)]


class BfsSummary(C.Structure):
    """rg_bfs_summary: rounds, vertices reached, out-degree sum of the
    reached vertices, level-array checksum."""
    _fields_ = [("iters", C.c_int64), ("reached", C.c_int64), ("edges", C.c_int64),
                ("checksum", C.c_uint64)]

    def astuple(self):
        return (self.iters, self.reached, self.edges, self.checksum)


class BcSummary(C.Structure):
    """rg_bc_summary: forward rounds, vertices reached, weighted sum of the
    finite dependencies, count of the non-finite ones."""
    _fields_ = [("rounds", C.c_int64), ("reached", C.c_int64), ("wsum", C.c_double),
                ("nonfinite", C.c_int64)]


Q_BFS, Q_BC, Q_PAGERANK = range(3)


class Query(C.Structure):
    _fields_ = [("kind", C.c_int), ("source", C.c_uint32), ("iters", C.c_int32), ("arrival_ms", C.c_double)]


class QueryResult(C.Structure):
    _fields_ = [("kind", C.c_int), ("slot", C.c_int), ("batch_id", C.c_int64), ("batch_size", C.c_int64),
                ("bfs", BfsSummary), ("iters", C.c_int64), ("l1", C.c_double), ("wsum", C.c_double),
                ("nonfinite", C.c_int64), ("arrival_ms", C.c_double), ("start_ms", C.c_double),
                ("end_ms", C.c_double), ("latency_ms", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("slots", C.c_int), ("max_batch", C.c_int), ("one_at_a_time", C.c_int),
                ("hbm_reserve", C.c_int64), ("opts", Opts), ("bfs_opts", Opts), ("alone_kinds", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("queries", C.c_int64), ("batches", C.c_int64), ("slots", C.c_int),
                ("makespan_ms", C.c_double), ("qps", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("mean_batch", C.c_double),
                ("hbm_budget", C.c_int64), ("slot_bytes", C.c_int64), ("peak_bytes", C.c_int64),
                ("mem_waits", C.c_int64), ("mean_bc_batch", C.c_double)]

    def asdict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("rg_last_error", C.c_char_p)
_sig("rg_version", C.c_char_p)
_sig("rg_generate_edge_count", C.c_int64, C.c_int, C.c_int)
_sig("rg_generate_edges", C.c_int, C.c_int, C.c_int, C.c_int, _u32p, _u32p)
_sig("rg_hgraph_generate", C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_vp))
_sig("rg_hgraph_from_edges", C.c_int, C.c_int64, C.c_int64, _u32p, _u32p, C.c_int, C.POINTER(_vp))
_sig("rg_hgraph_from_csr", C.c_int, C.c_int64, C.c_int64, _u64p, _u32p, C.c_int, C.c_int, C.POINTER(_vp))
_sig("rg_hgraph_read", C.c_int, C.c_char_p, C.c_int, C.c_int, C.POINTER(_vp))
_sig("rg_hgraph_write", C.c_int, _vp, C.c_char_p, C.c_int)
_sig("rg_hgraph_info", C.c_int, _vp, _i64p, _i64p, C.POINTER(C.c_int))
_sig("rg_hgraph_offsets", _u64p, _vp, C.c_int)
_sig("rg_hgraph_neighbors", _u32p, _vp, C.c_int)
_sig("rg_hgraph_free", None, _vp)
_sig("rg_pick_sources_gap", C.c_int, _vp, C.c_int, _u32p)
_sig("rg_pick_sources_bc", C.c_int, _vp, C.c_int, _u32p)
_sig("rg_hash_int", C.c_uint64, C.c_uint64)
_sig("rg_device_count", C.c_int, C.POINTER(C.c_int))
_sig("rg_ctx_create", C.c_int, C.c_int, C.POINTER(_vp))
_sig("rg_ctx_destroy", C.c_int, _vp)
_sig("rg_ctx_sync", C.c_int, _vp)
_sig("rg_ctx_timing", C.c_int, _vp, C.c_int)
_sig("rg_ctx_graphs", C.c_int, _vp, C.c_int)
_sig("rg_ctx_kernel_time", C.c_int, _vp, _i64p, _f64p)
_sig("rg_device_copy_bandwidth", C.c_int, _vp, C.c_int64, C.c_int, _f64p, _f64p)
_sig("rg_device_read_bandwidth", C.c_int, _vp, C.c_int64, C.c_int, C.c_int64, _f64p, _f64p, _f64p)
_sig("rg_graph_upload", C.c_int, _vp, _vp, C.POINTER(_vp))
_sig("rg_graph_info", C.c_int, _vp, _i64p, _i64p, C.POINTER(C.c_int))
_sig("rg_graph_download", C.c_int, _vp, C.c_int, _u64p, _u32p)
_sig("rg_graph_load", C.c_int, _vp, C.c_char_p, C.c_int, C.c_int, _u32p, C.POINTER(_vp))
_sig("rg_graph_file_info", C.c_int, C.c_char_p, C.c_int, _i64p, _i64p, C.POINTER(C.c_int), C.POINTER(C.c_int))
_sig("rg_graph_save", C.c_int, _vp, C.c_char_p, C.c_int, _u32p)
_sig("rg_graph_transpose", C.c_int, _vp)
_sig("rg_graph_free", C.c_int, _vp)
_sig("rg_opts_default", None, C.POINTER(Opts))
_sig("rg_pagerank", C.c_int, _vp, C.c_int64, C.POINTER(Opts), _f64p, _i64p, _f64p, _f64p)
_sig("rg_pagerank_delta", C.c_int, _vp, C.c_int64, C.POINTER(Opts), _f64p, _i64p, _f64p, _f64p)
_sig("rg_bfs", C.c_int, _vp, C.c_uint32, C.POINTER(Opts), _u32p, _i64p, _i64p, _i64p, _f64p)
_sig("rg_level_hash", C.c_uint64, C.c_uint32, C.c_uint32)
_sig("rg_bfs_query", C.c_int, _vp, C.c_uint32, C.POINTER(Opts), C.POINTER(BfsSummary), _f64p)
_sig("rg_bfs_batch", C.c_int, _vp, _u32p, C.c_int, C.POINTER(Opts), _i32p, C.POINTER(BfsSummary), _f64p)
_sig("rg_serve_config_default", None, C.POINTER(ServeConfig))
_sig("rg_serve", C.c_int, _vp, C.POINTER(ServeConfig), C.POINTER(Query), C.c_int64, C.POINTER(QueryResult),
     C.POINTER(ServeStats))
_sig("rg_radii", C.c_int, _vp, _u32p, C.c_int, C.POINTER(Opts), _i32p, _i64p, _f64p)
_sig("rg_bc", C.c_int, _vp, _u32p, C.c_int, C.POINTER(Opts), _f64p, _i64p, _f64p)
_sig("rg_bc_batch", C.c_int, _vp, _u32p, C.c_int, C.POINTER(Opts), _f64p, C.POINTER(BcSummary), _f64p)
# Engine level: vertexSubset / edgeMap / vertexMap / vertexFilter over device state arrays.
_sig("rg_subset_create", C.c_int, _vp, C.POINTER(_vp))
_sig("rg_subset_set", C.c_int, _vp, _u32p, C.c_int64)
_sig("rg_subset_set_all", C.c_int, _vp)
_sig("rg_subset_size", C.c_int, _vp, _i64p)
_sig("rg_subset_get", C.c_int, _vp, _u32p)
_sig("rg_subset_free", C.c_int, _vp)
_sig("rg_array_alloc", C.c_int, _vp, C.c_int64, C.c_int, C.POINTER(_vp))
_sig("rg_array_write", C.c_int, _vp, _vp, C.c_int64, C.c_int64)
_sig("rg_array_read", C.c_int, _vp, _vp, C.c_int64, C.c_int64)
_sig("rg_array_fill", C.c_int, _vp, C.c_uint64)
_sig("rg_array_free", C.c_int, _vp)
_sig("rg_array_sum", C.c_int, _vp, _f64p)
_sig("rg_array_device_ptr", C.c_int, _vp, C.POINTER(_vp))
_sig("rg_edgemap", C.c_int, _vp, C.c_int, _vp, C.POINTER(_vp), C.c_int, C.c_int64, C.c_int64, C.c_int,
     C.POINTER(Opts), _vp)
_sig("rg_vertexmap", C.c_int, _vp, C.c_int, _vp, C.POINTER(_vp), C.c_int, C.c_double, C.c_double)
_sig("rg_vertexfilter", C.c_int, _vp, C.c_int, _vp, C.POINTER(_vp), C.c_int, _f64p, C.c_int, _vp)
_sig("rg_bfs_noalloc", C.c_int, _vp, C.c_uint32, _vp, C.POINTER(BfsSummary))
FIXED_SCRATCH_BYTES = 64 << 20
NO_OUTPUT, DENSE_FORWARD = 1, 8
OP_BFS, OP_PAGERANK, OP_RADII, OP_BC_FWD, OP_BC_BACK = range(5)
(VOP_PR_UPDATE, VOP_PR_CONTRIB, VOP_RADII_SYNC, VOP_MARK, VOP_BC_BACK, VOP_INVERT, VOP_BC_FINAL,
 VOP_ABS_DIFF, VOP_ABS) = range(9)
FOP_PRD_FIRST, FOP_PRD, FOP_U32_EQ, FOP_U8_NONZERO = range(4)

_OPTIONAL = {}


def _opt_sig(name, res, *args):
    if hasattr(lib, name):
        _OPTIONAL[name] = _sig(name, res, *args)


_opt_sig("rg_graph_degree_sort", C.c_int, _vp, C.c_int, C.c_int, C.POINTER(_vp), _u32p)
_opt_sig("rg_tc", C.c_int, _vp, C.POINTER(Opts), _u64p, _u64p, _f64p)
_sig("rg_autotune", C.c_int, _vp, C.c_int, _u32p, C.c_int, C.POINTER(Opts), C.POINTER(Opts), _f64p)
APP_PAGERANK, APP_BFS = 0, 1


def exported_symbols():
    """Names the loaded library exports among those the header declares."""
    return sorted(n for n in dir(lib) if n.startswith("rg_"))


def _check(rc: int):
    if rc != RG_OK:
        raise RGError(rc, lib.rg_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def opts(variant="push", hubs=0, threshold=-1, wg_size=0, heavy_degree=0, hub_placement=HUBS_LDS,
         l2_hubs=0, dst_block_bits=0, src_segment_bits=0, unroll=0, hub_min_arcs=0,
         unit_order=0, unit_queue=0, batch_launch=0) -> Opts:
    o = Opts()
    lib.rg_opts_default(C.byref(o))
    o.variant = VARIANTS[variant] if isinstance(variant, str) else int(variant)
    o.hubs = int(hubs)
    o.threshold = int(threshold)
    o.wg_size = int(wg_size)
    o.heavy_degree = int(heavy_degree)
    o.hub_placement = int(hub_placement)
    o.l2_hubs = int(l2_hubs)
    o.dst_block_bits = int(dst_block_bits)
    o.src_segment_bits = int(src_segment_bits)
    o.unroll = int(unroll)
    o.hub_min_arcs = int(hub_min_arcs)
    o.unit_order = int(unit_order)
    o.unit_queue = int(unit_queue)
    o.batch_launch = int(batch_launch)
    return o


def generate_edges(scale: int, degree: int, uniform: bool = False):
    ne = lib.rg_generate_edge_count(scale, degree)
    if ne < 0:
        raise RGError(RG_EINVAL, "bad scale/degree")
    s = np.empty(ne, np.uint32)
    d = np.empty(ne, np.uint32)
    _check(lib.rg_generate_edges(scale, degree, int(uniform), _ptr(s, _u32p), _ptr(d, _u32p)))
    return s, d


def hash_int(a: int) -> int:
    return int(lib.rg_hash_int(C.c_uint64(a)))


def serve_config(slots=2, max_batch=256, one_at_a_time=False, hbm_reserve=0, o: Opts = None,
                 bfs: Opts = None, alone_kinds=None) -> ServeConfig:
    c = ServeConfig()
    lib.rg_serve_config_default(C.byref(c))
    c.slots, c.max_batch, c.one_at_a_time = slots, max_batch, int(one_at_a_time)
    if alone_kinds is not None:
        c.alone_kinds = int(alone_kinds)
    if hbm_reserve:
        c.hbm_reserve = hbm_reserve
    if o is not None:
        c.opts = o
    if bfs is not None:
        c.bfs_opts = bfs
    return c


def poisson_workload(hg: "HostGraph", count: int, rate_qps: float, mix=(0.8, 0.15, 0.05), pr_iters=10,
                     seed=27491095):
    """Seeded query trace: exponential inter-arrival times at rate_qps
    (all at time 0 when rate_qps <= 0), kinds drawn with probabilities mix
    (BFS, BC, PageRank), sources uniform over vertices with out-degree > 0.
    Returns a ctypes array of Query."""
    rng = np.random.default_rng(seed)
    if rate_qps > 0:
        t = np.cumsum(rng.exponential(1e3 / rate_qps, size=count))
    else:
        t = np.zeros(count)
    kinds = rng.choice(3, size=count, p=np.asarray(mix, float) / sum(mix))
    deg = np.diff(hg.offsets().astype(np.int64))
    cand = np.nonzero(deg > 0)[0]
    src = cand[rng.integers(0, len(cand), size=count)]
    q = (Query * count)()
    for i in range(count):
        q[i].kind, q[i].source, q[i].iters, q[i].arrival_ms = int(kinds[i]), int(src[i]), pr_iters, float(t[i])
    return q


def level_hash(v: int, level: int) -> int:
    return int(lib.rg_level_hash(v, level))


def summary_weights(n: int) -> np.ndarray:
    """w(v) = (rg_level_hash(v, 0xffffffff) >> 11) * 2^-53 for v < n: the
    weights of the BC / PageRank result summaries (rg_bc_summary.wsum)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(0xffffffff) << np.uint64(32)) | np.arange(n, dtype=np.uint64)
        z = z + np.uint64(0x9e3779b97f4a7c15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def radii_starts(n: int, k: int = 64):
    """Radii's sources: hashInt(i) % n for i < min(n, 64) (Radii.C:117-127)."""
    return np.array([hash_int(i) % n for i in range(min(n, k))], np.uint32)


class HostGraph:
    """Host CSR with the reference readers' dedup semantics."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def generate(cls, scale, degree=16, uniform=False, symmetric=True):
        h = _vp()
        _check(lib.rg_hgraph_generate(scale, degree, int(uniform), int(symmetric), C.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, n, src, dst, symmetric=True):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        h = _vp()
        _check(lib.rg_hgraph_from_edges(n, len(src), _ptr(src, _u32p), _ptr(dst, _u32p),
                                        int(symmetric), C.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, offsets, nbrs, symmetric=True, dedup=False):
        offsets = np.ascontiguousarray(offsets, np.uint64)
        nbrs = np.ascontiguousarray(nbrs, np.uint32)
        h = _vp()
        _check(lib.rg_hgraph_from_csr(len(offsets) - 1, len(nbrs), _ptr(offsets, _u64p),
                                      _ptr(nbrs, _u32p), int(symmetric), int(dedup), C.byref(h)))
        return cls(h)

    @classmethod
    def read(cls, path, fmt=FMT_ADJ, symmetric=True):
        h = _vp()
        _check(lib.rg_hgraph_read(str(path).encode(), fmt, int(symmetric), C.byref(h)))
        return cls(h)

    def write(self, path, fmt=FMT_ADJ):
        _check(lib.rg_hgraph_write(self._h, str(path).encode(), fmt))

    def info(self):
        n, m, s = C.c_int64(), C.c_int64(), C.c_int()
        _check(lib.rg_hgraph_info(self._h, C.byref(n), C.byref(m), C.byref(s)))
        return n.value, m.value, bool(s.value)

    @property
    def n(self):
        return self.info()[0]

    @property
    def m(self):
        return self.info()[1]

    @property
    def symmetric(self):
        return self.info()[2]

    def offsets(self, in_edges=False) -> np.ndarray:
        n = self.n
        p = lib.rg_hgraph_offsets(self._h, int(in_edges))
        return np.ctypeslib.as_array(p, shape=(n + 1,)).copy()

    def neighbors(self, in_edges=False) -> np.ndarray:
        m = self.m
        if m == 0:
            return np.zeros(0, np.uint32)
        p = lib.rg_hgraph_neighbors(self._h, int(in_edges))
        return np.ctypeslib.as_array(p, shape=(m,)).copy()

    def csr(self, in_edges=False):
        return self.offsets(in_edges), self.neighbors(in_edges)

    def pick_sources(self, count):
        out = np.empty(count, np.uint32)
        _check(lib.rg_pick_sources_gap(self._h, count, _ptr(out, _u32p)))
        return out

    def pick_bc_sources(self, count):
        out = np.empty(count, np.uint32)
        _check(lib.rg_pick_sources_bc(self._h, count, _ptr(out, _u32p)))
        return out

    def close(self):
        if self._h:
            lib.rg_hgraph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count() -> int:
    c = C.c_int()
    _check(lib.rg_device_count(C.byref(c)))
    return c.value


class Context:
    """One device, one stream, reusable work pools (rg_ctx)."""

    def __init__(self, device=0):
        self._h = _vp()
        self._graphs = weakref.WeakSet()   # device graphs bound to this context
        self._arrays = weakref.WeakSet()   # device state arrays of this context
        _check(lib.rg_ctx_create(device, C.byref(self._h)))

    def sync(self):
        _check(lib.rg_ctx_sync(self._h))

    def graphs(self, enable=True, bfs_device_levels=False):
        """Iteration loops replayed as HIP graphs (on by default); with
        bfs_device_levels, the single-source pull BFS decides its level
        kinds on the device (rg_ctx_graphs flags)."""
        _check(lib.rg_ctx_graphs(self._h, (1 if enable else 0) | (2 if bfs_device_levels else 0)))

    def timing(self, enable=True):
        """Bracket every edge-map launch with HIP events (resets the record)."""
        _check(lib.rg_ctx_timing(self._h, int(enable)))

    def read_bandwidth(self, nbytes=1 << 30, reps=5, gathers=0):
        """(GB/s read, ms per read pass, ms per pass of `gathers` random
        8-byte loads) of the device read probes."""
        g, ms, msg = C.c_double(), C.c_double(), C.c_double()
        _check(lib.rg_device_read_bandwidth(self._h, int(nbytes), int(reps), int(gathers), C.byref(g),
                                            C.byref(ms), C.byref(msg)))
        return g.value, ms.value, msg.value

    def copy_bandwidth(self, nbytes=1 << 30, reps=10):
        """(GB/s read+write, ms per copy) of a plain device stream copy."""
        g, ms = C.c_double(), C.c_double()
        _check(lib.rg_device_copy_bandwidth(self._h, int(nbytes), int(reps), C.byref(g), C.byref(ms)))
        return g.value, ms.value

    def kernel_time(self):
        """(timed edge-map launches, their summed device ms) since timing()."""
        n, ms = C.c_int64(), C.c_double()
        _check(lib.rg_ctx_kernel_time(self._h, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def close(self):
        """Frees the context; the device graphs bound to it are freed first
        (a graph must not outlive its context)."""
        if self._h:
            for a in list(self._arrays):
                a.close()
            for g in list(self._graphs):
                g.close()
            lib.rg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_NP_OF_ELEM = {1: np.uint8, 4: np.uint32, 8: np.float64}


class DeviceArray:
    """A device state array (rg_array): the apps' newA arrays, the state of
    the engine-level operators.  Elements of 1, 4 or 8 bytes; `dtype` is the
    numpy type read() and write() use (its itemsize must match)."""

    def __init__(self, ctx: Context, count: int, dtype=np.float64):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self._h = _vp()
        _check(lib.rg_array_alloc(ctx._h, self.count, self.dtype.itemsize, C.byref(self._h)))
        ctx._arrays.add(self)

    def write(self, values, first=0):
        a = np.ascontiguousarray(values, self.dtype)
        _check(lib.rg_array_write(self._h, a.ctypes.data_as(_vp), int(first), a.size))

    def read(self, first=0, count=None):
        count = self.count - first if count is None else count
        out = np.empty(count, self.dtype)
        _check(lib.rg_array_read(self._h, out.ctypes.data_as(_vp), int(first), int(count)))
        return out

    def fill(self, value):
        bits = np.array([value], self.dtype).view({1: np.uint8, 4: np.uint32, 8: np.uint64}[self.dtype.itemsize])
        _check(lib.rg_array_fill(self._h, int(bits[0])))

    def sum(self) -> float:
        s = C.c_double()
        _check(lib.rg_array_sum(self._h, C.byref(s)))
        return s.value

    def device_ptr(self) -> int:
        p = _vp()
        _check(lib.rg_array_device_ptr(self._h, C.byref(p)))
        return p.value

    def close(self):
        if self._h:
            lib.rg_array_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Subset:
    """vertexSubset (rg_subset) of a device graph: a device id list or all
    vertices."""

    def __init__(self, dg: "DeviceGraph", ids=None, all_vertices=False):
        self.dg = dg
        self._h = _vp()
        _check(lib.rg_subset_create(dg._h, C.byref(self._h)))
        dg._subsets.add(self)
        if all_vertices:
            _check(lib.rg_subset_set_all(self._h))
        elif ids is not None:
            a = np.ascontiguousarray(ids, np.uint32)
            _check(lib.rg_subset_set(self._h, _ptr(a, _u32p), a.size))

    def size(self) -> int:
        k = C.c_int64()
        _check(lib.rg_subset_size(self._h, C.byref(k)))
        return k.value

    def ids(self) -> np.ndarray:
        out = np.empty(max(self.size(), 1), np.uint32)
        _check(lib.rg_subset_get(self._h, _ptr(out, _u32p)))
        return out[: self.size()]

    def close(self):
        if self._h:
            lib.rg_subset_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _state(arrays):
    return (_vp * max(len(arrays), 1))(*[a._h for a in arrays]), len(arrays)


@dataclass
class PageRankResult:
    ranks: np.ndarray
    iters: int
    l1: float
    ms: float


@dataclass
class BFSResult:
    parents: np.ndarray
    iters: int
    reached: int
    edges: int
    ms: float


class DeviceGraph:
    """Device CSR (+CSC) and the app entry points."""

    def __init__(self, ctx: Context, hg: HostGraph = None, handle=None):
        self.ctx = ctx
        if handle is not None:
            self._h = handle
        else:
            self._h = _vp()
            _check(lib.rg_graph_upload(ctx._h, hg._h, C.byref(self._h)))
        ctx._graphs.add(self)
        self._subsets = weakref.WeakSet()
        n, m, s = C.c_int64(), C.c_int64(), C.c_int()
        _check(lib.rg_graph_info(self._h, C.byref(n), C.byref(m), C.byref(s)))
        self.n, self.m, self.symmetric = n.value, m.value, bool(s.value)

    @classmethod
    def load(cls, ctx: Context, path, fmt=FMT_SG, symmetric=True):
        """Binary graph file straight to the device (rg_graph_load); returns
        (DeviceGraph, new_ids or None when the file has no new_ids sidecar)."""
        n, m, s, has = C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
        _check(lib.rg_graph_file_info(os.fsencode(path), int(fmt), C.byref(n), C.byref(m), C.byref(s), C.byref(has)))
        ids = np.empty(n.value, np.uint32) if has.value else None
        h = _vp()
        _check(lib.rg_graph_load(ctx._h, os.fsencode(path), int(fmt), int(symmetric),
                                 _ptr(ids, _u32p) if ids is not None else None, C.byref(h)))
        return cls(ctx, handle=h), ids

    def save(self, path, fmt=FMT_SG, new_ids=None):
        """Write the graph (rg_graph_save), with a new_ids sidecar if given."""
        ids = None if new_ids is None else np.ascontiguousarray(new_ids, np.uint32)
        _check(lib.rg_graph_save(self._h, os.fsencode(path), int(fmt), _ptr(ids, _u32p) if ids is not None else None))

    def download(self, in_edges=False):
        off = np.empty(self.n + 1, np.uint64)
        nb = np.empty(max(self.m, 1), np.uint32)
        _check(lib.rg_graph_download(self._h, int(in_edges), _ptr(off, _u64p), _ptr(nb, _u32p)))
        return off, nb[: self.m]

    def transpose(self):
        _check(lib.rg_graph_transpose(self._h))

    def degree_sort(self, by_outdeg=False, sort_neighbors=False):
        """GPU DegreeSorting; returns (new DeviceGraph, new_ids)."""
        f = _OPTIONAL.get("rg_graph_degree_sort")
        if f is None:
            raise RGError(RG_EUNSUPPORTED, "rg_graph_degree_sort not in this build")
        ids = np.empty(self.n, np.uint32)
        h = _vp()
        _check(f(self._h, int(by_outdeg), int(sort_neighbors), C.byref(h), _ptr(ids, _u32p)))
        return DeviceGraph(self.ctx, handle=h), ids

    def pagerank(self, maxiters=20, o: Opts = None, want_ranks=True) -> PageRankResult:
        o = o or opts()
        ranks = np.empty(self.n, np.float64) if want_ranks else None
        it, l1, ms = C.c_int64(), C.c_double(), C.c_double()
        _check(lib.rg_pagerank(self._h, maxiters, C.byref(o),
                               _ptr(ranks, _f64p) if want_ranks else None,
                               C.byref(it), C.byref(l1), C.byref(ms)))
        return PageRankResult(ranks, it.value, l1.value, ms.value)

    def pagerank_delta(self, maxiters=20, o: Opts = None, want_ranks=True) -> PageRankResult:
        o = o or opts()
        ranks = np.empty(self.n, np.float64) if want_ranks else None
        it, l1, ms = C.c_int64(), C.c_double(), C.c_double()
        _check(lib.rg_pagerank_delta(self._h, maxiters, C.byref(o), _ptr(ranks, _f64p) if want_ranks else None,
                                     C.byref(it), C.byref(l1), C.byref(ms)))
        return PageRankResult(ranks, it.value, l1.value, ms.value)

    def bfs(self, src, o: Opts = None, want_parents=True) -> BFSResult:
        o = o or opts()
        par = np.empty(self.n, np.uint32) if want_parents else None
        it, r, e, ms = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        _check(lib.rg_bfs(self._h, int(src), C.byref(o), _ptr(par, _u32p) if want_parents else None,
                          C.byref(it), C.byref(r), C.byref(e), C.byref(ms)))
        return BFSResult(par, it.value, r.value, e.value, ms.value)

    def bfs_query(self, src, o: Opts = None):
        """One BFS query on the single-source engine: (BfsSummary, ms)."""
        o = o or opts()
        out, ms = BfsSummary(), C.c_double()
        _check(lib.rg_bfs_query(self._h, int(src), C.byref(o), C.byref(out), C.byref(ms)))
        return out, ms.value

    def bfs_noalloc(self, src, parents: "DeviceArray") -> BfsSummary:
        """Push BFS that allocates nothing (rg_bfs_noalloc): parents into the
        device array `parents` (n uint32), scratch in the GPU's fixed array."""
        if parents.dtype.itemsize != 4 or parents.count < self.n:
            raise RGError(RG_EINVAL, "bfs_noalloc: parents must hold n 4-byte elements")
        out = BfsSummary()
        _check(lib.rg_bfs_noalloc(self._h, int(src), _vp(parents.device_ptr()), C.byref(out)))
        return out

    def bfs_batch(self, sources, o: Opts = None, want_levels=False):
        """Bit-parallel batch of BFS queries: (list of BfsSummary, levels
        [nsources, n] int32 or None, ms)."""
        o = o or opts()
        src = np.ascontiguousarray(sources, np.uint32)
        out = (BfsSummary * len(src))()
        lv = np.empty((len(src), self.n), np.int32) if want_levels else None
        ms = C.c_double()
        _check(lib.rg_bfs_batch(self._h, _ptr(src, _u32p), len(src), C.byref(o),
                                _ptr(lv, _i32p) if want_levels else None, out, C.byref(ms)))
        return list(out), lv, ms.value

    def serve(self, queries, cfg: ServeConfig = None):
        """Runs a query trace through the serving loop: (results, stats)."""
        cfg = cfg or serve_config()
        res = (QueryResult * len(queries))()
        st = ServeStats()
        _check(lib.rg_serve(self._h, C.byref(cfg), queries, len(queries), res, C.byref(st)))
        return res, st

    def radii(self, starts, o: Opts = None):
        o = o or opts()
        starts = np.ascontiguousarray(starts, np.uint32)
        rad = np.empty(self.n, np.int32)
        rounds, ms = C.c_int64(), C.c_double()
        _check(lib.rg_radii(self._h, _ptr(starts, _u32p), len(starts), C.byref(o), _ptr(rad, _i32p),
                            C.byref(rounds), C.byref(ms)))
        return rad, rounds.value, ms.value

    def bc(self, sources, o: Opts = None):
        o = o or opts()
        sources = np.ascontiguousarray(sources, np.uint32)
        deps = np.empty(self.n, np.float64)
        comb, ms = C.c_int64(), C.c_double()
        _check(lib.rg_bc(self._h, _ptr(sources, _u32p), len(sources), C.byref(o), _ptr(deps, _f64p),
                         C.byref(comb), C.byref(ms)))
        return deps, comb.value, ms.value

    def bc_batch(self, sources, o: Opts = None, want_deps=False):
        """1..256 single-source BC queries in batches of 64: (list of
        BcSummary, dependencies [nsources, n] float64 or None, ms)."""
        o = o or opts()
        src = np.ascontiguousarray(sources, np.uint32)
        out = (BcSummary * len(src))()
        deps = np.empty((len(src), self.n), np.float64) if want_deps else None
        ms = C.c_double()
        _check(lib.rg_bc_batch(self._h, _ptr(src, _u32p), len(src), C.byref(o),
                               _ptr(deps, _f64p) if want_deps else None, out, C.byref(ms)))
        return list(out), deps, ms.value

    def autotune(self, app="pagerank", sources=None, base: Opts = None):
        """Fastest engine configuration for `app` on this graph: (Opts, ms)."""
        best, ms = Opts(), C.c_double()
        if app == "pagerank":
            _check(lib.rg_autotune(self._h, APP_PAGERANK, None, 0, C.byref(base) if base else None,
                                   C.byref(best), C.byref(ms)))
        else:
            src = np.ascontiguousarray(sources, np.uint32)
            _check(lib.rg_autotune(self._h, APP_BFS, _ptr(src, _u32p), len(src),
                                   C.byref(base) if base else None, C.byref(best), C.byref(ms)))
        return best, ms.value

    def tc(self, o: Opts = None, per_vertex=False):
        f = _OPTIONAL.get("rg_tc")
        if f is None:
            raise RGError(RG_EUNSUPPORTED, "rg_tc not in this build")
        o = o or opts()
        tot, ms = C.c_uint64(), C.c_double()
        pv = np.empty(self.n, np.uint64) if per_vertex else None
        _check(f(self._h, C.byref(o), C.byref(tot), _ptr(pv, _u64p) if per_vertex else None, C.byref(ms)))
        return (tot.value, pv, ms.value) if per_vertex else (tot.value, ms.value)

    # ---- engine level (Ligra's edgeMap / vertexMap / vertexFilter) --------
    def edgemap(self, op, frontier: Subset, state, round_=0, threshold=-1, flags=0, o: Opts = None,
                out: Subset = None) -> Subset:
        """edgeMap(GA, frontier, F(state...), threshold, flags); returns the
        next frontier (a new Subset unless `out` is given; empty with
        NO_OUTPUT)."""
        out = out if out is not None else Subset(self)
        st, k = _state(state)
        _check(lib.rg_edgemap(self._h, int(op), frontier._h, st, k, int(round_), int(threshold), int(flags),
                              C.byref(o) if o is not None else None, out._h))
        return out

    def vertexmap(self, vop, subset: Subset, state, a=0.0, b=0.0):
        st, k = _state(state)
        _check(lib.rg_vertexmap(self._h, int(vop), subset._h, st, k, float(a), float(b)))

    def vertexfilter(self, fop, subset: Subset, state, params=()) -> Subset:
        out = Subset(self)
        st, k = _state(state)
        pr = np.ascontiguousarray(params if len(params) else [0.0], np.float64)
        _check(lib.rg_vertexfilter(self._h, int(fop), subset._h, st, k, _ptr(pr, _f64p), len(params), out._h))
        return out

    def close(self):
        if self._h:
            for s in list(self._subsets):
                s.close()
            lib.rg_graph_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
