"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of oracle/liboracle.so.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker.  The product library never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


if not os.path.exists(LIB):
    build()

_lib = C.CDLL(LIB)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)


def _s(name, res, *args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_s("orc_bfs_levels", C.c_int64, C.c_int64, _u64p, _u32p, C.c_uint32, _i32p)
_s("orc_bfs_check_parents", C.c_int64, C.c_int64, _u64p, _u32p, C.c_uint32, _i32p, _u32p)
_s("orc_pagerank", None, C.c_int64, _u64p, _u32p, C.c_int64, C.c_int, _f64p, _i64p, _f64p)
_s("orc_pagerank_delta", None, C.c_int64, _u64p, _u32p, C.c_int64, _f64p, _i64p, _f64p)
_s("orc_pagerank_pull", None, C.c_int64, _u64p, _u64p, _u32p, C.c_int64, _f64p, _i64p, _f64p)
_s("orc_radii", None, C.c_int64, _u64p, _u32p, _u32p, C.c_int, _i32p, _i64p)
_s("orc_bc", None, C.c_int64, _u64p, _u32p, _u64p, _u32p, _u32p, C.c_int, _f64p, _i64p)
_s("orc_tc", C.c_uint64, C.c_int64, _u64p, _u32p, _u64p)
_s("orc_tc_verify", C.c_uint64, C.c_int64, _u64p, _u32p)
_s("orc_degree_order", None, C.c_int64, _u32p, _u32p)
_s("orc_relabel_csr", None, C.c_int64, _u64p, _u32p, _u32p, C.c_int, _u64p, _u32p)
_s("orc_identify_hubs", C.c_int64, C.c_int64, _u32p, C.c_int64, _u8p)
_s("orc_hash_int", C.c_uint64, C.c_uint64)


def _p(a, t):
    return a.ctypes.data_as(t)


def _c(a, dt):
    return np.ascontiguousarray(a, dt)


def bfs_levels(off, nbr, src):
    off, nbr = _c(off, np.uint64), _c(nbr, np.uint32)
    n = len(off) - 1
    lv = np.empty(n, np.int32)
    ecc = _lib.orc_bfs_levels(n, _p(off, _u64p), _p(nbr, _u32p), int(src), _p(lv, _i32p))
    return lv, int(ecc)


_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def level_hash(v, level):
    """splitmix64's output function (Steele, Lea and Flood, OOPSLA 2014;
    Vigna's public-domain splitmix64.c) applied to (level << 32 | v) + the
    golden-ratio increment; vectorised over numpy arrays (uint64 wraps)."""
    with np.errstate(over="ignore"):
        z = (np.asarray(level, np.uint64) << np.uint64(32)) | np.asarray(v, np.uint64)
        z = z + _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def bfs_summary(off, level):
    """(iters, reached, edges, checksum) of a BFS level array: rounds with a
    nonempty frontier (BFS.C:107), tree size (BFS.C:108-115), out-degree sum
    of the reached vertices, and the sum of level_hash(v, level) mod 2^64."""
    level = np.asarray(level)
    reached = np.nonzero(level >= 0)[0]
    deg = np.diff(np.asarray(off, np.uint64).astype(np.int64))
    with np.errstate(over="ignore"):
        cs = int(np.sum(level_hash(reached, level[reached]), dtype=np.uint64))
    return (int(level.max()) + 1, len(reached), int(deg[reached].sum()), cs)


def check_parents(off, nbr, src, level, parents):
    off, nbr = _c(off, np.uint64), _c(nbr, np.uint32)
    level, parents = _c(level, np.int32), _c(parents, np.uint32)
    return int(_lib.orc_bfs_check_parents(len(off) - 1, _p(off, _u64p), _p(nbr, _u32p), int(src),
                                          _p(level, _i32p), _p(parents, _u32p)))


def pagerank(off, nbr, maxiters, nthreads=1):
    off, nbr = _c(off, np.uint64), _c(nbr, np.uint32)
    n = len(off) - 1
    r = np.empty(n, np.float64)
    it, l1 = C.c_int64(), C.c_double()
    _lib.orc_pagerank(n, _p(off, _u64p), _p(nbr, _u32p), maxiters, nthreads, _p(r, _f64p),
                      C.byref(it), C.byref(l1))
    return r, it.value, l1.value


def pagerank_delta(off, nbr, maxiters):
    off, nbr = _c(off, np.uint64), _c(nbr, np.uint32)
    n = len(off) - 1
    r = np.empty(n, np.float64)
    it, l1 = C.c_int64(), C.c_double()
    _lib.orc_pagerank_delta(n, _p(off, _u64p), _p(nbr, _u32p), maxiters, _p(r, _f64p), C.byref(it), C.byref(l1))
    return r, it.value, l1.value


def pagerank_pull(off, in_off, in_nbr, maxiters):
    off, in_off, in_nbr = _c(off, np.uint64), _c(in_off, np.uint64), _c(in_nbr, np.uint32)
    n = len(off) - 1
    r = np.empty(n, np.float64)
    it, l1 = C.c_int64(), C.c_double()
    _lib.orc_pagerank_pull(n, _p(off, _u64p), _p(in_off, _u64p), _p(in_nbr, _u32p), maxiters,
                           _p(r, _f64p), C.byref(it), C.byref(l1))
    return r, it.value, l1.value


def radii(off, nbr, starts):
    off, nbr, starts = _c(off, np.uint64), _c(nbr, np.uint32), _c(starts, np.uint32)
    n = len(off) - 1
    rad = np.empty(n, np.int32)
    rounds = C.c_int64()
    _lib.orc_radii(n, _p(off, _u64p), _p(nbr, _u32p), _p(starts, _u32p), len(starts), _p(rad, _i32p),
                   C.byref(rounds))
    return rad, rounds.value


def bc(off, nbr, in_off, in_nbr, sources):
    off, nbr = _c(off, np.uint64), _c(nbr, np.uint32)
    in_off, in_nbr = _c(in_off, np.uint64), _c(in_nbr, np.uint32)
    sources = _c(sources, np.uint32)
    n = len(off) - 1
    deps = np.empty(n, np.float64)
    comb = C.c_int64()
    _lib.orc_bc(n, _p(off, _u64p), _p(nbr, _u32p), _p(in_off, _u64p), _p(in_nbr, _u32p),
                _p(sources, _u32p), len(sources), _p(deps, _f64p), C.byref(comb))
    return deps, comb.value


def tc(off, nbr, per_vertex=False):
    off, nbr = _c(off, np.uint64), _c(nbr, np.uint32)
    n = len(off) - 1
    pv = np.empty(n, np.uint64) if per_vertex else None
    tot = _lib.orc_tc(n, _p(off, _u64p), _p(nbr, _u32p), _p(pv, _u64p) if per_vertex else None)
    return (int(tot), pv) if per_vertex else int(tot)


def tc_verify(off, nbr):
    off, nbr = _c(off, np.uint64), _c(nbr, np.uint32)
    return int(_lib.orc_tc_verify(len(off) - 1, _p(off, _u64p), _p(nbr, _u32p)))


def degree_order(key_deg):
    key_deg = _c(key_deg, np.uint32)
    ids = np.empty(len(key_deg), np.uint32)
    _lib.orc_degree_order(len(key_deg), _p(key_deg, _u32p), _p(ids, _u32p))
    return ids


def relabel_csr(off, nbr, new_ids, strict):
    off, nbr, new_ids = _c(off, np.uint64), _c(nbr, np.uint32), _c(new_ids, np.uint32)
    n = len(off) - 1
    noff = np.empty(n + 1, np.uint64)
    nn = np.empty(max(len(nbr), 1), np.uint32)
    _lib.orc_relabel_csr(n, _p(off, _u64p), _p(nbr, _u32p), _p(new_ids, _u32p), int(strict),
                         _p(noff, _u64p), _p(nn, _u32p))
    return noff, nn[: len(nbr)]


def identify_hubs(deg, num_hubs):
    deg = _c(deg, np.uint32)
    hub = np.empty(len(deg), np.uint8)
    c = _lib.orc_identify_hubs(len(deg), _p(deg, _u32p), int(num_hubs), _p(hub, _u8p))
    return hub.astype(bool), int(c)


def hash_int(a):
    return int(_lib.orc_hash_int(C.c_uint64(a)))


def csr_from_edges(n, src, dst, symmetric):
    """Numpy restatement of the reader's dedup (IO.h:364-467 / builder.h:
    134-163): arcs (u,v) (+ (v,u) if symmetric), self-loops dropped,
    duplicates removed, lists sorted."""
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    if symmetric:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
    keep = src != dst
    key = np.unique(src[keep] * n + dst[keep])
    u = key // n
    v = key % n
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(u, minlength=n), out=off[1:])
    return off, v.astype(np.uint32)


def transpose_csr(off, nbr):
    off = np.asarray(off, np.uint64)
    nbr = np.asarray(nbr, np.uint32)
    n = len(off) - 1
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(off).astype(np.int64))
    key = np.sort(nbr.astype(np.int64) * n + src)
    d = key // n
    s = key % n
    toff = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(d, minlength=n), out=toff[1:])
    return toff, s.astype(np.uint32)


def ref_available():
    return os.path.isdir(REF_DIR) and os.path.exists(os.path.join(REF_DIR, "cap_pno_BFS"))
