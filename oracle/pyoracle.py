"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check the GPU results against.  The product
package never imports it.

The pattern tables come from the committed fixtures tests/golden/dict_k.json,
which hold the reference's own dictionaries (dictionaries/k.csv).
"""
import ctypes
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "liboracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

_lib = None

i32p = ctypes.POINTER(ctypes.c_int32)
i64p = ctypes.POINTER(ctypes.c_int64)
u64p = ctypes.POINTER(ctypes.c_uint64)
u16p = ctypes.POINTER(ctypes.c_uint16)
U64_MAX = (1 << 64) - 1


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_read_edgelist.restype = ctypes.c_int64
        L.or_read_edgelist.argtypes = [ctypes.c_char_p, ctypes.POINTER(i32p), ctypes.POINTER(i32p)]
        L.or_csr_build.restype = ctypes.c_int
        L.or_csr_build.argtypes = [ctypes.c_int32, i32p, i32p, ctypes.c_int64, ctypes.c_int,
                                   i32p, ctypes.POINTER(i64p), ctypes.POINTER(i32p)]
        L.or_relabel.restype = ctypes.c_int
        L.or_relabel.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int, i32p,
                                 ctypes.POINTER(i64p), ctypes.POINTER(i32p)]
        L.or_clique_count.restype = ctypes.c_uint64
        L.or_clique_count.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, u64p]
        L.or_clique_count_roots.restype = ctypes.c_uint64
        L.or_clique_count_roots.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, i32p, ctypes.c_int64, u64p]
        L.or_motif_count.restype = ctypes.c_uint64
        L.or_motif_count.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int, u16p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, u64p, u64p]
        L.or_motif_count_roots.restype = ctypes.c_uint64
        L.or_motif_count_roots.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int, u16p,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, i32p,
                                           ctypes.c_int64, u64p, u64p]
        L.or_brute_clique.restype = ctypes.c_uint64
        L.or_brute_clique.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int]
        L.or_brute_motif.restype = ctypes.c_uint64
        L.or_brute_motif.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int, u16p, ctypes.c_int, u64p]
        L.or_near_clique_count.restype = ctypes.c_uint64
        L.or_near_clique_count.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_brute_near_clique.restype = ctypes.c_uint64
        L.or_brute_near_clique.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int]
        L.or_clique_bytes.restype = ctypes.c_uint64
        L.or_clique_bytes.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int, ctypes.c_int]
        L.or_motif_bytes.restype = ctypes.c_uint64
        L.or_motif_bytes.argtypes = [ctypes.c_int32, i64p, i32p, ctypes.c_int, ctypes.c_int]
        L.or_num_threads.restype = ctypes.c_int
        L.or_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t)


class CSR:
    """Symmetric (or oriented) CSR: n vertices, off[n+1] int64, adj int32."""

    def __init__(self, n, off, adj):
        self.n = int(n)
        self.off = np.ascontiguousarray(off, dtype=np.int64)
        self.adj = np.ascontiguousarray(adj, dtype=np.int32)

    @property
    def nnz(self):
        return int(self.off[-1])

    def degrees(self):
        return np.diff(self.off)

    def neighbours(self, v):
        return self.adj[self.off[v]:self.off[v + 1]]


def _take_csr(n, off_p, adj_p):
    L = lib()
    off = np.ctypeslib.as_array(off_p, shape=(n + 1,)).copy()
    nnz = int(off[-1])
    adj = np.ctypeslib.as_array(adj_p, shape=(max(nnz, 1),))[:nnz].copy()
    L.or_free(ctypes.cast(off_p, ctypes.c_void_p))
    L.or_free(ctypes.cast(adj_p, ctypes.c_void_p))
    return CSR(n, off, adj)


def read_edgelist(path):
    L = lib()
    s = i32p()
    d = i32p()
    m = L.or_read_edgelist(path.encode(), ctypes.byref(s), ctypes.byref(d))
    if m < 0:
        raise ValueError("cannot parse edgelist %s" % path)
    src = np.ctypeslib.as_array(s, shape=(max(m, 1),))[:m].copy()
    dst = np.ctypeslib.as_array(d, shape=(max(m, 1),))[:m].copy()
    L.or_free(ctypes.cast(s, ctypes.c_void_p))
    L.or_free(ctypes.cast(d, ctypes.c_void_p))
    return src, dst


def csr_from_edges(src, dst, n_hint=0, keep_loops=False):
    L = lib()
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    n = ctypes.c_int32()
    off = i64p()
    adj = i32p()
    rc = L.or_csr_build(int(n_hint), _ptr(src, i32p), _ptr(dst, i32p), len(src),
                        1 if keep_loops else 0, ctypes.byref(n), ctypes.byref(off), ctypes.byref(adj))
    if rc != 0:
        raise ValueError("negative vertex id")
    return _take_csr(n.value, off, adj)


def load_edgelist(path, keep_loops=False):
    s, d = read_edgelist(path)
    return csr_from_edges(s, d, 0, keep_loops)


def relabel(g, oriented):
    L = lib()
    perm = np.zeros(max(g.n, 1), dtype=np.int32)
    off = i64p()
    adj = i32p()
    L.or_relabel(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), 1 if oriented else 0,
                 _ptr(perm, i32p), ctypes.byref(off), ctypes.byref(adj))
    return perm[:g.n], _take_csr(g.n, off, adj)


def clique_count(g, k, nthreads=0, faithful=False, per_root=False):
    L = lib()
    pr = np.zeros(max(g.n, 1), dtype=np.uint64) if per_root else None
    r = L.or_clique_count(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), k, nthreads,
                          1 if faithful else 0, _ptr(pr, u64p) if per_root else None)
    if r == U64_MAX:
        raise ValueError("bad k")
    return (int(r), pr[:g.n]) if per_root else int(r)


def clique_count_roots(g, k, roots, nthreads=0, faithful=False):
    L = lib()
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    pr = np.zeros(max(len(roots), 1), dtype=np.uint64)
    r = L.or_clique_count_roots(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), k, nthreads,
                                1 if faithful else 0, _ptr(roots, i32p), len(roots), _ptr(pr, u64p))
    return int(r), pr[:len(roots)]


def _table(k, table):
    """(lut, ncls): the reference dictionary of k, or a caller-given
    (quick -> class) table with its class count (k = 7 has no reference
    dictionary)."""
    if table is None:
        return load_lut(k)
    lut, ncls = table
    return np.ascontiguousarray(lut, dtype=np.uint16), int(ncls)


def motif_count(g, k, nthreads=0, faithful=False, per_root=False, table=None):
    L = lib()
    lut, ncls = _table(k, table)
    hist = np.zeros(ncls, dtype=np.uint64)
    pr = np.zeros(max(g.n, 1), dtype=np.uint64) if per_root else None
    r = L.or_motif_count(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), k, _ptr(lut, u16p), ncls,
                         nthreads, 1 if faithful else 0, _ptr(hist, u64p),
                         _ptr(pr, u64p) if per_root else None)
    if r == U64_MAX:
        raise ValueError("bad k")
    return (int(r), hist, pr[:g.n]) if per_root else (int(r), hist)


def motif_count_roots(g, k, roots, nthreads=0, faithful=False, table=None):
    L = lib()
    lut, ncls = _table(k, table)
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    hist = np.zeros(ncls, dtype=np.uint64)
    pr = np.zeros(max(len(roots), 1), dtype=np.uint64)
    r = L.or_motif_count_roots(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), k, _ptr(lut, u16p), ncls,
                               nthreads, 1 if faithful else 0, _ptr(roots, i32p), len(roots),
                               _ptr(hist, u64p), _ptr(pr, u64p))
    return int(r), hist, pr[:len(roots)]


def near_clique_count(g, k, nthreads=0, faithful=False):
    """The pattern-aware query qk (q4_PA .. q13_PA)."""
    r = lib().or_near_clique_count(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), int(k), nthreads,
                                   1 if faithful else 0)
    if r == U64_MAX:
        raise ValueError("bad k")
    return int(r)


def brute_near_clique(g, k):
    return int(lib().or_brute_near_clique(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), int(k)))


def brute_clique(g, k):
    return int(lib().or_brute_clique(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), k))


def brute_motif(g, k, table=None):
    lut, ncls = _table(k, table)
    hist = np.zeros(ncls, dtype=np.uint64)
    t = lib().or_brute_motif(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), k, _ptr(lut, u16p), ncls,
                             _ptr(hist, u64p))
    return int(t), hist


def clique_bytes(g_oriented, k, nthreads=0):
    return int(lib().or_clique_bytes(g_oriented.n, _ptr(g_oriented.off, i64p),
                                     _ptr(g_oriented.adj, i32p), k, nthreads))


def motif_bytes(g, k, nthreads=0):
    return int(lib().or_motif_bytes(g.n, _ptr(g.off, i64p), _ptr(g.adj, i32p), k, nthreads))


def num_threads():
    return int(lib().or_num_threads())


_dicts = {}


def load_dict(k):
    """The reference dictionary for k (quick -> cgLocal, cgLocal -> cgGlobal)."""
    if k not in _dicts:
        with open(os.path.join(GOLDEN, "dict_%d.json" % k)) as f:
            _dicts[k] = json.load(f)
    return _dicts[k]


def load_lut(k):
    d = load_dict(k)
    lut = np.asarray(d["quick_to_cg_local"], dtype=np.uint16)
    return lut, len(d["cg_local_to_global"])
