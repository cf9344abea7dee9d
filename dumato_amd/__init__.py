"""dumato_amd: MI355X graph pattern mining (k-clique counts, k-motif histograms).

Thin ctypes binding of libdumato_amd.so (C ABI in include/dumato.h).  All
counting runs in the library's HIP kernels; this module only marshals
arguments.  There is no CPU fallback: if the library or a GPU is missing, the
calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdumato_amd.so")

DM_SCHED_PLAIN = 0
DM_SCHED_LEARNED = 1

_ERR = {-1: "DM_ERR_ARG", -2: "DM_ERR_IO", -3: "DM_ERR_HIP", -4: "DM_ERR_UNSUPPORTED",
        -5: "DM_ERR_NOMEM", -6: "DM_ERR_DEVICE"}


class DMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (_ERR.get(code, code), msg))
        self.code = code


class Config(ctypes.Structure):
    """Mirror of struct dm_config."""
    _fields_ = [("sched", ctypes.c_int32), ("vwave", ctypes.c_int32), ("steal", ctypes.c_int32),
                ("steal_min", ctypes.c_int32), ("steal_period", ctypes.c_int32),
                ("lds_bytes", ctypes.c_int32), ("waves_per_block", ctypes.c_int32),
                ("blocks", ctypes.c_int32), ("record_costs", ctypes.c_int32),
                ("per_root", ctypes.c_int32), ("root_begin", ctypes.c_int32),
                ("root_end", ctypes.c_int32), ("clique_bitmap", ctypes.c_int32),
                ("repartition", ctypes.c_int32), ("watchdog_s", ctypes.c_double)]


class Stats(ctypes.Structure):
    """Mirror of struct dm_stats."""
    _fields_ = [("device_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("sched_ms", ctypes.c_double), ("subgraphs", ctypes.c_uint64),
                ("alg_bytes", ctypes.c_uint64), ("donations", ctypes.c_uint64),
                ("jobs", ctypes.c_uint64), ("busy_sum_cycles", ctypes.c_uint64),
                ("busy_max_cycles", ctypes.c_uint64), ("vwave", ctypes.c_int32), ("workers", ctypes.c_int32),
                ("blocks", ctypes.c_int32), ("lds_bytes", ctypes.c_int32),
                ("first_job_ms", ctypes.c_double), ("static_done_ms", ctypes.c_double),
                ("last_job_ms", ctypes.c_double), ("exit_ms", ctypes.c_double),
                ("job_max_cycles", ctypes.c_uint64), ("infer_ms", ctypes.c_double), ("bin_ms", ctypes.c_double),
                ("active_workers", ctypes.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class DfsConfig(ctypes.Structure):
    """Mirror of struct dm_dfs_config."""
    _fields_ = [("vwave", ctypes.c_int32), ("blocks", ctypes.c_int32), ("jobs_per_worker", ctypes.c_int32),
                ("threshold", ctypes.c_int32), ("poll_us", ctypes.c_int32), ("stop_every_us", ctypes.c_int32),
                ("max_rebalances", ctypes.c_int32), ("reserved", ctypes.c_int32), ("watchdog_s", ctypes.c_double)]


class DfsStats(ctypes.Structure):
    """Mirror of struct dm_dfs_stats."""
    _fields_ = [("total_ms", ctypes.c_double), ("lb_ms", ctypes.c_double), ("launches", ctypes.c_int32),
                ("rebalances", ctypes.c_int32), ("workers", ctypes.c_int32), ("vwave", ctypes.c_int32),
                ("donated_jobs", ctypes.c_uint64), ("subgraphs", ctypes.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


DM_DFS_CLIQUE_PA = 0
DM_DFS_CLIQUE_WC = 1
DM_DFS_MOTIF = 2
DM_DFS_NEAR_CLIQUE = 3
DM_PATTERN_CLIQUE = 0
DM_PATTERN_MOTIF = 1
DM_PATTERN_NEAR_CLIQUE = 2

_lib = None

i32p = ctypes.POINTER(ctypes.c_int32)
i64p = ctypes.POINTER(ctypes.c_int64)
u64p = ctypes.POINTER(ctypes.c_uint64)
u16p = ctypes.POINTER(ctypes.c_uint16)
f32p = ctypes.POINTER(ctypes.c_float)
GP = ctypes.c_void_p

# (name, restype, argtypes) of every exported entry point (include/dumato.h)
SYMBOLS = [
    ("dm_version", ctypes.c_int, []),
    ("dm_last_error", ctypes.c_char_p, []),
    ("dm_config_init", ctypes.c_int, [ctypes.POINTER(Config)]),
    ("dm_graph_load_edgelist", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(GP)]),
    ("dm_graph_from_edges", ctypes.c_int, [i32p, i32p, ctypes.c_size_t, ctypes.c_int32, ctypes.POINTER(GP)]),
    ("dm_graph_info", ctypes.c_int, [GP, i32p, i64p, i32p]),
    ("dm_graph_csr", ctypes.c_int, [GP, i64p, ctypes.c_size_t, i32p, ctypes.c_size_t]),
    ("dm_graph_free", None, [GP]),
    ("dm_gen_erdos_renyi", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.POINTER(GP)]),
    ("dm_gen_chung_lu", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_uint64, ctypes.POINTER(GP)]),
    ("dm_gen_rmat", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_uint64, ctypes.POINTER(GP)]),
    ("dm_gen_barabasi_albert", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(GP)]),
    ("dm_clique_count", ctypes.c_int, [GP, ctypes.c_int, ctypes.POINTER(Config), u64p]),
    ("dm_motif_count", ctypes.c_int, [GP, ctypes.c_int, ctypes.POINTER(Config), u64p, ctypes.c_size_t, u64p]),
    ("dm_count_noalloc", ctypes.c_int, [GP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Config), u64p,
                                        ctypes.c_size_t, u64p]),
    ("dm_last_stats", ctypes.c_int, [GP, ctypes.POINTER(Stats)]),
    ("dm_last_root_costs", ctypes.c_int, [GP, u64p, ctypes.c_size_t]),
    ("dm_last_root_counts", ctypes.c_int, [GP, u64p, ctypes.c_size_t]),
    ("dm_motif_num_classes", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]),
    ("dm_pattern_table", ctypes.c_int, [ctypes.c_int, u16p, ctypes.c_size_t, u64p, ctypes.c_size_t]),
    ("dm_sched_train", ctypes.c_int, [GP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_double)]),
    ("dm_sched_predict", ctypes.c_int, [GP, f32p, ctypes.c_size_t]),
    ("dm_autotune", ctypes.c_int, [GP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(Config)]),
    ("dm_batch_create", ctypes.c_int, [ctypes.POINTER(GP), ctypes.c_size_t, ctypes.POINTER(GP)]),
    ("dm_batch_num_results", ctypes.c_int, [GP, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("dm_batch_run", ctypes.c_int, [GP, ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(Config), u64p,
                                    ctypes.c_size_t, u64p, ctypes.POINTER(ctypes.c_double)]),
    ("dm_batch_autotune", ctypes.c_int, [GP, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_double,
                                         ctypes.POINTER(Config)]),
    ("dm_batch_last_stats", ctypes.c_int, [GP, ctypes.POINTER(Stats)]),
    ("dm_batch_free", None, [GP]),
    ("dm_mlp_grad", ctypes.c_int, [f32p, f32p, f32p, ctypes.c_int, f32p, f32p, f32p]),
    ("dm_mlp_adam", ctypes.c_int, [f32p, f32p, f32p, ctypes.c_int, ctypes.c_int, ctypes.c_float, f32p]),
    ("dm_dfs_config_init", ctypes.c_int, [ctypes.POINTER(DfsConfig)]),
    ("dm_dfs_create", ctypes.c_int, [GP, ctypes.c_int, ctypes.c_int, ctypes.POINTER(DfsConfig), ctypes.POINTER(GP)]),
    ("dm_dfs_kernel", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("dm_dfs_run", ctypes.c_int, [GP, ctypes.c_void_p, u64p, ctypes.c_size_t, ctypes.POINTER(DfsStats)]),
    ("dm_dfs_free", None, [GP]),
    ("dm_validate_counter", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_uint64,
                                           ctypes.POINTER(ctypes.c_int)]),
    ("dm_validate_pattern", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, u64p, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), u64p,
                                           ctypes.POINTER(ctypes.c_size_t)]),
]


def lib():
    """Load libdumato_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libdumato_amd.so not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SYMBOLS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise DMError(rc, lib().dm_last_error().decode(errors="replace"))


def _ptr(a, t):
    return a.ctypes.data_as(t)


def config(**kw):
    """A dm_config with the library defaults, overridden by keyword."""
    c = Config()
    _check(lib().dm_config_init(ctypes.byref(c)))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class Graph:
    """Owns a dm_graph handle (host CSR + device copy on first use)."""

    def __init__(self, handle):
        self._h = GP(handle)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.dm_graph_free(h)
            self._h = None

    @classmethod
    def _make(cls, fn, *args):
        h = GP()
        _check(fn(*args, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edgelist(cls, path):
        return cls._make(lib().dm_graph_load_edgelist, str(path).encode())

    @classmethod
    def from_edges(cls, src, dst, n=0):
        s = np.ascontiguousarray(src, dtype=np.int32)
        d = np.ascontiguousarray(dst, dtype=np.int32)
        return cls._make(lib().dm_graph_from_edges, _ptr(s, i32p), _ptr(d, i32p), len(s), int(n))

    @classmethod
    def erdos_renyi(cls, n, m, seed):
        return cls._make(lib().dm_gen_erdos_renyi, n, m, seed)

    @classmethod
    def chung_lu(cls, n, m, gamma, wmax, seed):
        return cls._make(lib().dm_gen_chung_lu, n, m, gamma, wmax, seed)

    @classmethod
    def rmat(cls, scale, edge_factor, a, b, c, seed):
        return cls._make(lib().dm_gen_rmat, scale, edge_factor, a, b, c, seed)

    @classmethod
    def barabasi_albert(cls, n, m, planted, kmin, kmax, seed):
        return cls._make(lib().dm_gen_barabasi_albert, n, m, planted, kmin, kmax, seed)

    def info(self):
        n = ctypes.c_int32()
        m = ctypes.c_int64()
        md = ctypes.c_int32()
        _check(lib().dm_graph_info(self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(md)))
        return n.value, m.value, md.value

    @property
    def n(self):
        return self.info()[0]

    @property
    def m(self):
        return self.info()[1]

    def csr(self):
        n, m, _ = self.info()
        off = np.zeros(n + 1, np.int64)
        adj = np.zeros(max(2 * m, 1), np.int32)
        _check(lib().dm_graph_csr(self._h, _ptr(off, i64p), len(off), _ptr(adj, i32p), len(adj)))
        return off, adj[:2 * m]

    # ---- counting ----
    def clique_count(self, k, cfg=None):
        out = ctypes.c_uint64()
        _check(lib().dm_clique_count(self._h, int(k), ctypes.byref(cfg) if cfg else None, ctypes.byref(out)))
        return int(out.value)

    def motif_count(self, k, cfg=None):
        ncls = num_classes(k)
        hist = np.zeros(ncls, np.uint64)
        tot = ctypes.c_uint64()
        _check(lib().dm_motif_count(self._h, int(k), ctypes.byref(cfg) if cfg else None,
                                    _ptr(hist, u64p), ncls, ctypes.byref(tot)))
        return int(tot.value), hist

    def count_noalloc(self, motif, k, cfg=None):
        """dm_count_noalloc: (total, counts) -- counts = [clique count] or the
        motif histogram -- computed without device allocation."""
        n = num_classes(k) if motif else 1
        out = np.zeros(n, np.uint64)
        tot = ctypes.c_uint64()
        _check(lib().dm_count_noalloc(self._h, 1 if motif else 0, int(k), ctypes.byref(cfg) if cfg else None,
                                      _ptr(out, u64p), n, ctypes.byref(tot)))
        return int(tot.value), out

    def last_stats(self):
        s = Stats()
        _check(lib().dm_last_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def last_root_costs(self):
        out = np.zeros(max(self.n, 1), np.uint64)
        _check(lib().dm_last_root_costs(self._h, _ptr(out, u64p), len(out)))
        return out[:self.n]

    def last_root_counts(self):
        out = np.zeros(max(self.n, 1), np.uint64)
        _check(lib().dm_last_root_counts(self._h, _ptr(out, u64p), len(out)))
        return out[:self.n]

    # ---- learned scheduler ----
    def sched_train(self, motif, k_train, epochs=30, seed=1):
        loss = ctypes.c_double()
        _check(lib().dm_sched_train(self._h, 1 if motif else 0, int(k_train), int(epochs), int(seed),
                                    ctypes.byref(loss)))
        return loss.value

    def autotune(self, motif, k, budget_ms=200.0):
        """Tune width / LDS stack / steal granularity for (mode, k); returns the Config."""
        c = Config()
        _check(lib().dm_autotune(self._h, 1 if motif else 0, int(k), float(budget_ms), ctypes.byref(c)))
        return c

    def sched_predict(self):
        out = np.zeros(max(self.n, 1), np.float32)
        _check(lib().dm_sched_predict(self._h, _ptr(out, f32p), len(out)))
        return out[:self.n]


def num_classes(k):
    n = ctypes.c_size_t()
    _check(lib().dm_motif_num_classes(int(k), ctypes.byref(n)))
    return int(n.value)


def pattern_table(k):
    """(quick -> class id, class id -> canonical code) regenerated by the library."""
    ncls = num_classes(k)
    nq = 1 << ((k + 1) * (k - 2) // 2)
    q2c = np.zeros(nq, np.uint16)
    codes = np.zeros(ncls, np.uint64)
    _check(lib().dm_pattern_table(int(k), _ptr(q2c, u16p), nq, _ptr(codes, u64p), ncls))
    return q2c, codes


def validate_counter(graph_path, k, result):
    """True when "<graph_path>.<k>.aggregateCounter" (Peregrine output) sums to
    result; DuMatoCPU::validateAggregateCounter, src/main/DuMatoCPU.cu:564-595.
    Raises DMError when the file cannot be opened."""
    m = ctypes.c_int()
    _check(lib().dm_validate_counter(os.fsencode(graph_path), int(k), int(result), ctypes.byref(m)))
    return bool(m.value)


def validate_pattern(graph_path, k, hist):
    """(count_match, pattern_match) of hist against "<graph_path>.<k>.aggregatePattern":
    equal totals, and equal sorted multisets of the positive counts (class ids are
    not compared); DuMatoCPU::validateAggregatePattern, src/main/DuMatoCPU.cu:597-661."""
    h = np.ascontiguousarray(hist, dtype=np.uint64)
    cm, pm = ctypes.c_int(), ctypes.c_int()
    _check(lib().dm_validate_pattern(os.fsencode(graph_path), int(k), _ptr(h, u64p), len(h),
                                     ctypes.byref(cm), ctypes.byref(pm), None, None))
    return bool(cm.value), bool(pm.value)


MLP_PARAMS = 4801


def mlp_grad(params, X, Y):
    """(loss, grad[4801], outputs[n]) of the library's MLP kernels."""
    P = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    n = len(Y)
    out = np.zeros(n, np.float32)
    grad = np.zeros(MLP_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    _check(lib().dm_mlp_grad(_ptr(P, f32p), _ptr(X, f32p), _ptr(Y, f32p), n, _ptr(out, f32p),
                             _ptr(grad, f32p), _ptr(loss, f32p)))
    return float(loss[0]), grad, out


def mlp_adam(params, X, Y, steps, lr):
    """Parameters after `steps` full-batch Adam steps, and the losses before each."""
    P = np.array(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    losses = np.zeros(steps, np.float32)
    _check(lib().dm_mlp_adam(_ptr(P, f32p), _ptr(X, f32p), _ptr(Y, f32p), len(Y), int(steps),
                             float(lr), _ptr(losses, f32p)))
    return P, losses


def dfs_config(**kw):
    """A dm_dfs_config with the library defaults, overridden by keyword."""
    c = DfsConfig()
    _check(lib().dm_dfs_config_init(ctypes.byref(c)))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def dfs_kernel(kind, vwave):
    """Host stub of one of the library's step-API kernels (DM_DFS_*)."""
    p = ctypes.c_void_p()
    _check(lib().dm_dfs_kernel(int(kind), int(vwave), ctypes.byref(p)))
    return p.value


class Dfs:
    """Per-worker DFS state of the step API (include/dumato_device.h) for one
    (graph, k, pattern), run with host-driven load balancing.  pattern:
    DM_PATTERN_CLIQUE / DM_PATTERN_MOTIF / DM_PATTERN_NEAR_CLIQUE (False /
    True stand for clique / motif)."""

    def __init__(self, graph, k, pattern, cfg=None):
        pattern = int(pattern)
        self.graph, self.k, self.pattern = graph, int(k), pattern
        self.motif = pattern == DM_PATTERN_MOTIF
        self.cfg = cfg if cfg is not None else dfs_config()
        h = GP()
        _check(lib().dm_dfs_create(graph._h, self.k, pattern, ctypes.byref(self.cfg), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.dm_dfs_free(h)
            self._h = None

    def run(self, kernel=None):
        """Runs `kernel` (a host stub; default: the library's kernel of the
        state's kind and width).  Returns (total, counts, stats)."""
        if kernel is None:
            kind = {DM_PATTERN_CLIQUE: DM_DFS_CLIQUE_PA, DM_PATTERN_MOTIF: DM_DFS_MOTIF,
                    DM_PATTERN_NEAR_CLIQUE: DM_DFS_NEAR_CLIQUE}[self.pattern]
            kernel = dfs_kernel(kind, self.cfg.vwave or 32)
        n = num_classes(self.k) if self.motif else 1
        out = np.zeros(n, np.uint64)
        st = DfsStats()
        _check(lib().dm_dfs_run(self._h, ctypes.c_void_p(kernel), _ptr(out, u64p), n, ctypes.byref(st)))
        return int(out.sum()), out, st.as_dict()


class Query(ctypes.Structure):
    """Mirror of struct dm_query."""
    _fields_ = [("graph", ctypes.c_int32), ("motif", ctypes.c_int32), ("k", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Batch:
    """Many graphs and (pattern, k) queries in one persistent launch (hipGraph-replayed)."""

    def __init__(self, graphs):
        arr = (GP * len(graphs))(*[g._h.value for g in graphs])
        h = GP()
        _check(lib().dm_batch_create(arr, len(graphs), ctypes.byref(h)))
        self._h = h
        self.ngraphs = len(graphs)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.dm_batch_free(h)
            self._h = None

    @staticmethod
    def _queries(queries):
        q = (Query * len(queries))()
        for i, (g, mode, k) in enumerate(queries):
            q[i].graph, q[i].motif, q[i].k = int(g), 1 if mode == "motif" else 0, int(k)
        return q

    def run(self, queries, cfg=None):
        """queries: [(graph index, "clique"|"motif", k)].  Returns (list of results
        -- int for cliques, uint64 histogram for motifs --, per-graph latency ms)."""
        q = self._queries(queries)
        nres = ctypes.c_size_t()
        _check(lib().dm_batch_num_results(self._h, q, len(queries), ctypes.byref(nres)))
        res = np.zeros(max(nres.value, 1), np.uint64)
        offs = np.zeros(len(queries) + 1, np.uint64)
        lat = (ctypes.c_double * self.ngraphs)()
        _check(lib().dm_batch_run(self._h, q, len(queries), ctypes.byref(cfg) if cfg else None,
                                  _ptr(res, u64p), len(res), _ptr(offs, u64p), lat))
        out = []
        for i, (_, mode, _) in enumerate(queries):
            a, b = int(offs[i]), int(offs[i + 1])
            out.append(res[a:b].copy() if mode == "motif" else int(res[a]))
        return out, np.array(lat[:], dtype=np.float64)

    def autotune(self, queries, budget_ms=2000.0):
        q = self._queries(queries)
        c = Config()
        _check(lib().dm_batch_autotune(self._h, q, len(queries), float(budget_ms), ctypes.byref(c)))
        return c

    def last_stats(self):
        s = Stats()
        _check(lib().dm_batch_last_stats(self._h, ctypes.byref(s)))
        return s.as_dict()
