"""ctypes binding of the C ABI in include/tm_abi.h (libtm_hip.so).

The library is loaded from everest_amd/lib/; if it is missing the import of
this module succeeds but every call raises, so no code path silently falls
back to a CPU implementation.
"""
import ctypes
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
# EVEREST_TM_LIB points at another build of the same library (tuning experiments)
LIB_PATH = os.environ.get("EVEREST_TM_LIB") or os.path.join(LIB_DIR, "libtm_hip.so")
MAX_E = 8

TM_OK = 0
STATUS = {0: "TM_OK", 1: "TM_ERR_ARG", 2: "TM_ERR_MOTIF", 3: "TM_ERR_HIP", 4: "TM_ERR_NOMEM",
          5: "TM_ERR_INTERNAL", 6: "TM_ERR_NODEVICE", 7: "TM_ERR_TIMEOUT"}
TM_ERR_TIMEOUT = 7


class TMError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


class HostGraph(ctypes.Structure):
    _fields_ = [("num_edges", ctypes.c_int32), ("num_vertices", ctypes.c_int32),
                ("Eg", ctypes.c_void_p), ("inEdgesV", ctypes.c_void_p), ("outEdgesV", ctypes.c_void_p),
                ("inEdgesR", ctypes.c_void_p), ("outEdgesR", ctypes.c_void_p), ("Euid", ctypes.c_void_p)]


class Tuning(ctypes.Structure):
    _fields_ = [("waves_per_cu", ctypes.c_int32), ("claim_blocks", ctypes.c_int32),
                ("scan_max", ctypes.c_int32), ("time_limit_ms", ctypes.c_int32),
                ("steal_every", ctypes.c_int32), ("steal_max", ctypes.c_int32),
                ("steal_min", ctypes.c_int32), ("steal_after_us", ctypes.c_int32),
                ("waves_per_simd_hint", ctypes.c_int32), ("sb_flush", ctypes.c_int32),
                ("sb_inline", ctypes.c_int32), ("tail_swap", ctypes.c_int32), ("batch_mode", ctypes.c_int32),
                ("sweep", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int32), ("max_batch", ctypes.c_int32), ("arena_bytes", ctypes.c_int64),
                ("tune", Tuning)]


class ServerResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("batch", ctypes.c_int32), ("count", ctypes.c_uint64),
                ("rows", ctypes.c_int64), ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libtm_hip.so is not built (expected at %s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P, I32, I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        L.tm_abi_version.restype = I32
        L.tm_last_error.restype = ctypes.c_char_p
        L.tm_device_count.restype = I32
        L.tm_graph_upload.argtypes = [I32, ctypes.POINTER(HostGraph), ctypes.POINTER(P)]
        L.tm_graph_free.argtypes = [P]
        L.tm_time_order.argtypes = [I32, I64, P, P]
        L.tm_csr_build.argtypes = [I32, I64, P, P, P, P, P, P, P, P, P, ctypes.POINTER(I32)]
        L.tm_graph_info.argtypes = [P, ctypes.POINTER(I64), ctypes.POINTER(I32), ctypes.POINTER(I32),
                                    ctypes.POINTER(I32)]
        L.tm_count.argtypes = [P, I32, P, P, I32, I32, I32, ctypes.POINTER(Tuning), P]
        L.tm_enum.argtypes = [P, I32, P, P, I32, I32, I32, I64, ctypes.POINTER(Tuning), P, P,
                              ctypes.POINTER(I64)]
        L.tm_batch_count.argtypes = [P, I32, P, P, P, P, ctypes.POINTER(Tuning), P]
        L.tm_batch_count_noalloc.argtypes = [P, I32, P, P, P, P, ctypes.POINTER(Tuning), P]
        L.tm_last_stats.argtypes = [P, P, I32]
        L.tm_last_kernel_ms.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
        L.tm_count_work.argtypes = [P, I32, P, P, I32, I32, I32, ctypes.POINTER(Tuning), P, P]
        L.tm_count_bytes.argtypes = [P, I32, P, P, I32, I32, I32, ctypes.POINTER(Tuning), P, P]
        L.tm_shape_class.argtypes = [I32, P, P, ctypes.POINTER(I32)]
        L.tm_set_class_tuning.argtypes = [P, I32, ctypes.POINTER(Tuning)]
        L.tm_batch_prepare.argtypes = [P, I32, P, P, P, P, ctypes.POINTER(Tuning), ctypes.POINTER(P)]
        L.tm_batch_run.argtypes = [P, P]
        L.tm_batch_free.argtypes = [P]
        L.tm_server_create.argtypes = [P, ctypes.POINTER(ServerConfig), ctypes.POINTER(P)]
        L.tm_server_submit.argtypes = [P, I32, P, P, I32, I64, ctypes.POINTER(ctypes.c_uint64)]
        L.tm_server_wait.argtypes = [P, ctypes.c_uint64, I32, ctypes.POINTER(ServerResult)]
        L.tm_server_take_rows.argtypes = [P, ctypes.c_uint64, P, I64, ctypes.POINTER(I64)]
        L.tm_server_release.argtypes = [P, ctypes.c_uint64]
        L.tm_server_stats.argtypes = [P, P, I32]
        L.tm_server_destroy.argtypes = [P]
        for name in ("tm_graph_upload", "tm_graph_free", "tm_graph_info", "tm_count", "tm_enum",
                     "tm_batch_count", "tm_batch_count_noalloc", "tm_last_stats", "tm_last_kernel_ms", "tm_count_work", "tm_count_bytes",
                     "tm_batch_prepare", "tm_batch_run", "tm_batch_free", "tm_shape_class", "tm_set_class_tuning",
                     "tm_server_create", "tm_server_submit", "tm_server_wait", "tm_server_take_rows",
                     "tm_server_release", "tm_server_stats", "tm_server_destroy"):
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def check(status):
    if status != TM_OK:
        raise TMError(status, lib().tm_last_error().decode(errors="replace"))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


TUNING_FIELDS = [f for f, _ in Tuning._fields_]


def tuning(**kw):
    """A Tuning with the given fields set (zero fields take library defaults)."""
    t = Tuning()
    for k, v in kw.items():
        if k not in TUNING_FIELDS:
            raise TypeError("unknown tuning field %r" % k)
        setattr(t, k, int(v))
    return t


def tuning_dict(t):
    return {f: int(getattr(t, f)) for f in TUNING_FIELDS}


def time_order(T, device=0):
    """Stable order of positions by time, computed on the GPU (tm_time_order)."""
    T = np.ascontiguousarray(T, dtype=np.int32)
    out = np.zeros(T.size, dtype=np.int32)
    check(lib().tm_time_order(int(device), ctypes.c_int64(T.size), _p(T), _p(out)))
    return out


def csr_build(U, V, T, device=0):
    """The reference's graph arrays of a time-sorted edge list, built on the
    GPU (tm_csr_build): dict with Eg (n x 3), outV, inV, outR, inR, rename
    (raw id of each new id), V (vertex count)."""
    U = np.ascontiguousarray(U, dtype=np.int32)
    V = np.ascontiguousarray(V, dtype=np.int32)
    T = np.ascontiguousarray(T, dtype=np.int32)
    n = U.size
    Eg = np.zeros(3 * n, dtype=np.int32)
    outV, inV = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    outR, inR, raw = (np.zeros(2 * n + 1, dtype=np.int32) for _ in range(3))
    nv = ctypes.c_int32()
    check(lib().tm_csr_build(int(device), ctypes.c_int64(n), _p(U), _p(V), _p(T), _p(Eg), _p(outV), _p(inV),
                             _p(outR), _p(inR), _p(raw), ctypes.byref(nv)))
    k = nv.value
    return dict(Eg=Eg.reshape(-1, 3), outV=outV, inV=inV, outR=outR[:k + 1].copy(), inR=inR[:k + 1].copy(),
                rename=raw[:k].copy(), V=k)


def shape_class(edges):
    """Shape class of a motif (stack depth and node-map width of its launch)."""
    src, dst = motif_arrays(edges)
    c = ctypes.c_int32()
    check(lib().tm_shape_class(len(edges), _p(src), _p(dst), ctypes.byref(c)))
    return c.value


def motif_arrays(edges):
    """(src, dst) int32 arrays of a renamed motif given as (a, b[, k]) tuples."""
    src = np.array([e[0] for e in edges], dtype=np.int32)
    dst = np.array([e[1] for e in edges], dtype=np.int32)
    return src, dst


class DeviceGraph:
    """A graph resident in HBM (tm_graph handle)."""

    def __init__(self, Eg, inV, outV, inR, outR, Euid, device=0):
        L = lib()
        self._keep = [np.ascontiguousarray(a, dtype=np.int32) for a in (Eg, inV, outV, inR, outR, Euid)]
        Eg, inV, outV, inR, outR, Euid = self._keep
        hg = HostGraph(int(Eg.size // 3), int(inR.size - 1), _p(Eg), _p(inV), _p(outV), _p(inR), _p(outR),
                       _p(Euid))
        h = ctypes.c_void_p()
        check(L.tm_graph_upload(int(device), ctypes.byref(hg), ctypes.byref(h)))
        self.handle = h
        self.num_edges = int(Eg.size // 3)
        self._keep = None

    @classmethod
    def from_dict(cls, g, device=0):
        return cls(g["Eg"], g["inV"], g["outV"], g["inR"], g["outR"], g["Euid"], device)

    def info(self):
        b, e, v, d = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(lib().tm_graph_info(self.handle, ctypes.byref(b), ctypes.byref(e), ctypes.byref(v),
                                  ctypes.byref(d)))
        return dict(device_bytes=b.value, num_edges=e.value, num_vertices=v.value, device=d.value)

    def last_stats(self):
        st = np.zeros(17, dtype=np.uint64)
        check(lib().tm_last_stats(self.handle, _p(st), 17))
        return dict(steps=int(st[0]), items=int(st[1]), roots=int(st[2]), leaves=int(st[3]), pushes=int(st[4]),
                    donated=int(st[8]), taken=int(st[9]), capexit=int(st[10]), work_ticks=int(st[11]),
                    grid=int(st[12]), claim_blocks=int(st[13]), lds=int(st[14]), wpe=int(st[15]),
                    sweep=int(st[16]))

    def last_kernel_ms(self):
        ms = ctypes.c_double()
        check(lib().tm_last_kernel_ms(self.handle, ctypes.byref(ms)))
        return ms.value

    def count_work(self, edges, delta, beg=0, end=-1, tune=None):
        """(count, [R, X1, X2, P]) with the reference's work counted on the GPU."""
        src, dst = motif_arrays(edges)
        out = np.zeros(1, dtype=np.uint64)
        work = np.zeros(4, dtype=np.uint64)
        t = tune if tune is not None else Tuning()
        check(lib().tm_count_work(self.handle, len(edges), _p(src), _p(dst), int(delta), int(beg), int(end),
                                  ctypes.byref(t), _p(out), _p(work)))
        return int(out[0]), [int(x) for x in work]

    def count_bytes(self, edges, delta, beg=0, end=-1, tune=None):
        """(count, bytes read by this library's own search), from the metered kernel."""
        src, dst = motif_arrays(edges)
        out = np.zeros(2, dtype=np.uint64)
        t = tune if tune is not None else Tuning()
        check(lib().tm_count_bytes(self.handle, len(edges), _p(src), _p(dst), int(delta), int(beg), int(end),
                                   ctypes.byref(t), _p(out[:1]), _p(out[1:])))
        return int(out[0]), int(out[1])

    def count(self, edges, delta, beg=0, end=-1, tune=None):
        src, dst = motif_arrays(edges)
        out = np.zeros(1, dtype=np.uint64)
        t = tune if tune is not None else Tuning()
        check(lib().tm_count(self.handle, len(edges), _p(src), _p(dst), int(delta), int(beg), int(end),
                             ctypes.byref(t), _p(out)))
        return int(out[0])

    def enum(self, edges, delta, cap, beg=0, end=-1, tune=None):
        src, dst = motif_arrays(edges)
        M = len(edges)
        rows = np.zeros((max(int(cap), 1), M), dtype=np.int32)
        cnt = np.zeros(1, dtype=np.uint64)
        nw = ctypes.c_int64()
        t = tune if tune is not None else Tuning()
        check(lib().tm_enum(self.handle, M, _p(src), _p(dst), int(delta), int(beg), int(end), int(cap),
                            ctypes.byref(t), _p(rows), _p(cnt), ctypes.byref(nw)))
        return int(cnt[0]), rows[:nw.value].copy()

    def batch_count(self, queries, tune=None):
        """queries: list of (edges, delta)."""
        ne, src, dst, dl = _batch_arrays(queries)
        out = np.zeros(len(queries), dtype=np.uint64)
        t = tune if tune is not None else Tuning()
        check(lib().tm_batch_count(self.handle, len(queries), _p(ne), _p(src), _p(dst), _p(dl), ctypes.byref(t),
                                   _p(out)))
        return [int(x) for x in out]

    def batch_count_noalloc(self, queries, tune=None):
        """batch_count through the entry point that allocates no device memory
        (its buffers are a static device array; at most TM_NOALLOC_MAXQ queries)."""
        ne, src, dst, dl = _batch_arrays(queries)
        out = np.zeros(max(len(queries), 1), dtype=np.uint64)
        t = tune if tune is not None else Tuning()
        check(lib().tm_batch_count_noalloc(self.handle, len(queries), _p(ne), _p(src), _p(dst), _p(dl),
                                           ctypes.byref(t), _p(out)))
        return [int(x) for x in out[:len(queries)]]

    def set_class_tuning(self, cls, tune):
        """Install (tune: a Tuning) or clear (None) the tuning of a shape class."""
        check(lib().tm_set_class_tuning(self.handle, int(cls), ctypes.byref(tune) if tune is not None else None))

    def prepare_batch(self, queries, tune=None):
        """A PreparedBatch: the batch's launches captured into a HIP graph."""
        return PreparedBatch(self, queries, tune)

    def close(self):
        if getattr(self, "handle", None):
            lib().tm_graph_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """A serving loop over a DeviceGraph (tm_server_*): submit from any
    thread, wait for results; launches are formed by continuous batching."""

    def __init__(self, graph, slots=0, max_batch=0, arena_bytes=0, tune=None):
        cfg = ServerConfig()
        cfg.slots, cfg.max_batch, cfg.arena_bytes = slots, max_batch, arena_bytes
        if tune is not None:
            cfg.tune = tune
        h = ctypes.c_void_p()
        check(lib().tm_server_create(graph.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.graph = graph

    def submit(self, edges, delta, enum_cap=-1):
        src, dst = motif_arrays(edges)
        t = ctypes.c_uint64()
        check(lib().tm_server_submit(self.handle, len(edges), _p(src), _p(dst), int(delta), int(enum_cap),
                                     ctypes.byref(t)))
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """dict(status, batch, count, rows, latency_ms, queue_ms); None on timeout."""
        r = ServerResult()
        st = lib().tm_server_wait(self.handle, ctypes.c_uint64(ticket), int(timeout_ms), ctypes.byref(r))
        if st == TM_ERR_TIMEOUT:
            return None
        check(st)
        return dict(status=r.status, batch=r.batch, count=int(r.count), rows=int(r.rows), latency_ms=r.latency_ms,
                    queue_ms=r.queue_ms)

    def take_rows(self, ticket, M, max_rows):
        rows = np.zeros((max(int(max_rows), 1), M), dtype=np.int32)
        n = ctypes.c_int64()
        check(lib().tm_server_take_rows(self.handle, ctypes.c_uint64(ticket), _p(rows), int(max_rows),
                                        ctypes.byref(n)))
        return rows[:n.value].copy()

    def release(self, ticket):
        check(lib().tm_server_release(self.handle, ctypes.c_uint64(ticket)))

    def stats(self):
        st = np.zeros(6, dtype=np.uint64)
        check(lib().tm_server_stats(self.handle, _p(st), 6))
        return dict(launches=int(st[0]), count_served=int(st[1]), enum_served=int(st[2]), max_batch=int(st[3]),
                    waiting=int(st[4]), running=int(st[5]))

    def close(self):
        if getattr(self, "handle", None):
            lib().tm_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _batch_arrays(queries):
    nq = len(queries)
    ne = np.zeros(nq, dtype=np.int32)
    src = np.zeros((nq, MAX_E), dtype=np.int32)
    dst = np.zeros((nq, MAX_E), dtype=np.int32)
    dl = np.zeros(nq, dtype=np.int32)
    for q, (edges, delta) in enumerate(queries):
        ne[q] = len(edges)
        s, d = motif_arrays(edges)
        src[q, :len(edges)] = s
        dst[q, :len(edges)] = d
        dl[q] = delta
    return ne, src, dst, dl


class PreparedBatch:
    """tm_batch_prepare / tm_batch_run / tm_batch_free."""

    def __init__(self, graph, queries, tune=None):
        ne, src, dst, dl = _batch_arrays(queries)
        t = tune if tune is not None else Tuning()
        h = ctypes.c_void_p()
        check(lib().tm_batch_prepare(graph.handle, len(queries), _p(ne), _p(src), _p(dst), _p(dl), ctypes.byref(t),
                                     ctypes.byref(h)))
        self.handle, self.nq, self.graph = h, len(queries), graph

    def run(self):
        out = np.zeros(self.nq, dtype=np.uint64)
        check(lib().tm_batch_run(self.handle, _p(out)))
        return [int(x) for x in out]

    def close(self):
        if getattr(self, "handle", None):
            lib().tm_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
