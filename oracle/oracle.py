"""ctypes wrapper over the compiled C oracle (test infrastructure only)."""
import ctypes
import os
import subprocess

import numpy as np

from . import motif as omotif

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build(force=False):
    so = os.path.join(_HERE, "libtm_oracle.so")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(
            os.path.join(_HERE, "tm_oracle.c")):
        subprocess.check_call(["make", "-s", "-C", _HERE, "libtm_oracle.so"])
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "libtm_oracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        P = ctypes.c_void_p
        L.tm_oracle_count.argtypes = [P, ctypes.c_int, P, P, P, P, P, ctypes.c_int,
                                      ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, P, P]
        L.tm_oracle_count.restype = ctypes.c_int
        L.tm_oracle_enum.argtypes = [P, ctypes.c_int, P, P, P, P, P, ctypes.c_int,
                                     ctypes.c_longlong, ctypes.c_int, ctypes.c_int, P,
                                     ctypes.c_longlong, P, P]
        L.tm_oracle_enum.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def minfo_array(edges):
    mi = omotif.predecode(edges)
    arr = np.array([[m.baseNode, m.constraintNode, m.mappedNodes, m.io, int(m.newNode)]
                    for m in mi], dtype=np.int32).reshape(-1, 5)
    return np.ascontiguousarray(arr)


def count(g, motif_edges, delta, beg=0, end=-1, nthreads=None, with_work=False):
    """Exact uint64 count of ``motif_edges`` (renamed, sorted) on graph dict ``g``
    produced by oracle.prep.preprocess."""
    L = lib()
    mi = minfo_array(motif_edges)
    cnt = ctypes.c_ulonglong(0)
    work = np.zeros(4, dtype=np.uint64)
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    rc = L.tm_oracle_count(_ptr(g["Eg"]), int(g["E"]), _ptr(g["inV"]), _ptr(g["inR"]),
                           _ptr(g["outV"]), _ptr(g["outR"]), _ptr(mi), len(motif_edges),
                           int(delta), int(beg), int(end), int(nthreads),
                           ctypes.byref(cnt), _ptr(work))
    if rc != 0:
        raise ValueError("oracle rejected the motif")
    if with_work:
        return int(cnt.value), [int(x) for x in work]
    return int(cnt.value)


def enum(g, motif_edges, delta, cap, beg=0, end=-1):
    """Returns (total_count, rows[min(total,cap), M]) in root/DFS order."""
    L = lib()
    mi = minfo_array(motif_edges)
    M = len(motif_edges)
    rows = np.zeros((max(int(cap), 1), M), dtype=np.int32)
    cnt = ctypes.c_ulonglong(0)
    rc = L.tm_oracle_enum(_ptr(g["Eg"]), int(g["E"]), _ptr(g["inV"]), _ptr(g["inR"]),
                          _ptr(g["outV"]), _ptr(g["outR"]), _ptr(mi), M, int(delta),
                          int(beg), int(end), _ptr(g["Euid"]), int(cap), _ptr(rows),
                          ctypes.byref(cnt))
    if rc != 0:
        raise ValueError("oracle rejected the motif")
    n = min(int(cnt.value), int(cap))
    return int(cnt.value), rows[:n].copy()
