"""ORACLE — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker (or, for bench.py, as the timed
CPU baseline).  The product path never imports or links it.

ctypes bindings to liboracle.so:
  tarjan(graph)        iterative Tarjan, baselines/tarjan_scc/tarjan_scc.hpp:39-75
  canonical(labels)    min-vertex-id labels of any SCC labelling
  status(canon)        SCC_status, src/scc.cpp:13-29
  scc_ref(graph, ...)  CPU restatement of src/scc.hpp / reach.hpp /
                       hash_bag.h / resizable_table.h (scc_ref.cpp)
  reach(...)           REACH::reach, src/reach.hpp:208-243
  le_lists(graph, order)  LE-lists, seq_le_list of src/LeList.hpp:76-115
                       (le_list.c)
  cc(graph)            connected components (min vertex id per component)
                       of the undirected graph under the out-CSR: the
                       partition of src/Connectivity.h:37-119 (cc.c)
  Table / bag_roundtrip / pivot_perm   per-function hooks

Parity of these functions is pinned in tests/test_oracle.py against
networkx and scipy (independent libraries), brute-force transitive closure
on small graphs, closed-form graph families and Python sets.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


if not os.path.exists(LIB_PATH):
    build()

lib = C.CDLL(LIB_PATH)
u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)


class OParams(C.Structure):
    _fields_ = [("beta", C.c_double), ("local_reach", C.c_int32),
                ("local_scc", C.c_int32), ("tau", C.c_uint64),
                ("min_bag", C.c_uint64), ("dense_div", C.c_uint64),
                ("seed", C.c_uint64), ("table_init", C.c_uint64),
                ("threads", C.c_int32)]


class OStats(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("t_total", "t_init", "t_first", "t_multi", "t_rounds")] + \
               [(k, C.c_uint64) for k in
                ("n_trimmed", "first_scc_size", "rounds", "resizes",
                 "levels_fwd", "levels_bwd", "sub_rounds")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)


_sig("psc_oracle_tarjan", C.c_int64, C.c_uint64, u64p, u32p, u32p)
_sig("psc_oracle_cc", C.c_int, C.c_uint64, u64p, u32p, u32p)
_sig("psc_oracle_le_lists", C.c_uint64, C.c_uint64, u64p, u32p, u32p,
     C.POINTER(u32p))
_sig("psc_oracle_free", None, C.c_void_p)
_sig("psc_oracle_canonical", C.c_int, C.c_uint64, u64p, u32p)
_sig("psc_oracle_status", C.c_int, C.c_uint64, u32p, u64p, u64p)
_sig("psc_oracle_scc_create", C.c_void_p, C.c_uint64, C.c_uint64, u64p, u32p,
     u64p, u32p, C.POINTER(OParams))
_sig("psc_oracle_scc_run", C.c_int, C.c_void_p, u64p, C.POINTER(OStats))
_sig("psc_oracle_scc_destroy", None, C.c_void_p)
_sig("psc_oracle_max_threads", C.c_int)
_sig("psc_oracle_reach", C.c_uint64, C.c_uint64, C.c_uint64, u64p, u32p, u64p,
     u32p, C.c_uint32, C.c_int, C.c_int, C.c_uint64, u8p)
_sig("psc_oracle_table_new", C.c_void_p, C.c_uint64)
_sig("psc_oracle_table_free", None, C.c_void_p)
_sig("psc_oracle_table_insert", C.c_int, C.c_void_p, C.c_uint64, u32p, u32p, u8p)
_sig("psc_oracle_table_contains", C.c_int, C.c_void_p, C.c_uint32, C.c_uint32)
_sig("psc_oracle_table_values", C.c_uint64, C.c_void_p, C.c_uint32, u32p, C.c_uint64)
_sig("psc_oracle_table_size", C.c_uint64, C.c_void_p)
_sig("psc_oracle_bag_roundtrip", C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
     u32p, u32p, C.c_int)
_sig("psc_oracle_pivot_perm", C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None and a.size else None


def _arrays(g):
    off = np.ascontiguousarray(g.offset, dtype=np.uint64)
    e = np.ascontiguousarray(g.edges, dtype=np.uint32)
    ioff = np.ascontiguousarray(g.in_offset, dtype=np.uint64)
    ie = np.ascontiguousarray(g.in_edges, dtype=np.uint32)
    return off, e, ioff, ie


def cc_csr(n, offset, edges):
    off = np.ascontiguousarray(offset, dtype=np.uint64)
    e = np.ascontiguousarray(edges, dtype=np.uint32)
    out = np.empty(n, dtype=np.uint32)
    rc = lib.psc_oracle_cc(n, _p(off, u64p), _p(e, u32p), _p(out, u32p))
    if rc != 0:
        raise ValueError("psc_oracle_cc failed")
    return out


def cc(g):
    return cc_csr(g.n, g.offset, g.edges)


def le_lists_csr(n, offset, edges, order):
    """LE-lists of seq_le_list (src/LeList.hpp:76-115): (count, 3) uint32
    rows (target, source vertex, distance) sorted by (target, position of
    the source in `order`)."""
    off = np.ascontiguousarray(offset, dtype=np.uint64)
    e = np.ascontiguousarray(edges, dtype=np.uint32)
    o = np.ascontiguousarray(order, dtype=np.uint32)
    ptr = u32p()
    cnt = lib.psc_oracle_le_lists(n, _p(off, u64p), _p(e, u32p), _p(o, u32p), C.byref(ptr))
    if cnt == 2**64 - 1:
        raise ValueError("psc_oracle_le_lists: out of memory or order not a permutation")
    if cnt == 0:
        return np.zeros((0, 3), dtype=np.uint32)
    try:
        return np.ctypeslib.as_array(ptr, shape=(cnt, 3)).copy()
    finally:
        lib.psc_oracle_free(ptr)


def le_lists(g, order):
    return le_lists_csr(g.n, g.offset, g.edges, order)


def tarjan_csr(n, offset, edges):
    """Tarjan completion-order labels 1..cnt (uint32) and cnt."""
    off = np.ascontiguousarray(offset, dtype=np.uint64)
    e = np.ascontiguousarray(edges, dtype=np.uint32)
    comp = np.zeros(max(n, 1), dtype=np.uint32)
    cnt = lib.psc_oracle_tarjan(n, _p(off, u64p), _p(e, u32p), _p(comp, u32p))
    if cnt < 0:
        raise MemoryError("oracle tarjan")
    return comp[:n], int(cnt)


def tarjan(g):
    return tarjan_csr(g.n, g.offset, g.edges)


def canonical(labels):
    lab = np.ascontiguousarray(labels, dtype=np.uint64)
    out = np.zeros(max(lab.size, 1), dtype=np.uint32)
    if lib.psc_oracle_canonical(lab.size, _p(lab, u64p), _p(out, u32p)) != 0:
        raise MemoryError("oracle canonical")
    return out[:lab.size]


def tarjan_canonical(g):
    comp, _ = tarjan(g)
    return canonical(comp.astype(np.uint64))


def status(canon):
    c = np.ascontiguousarray(canon, dtype=np.uint32)
    a, b = C.c_uint64(), C.c_uint64()
    lib.psc_oracle_status(c.size, _p(c, u32p), C.byref(a), C.byref(b))
    return a.value, b.value


def params(beta=1.5, local_reach=True, local_scc=True, tau=512, min_bag=1024,
           dense_div=20, seed=1, table_init=0, threads=0):
    return OParams(beta, int(local_reach), int(local_scc), tau, min_bag,
                   dense_div, seed, table_init, threads)


class RefSCC:
    """The CPU restatement as an object: construct once, run many times
    (src/scc.cpp:57-67)."""

    def __init__(self, g, **kw):
        self._arr = _arrays(g)
        self.n = g.n
        self._prm = params(**kw)
        off, e, ioff, ie = self._arr
        self._h = lib.psc_oracle_scc_create(g.n, e.size, _p(off, u64p), _p(e, u32p),
                                            _p(ioff, u64p), _p(ie, u32p),
                                            C.byref(self._prm))

    def run(self):
        labels = np.zeros(max(self.n, 1), dtype=np.uint64)
        st = OStats()
        lib.psc_oracle_scc_run(self._h, _p(labels, u64p), C.byref(st))
        return labels[:self.n], st.as_dict()

    def close(self):
        if self._h:
            lib.psc_oracle_scc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def scc_ref(g, **kw):
    r = RefSCC(g, **kw)
    try:
        return r.run()
    finally:
        r.close()


def reach(g, src, direction=0, local=True, tau=512):
    off, e, ioff, ie = _arrays(g)
    dst = np.zeros(max(g.n, 1), dtype=np.uint8)
    levels = lib.psc_oracle_reach(g.n, e.size, _p(off, u64p), _p(e, u32p),
                                  _p(ioff, u64p), _p(ie, u32p), src, direction,
                                  int(local), tau, _p(dst, u8p))
    return dst[:g.n].astype(bool), int(levels)


class Table:
    def __init__(self, m):
        self._h = lib.psc_oracle_table_new(m)

    def insert(self, keys, vals):
        k = np.ascontiguousarray(keys, dtype=np.uint32)
        v = np.ascontiguousarray(vals, dtype=np.uint32)
        fresh = np.zeros(max(k.size, 1), dtype=np.uint8)
        full = lib.psc_oracle_table_insert(self._h, k.size, _p(k, u32p), _p(v, u32p),
                                           _p(fresh, u8p))
        return fresh[:k.size].astype(bool), bool(full)

    def contains(self, k, v):
        return bool(lib.psc_oracle_table_contains(self._h, k, v))

    def values(self, k):
        cap = 1 << 16
        out = np.zeros(cap, dtype=np.uint32)
        c = lib.psc_oracle_table_values(self._h, k, _p(out, u32p), cap)
        return out[:min(c, cap)]

    def size(self):
        return int(lib.psc_oracle_table_size(self._h))

    def __del__(self):
        if self._h:
            lib.psc_oracle_table_free(self._h)
            self._h = None


def bag_roundtrip(n, items, min_bag=1024, stamped=False):
    it = np.ascontiguousarray(items, dtype=np.uint32)
    out = np.zeros(max(it.size, 1), dtype=np.uint32)
    c = lib.psc_oracle_bag_roundtrip(n, min_bag, it.size, _p(it, u32p),
                                     _p(out, u32p), int(stamped))
    return out[:c]


def pivot_perm(N, seed, i):
    return int(lib.psc_oracle_pivot_perm(N, seed, i))


def max_threads():
    return int(lib.psc_oracle_max_threads())
