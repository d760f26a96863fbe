"""MI355X strongly-connected-components path, Python side.

A thin ctypes layer over ``libpsc.so`` (C ABI in ``include/psc.h``).  It
mirrors the reference's interface: ``Graph`` is ``struct Graph``
(src/graph.hpp:35-43) with ``read_bin`` for ``read_graph`` /
``read_large_graph`` (src/graph.hpp:248-315); ``SCC(graph).scc(...)`` is
``SCC::SCC`` + ``SCC::scc(labels, beta, local_reach, local_scc)``
(src/scc.hpp:79-92,277); ``status`` is ``SCC_status`` (src/scc.cpp:13-29).

The GPU library is the only compute path: when ``libpsc.so`` is missing the
import fails loudly instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsc.so")

PSC_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "out of memory",
          -4: "I/O error", -5: "over HBM budget"}


class PscError(RuntimeError):
    def __init__(self, where, code, msg):
        super().__init__(f"{where} failed ({code}, {ERRORS.get(code, '?')}): {msg}")
        self.code = code


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run parallel-strong-connectivity_amd/build.py "
        "(or __graft_entry__.build()) first")

lib = C.CDLL(LIB_PATH)

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)


class CGraph(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint64),
                ("offset", u64p), ("edges", u32p),
                ("in_offset", u64p), ("in_edges", u32p)]


class CParams(C.Structure):
    _fields_ = [("beta", C.c_double), ("local_reach", C.c_int32),
                ("local_scc", C.c_int32), ("tau", C.c_uint64),
                ("min_bag", C.c_uint64), ("dense_div", C.c_uint64),
                ("seed", C.c_uint64), ("table_init", C.c_uint64),
                ("canonical", C.c_int32), ("timing", C.c_int32),
                ("tail_max", C.c_uint64),
                ("pivot_reach", C.c_int32), ("capture", C.c_int32),
                ("vgc_batches", C.c_uint64)]


class CStats(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                ("ms_total", "ms_init", "ms_first", "ms_multi", "ms_resize",
                 "ms_finalize")] + \
               [(k, C.c_uint64) for k in
                ("n_trimmed", "first_scc_size", "rounds", "sub_rounds",
                 "levels_fwd", "levels_bwd", "resizes", "host_syncs",
                 "bytes_csr", "bytes_label", "bytes_table", "bytes_front",
                 "edges_scanned", "table_inserts")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CKernelProf(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64),
                ("total_ms", C.c_double), ("bytes", C.c_double)]


class CHostGraph(C.Structure):
    _fields_ = [("g", CGraph), ("offset", u64p), ("edges", u32p),
                ("in_offset", u64p), ("in_edges", u32p)]


class CDevGraph(C.Structure):
    _fields_ = [("g", CGraph)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


_sig("psc_params_default", None, C.POINTER(CParams))
_sig("psc_last_error", C.c_char_p)
_sig("psc_create", C.c_int, C.POINTER(CGraph), C.POINTER(C.c_void_p))
_sig("psc_create_device", C.c_int, C.POINTER(CGraph), C.POINTER(C.c_void_p))
_sig("psc_run", C.c_int, C.c_void_p, C.POINTER(CParams), u64p, C.POINTER(CStats))
_sig("psc_run_device", C.c_int, C.c_void_p, C.POINTER(CParams), C.c_void_p,
     C.POINTER(CStats))
_sig("psc_status_device", C.c_int, C.c_void_p, C.c_void_p, u64p, u64p)
_sig("psc_destroy", None, C.c_void_p)
_sig("psc_kernel_profile", C.c_int, C.c_void_p, C.POINTER(CKernelProf), C.c_int)
_sig("psc_handle_stream", C.c_void_p, C.c_void_p)
_sig("psc_device_alloc", C.c_int, C.c_size_t, C.POINTER(C.c_void_p))
_sig("psc_device_free", C.c_int, C.c_void_p)
_sig("psc_device_copy", C.c_int, C.c_void_p, C.c_void_p, C.c_size_t)
_sig("psc_host_alloc", C.c_int, C.c_size_t, C.POINTER(C.c_void_p))
_sig("psc_pivot_perm", C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int)
_sig("psc_host_free", C.c_int, C.c_void_p)
_sig("psc_host_graph_free", None, C.POINTER(CHostGraph))
_sig("psc_read_bin", C.c_int, C.c_char_p, C.c_int, C.POINTER(C.POINTER(CHostGraph)))
_sig("psc_write_bin", C.c_int, C.c_char_p, C.POINTER(CGraph), C.c_int)
_sig("psc_make_directed", C.c_int, C.c_uint64, C.c_uint64, u64p, u32p,
     C.POINTER(C.POINTER(CHostGraph)))
_sig("psc_gen_rmat", C.c_int, C.c_uint32, C.c_uint64, C.c_double, C.c_double,
     C.c_double, C.c_uint64, C.c_int, C.POINTER(C.POINTER(CHostGraph)))
_sig("psc_gen_grid", C.c_int, C.c_uint32, C.c_uint32, C.c_double, C.c_uint64,
     C.POINTER(C.POINTER(CHostGraph)))
_sig("psc_gen_knn3d", C.c_int, C.c_uint64, C.c_uint32, C.c_uint64,
     C.POINTER(C.POINTER(CHostGraph)))
_sig("psc_dev_gen_rmat", C.c_int, C.c_uint32, C.c_uint64, C.c_double,
     C.c_double, C.c_double, C.c_uint64, C.c_int, C.POINTER(C.POINTER(CDevGraph)))
_sig("psc_dev_gen_grid", C.c_int, C.c_uint32, C.c_uint32, C.c_double,
     C.c_uint64, C.POINTER(C.POINTER(CDevGraph)))
_sig("psc_dev_graph_free", None, C.POINTER(CDevGraph))
_sig("psc_connectivity", C.c_int, C.POINTER(CGraph), u32p)
_sig("psc_spanning_forest", C.c_int, C.POINTER(CGraph), u32p, u32p, C.POINTER(C.c_uint64))
_sig("psc_spanning_forest_device", C.c_int, C.POINTER(CGraph), C.c_void_p, C.c_void_p,
     C.POINTER(C.c_uint64), C.POINTER(C.c_double))
_sig("psc_le_lists", C.c_int, C.POINTER(CGraph), u32p, C.c_double, C.POINTER(u32p),
     C.POINTER(C.c_uint64))
_sig("psc_le_lists_device", C.c_int, C.POINTER(CGraph), C.c_void_p, C.c_double,
     C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.POINTER(C.c_double))
_sig("psc_le_free", None, u32p)
_sig("psc_le_free_device", C.c_int, C.c_void_p)
_sig("psc_connectivity_device", C.c_int, C.POINTER(CGraph), C.c_void_p,
     C.POINTER(C.c_double))
_sig("psc_dev_make_directed", C.c_int, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
     C.POINTER(C.POINTER(CDevGraph)))
_sig("psc_dev_graph_download", C.c_int, C.POINTER(CDevGraph),
     C.POINTER(C.POINTER(CHostGraph)))
_sig("psc_run_noalloc", C.c_int, C.POINTER(CGraph), C.POINTER(CParams), C.c_void_p,
     C.c_void_p, C.POINTER(CStats))
_sig("psc_noalloc_scratch_bytes", C.c_uint64)
_sig("psc_noalloc_scratch_peak", C.c_uint64)


class CBatchDesc(C.Structure):
    _fields_ = [("num_graphs", C.c_uint64), ("vbase", u64p), ("ebase", u64p),
                ("offset", u64p), ("edges", u32p), ("in_offset", u64p),
                ("in_edges", u32p)]


_sig("psc_batch_create", C.c_int, C.POINTER(CBatchDesc), C.c_int,
     C.POINTER(C.c_void_p))
_sig("psc_handle_graphs", C.c_uint64, C.c_void_p)
_sig("psc_batch_create_from_graphs", C.c_int, C.POINTER(CGraph), C.c_uint64,
     u64p, u64p, C.POINTER(C.c_void_p))
_sig("psc_batch_rebind", C.c_int, C.c_void_p, C.POINTER(CGraph), C.c_uint64,
     u64p, u64p)
_sig("psc_dev_batch_view", C.c_int, C.c_void_p, C.c_uint64, C.POINTER(CGraph),
     u64p, u64p)
_sig("psc_dev_batch_union", C.c_int, C.c_void_p, C.POINTER(CGraph))
_sig("psc_handle_vbase", C.c_int, C.c_void_p, u64p)
_sig("psc_batch_run", C.c_int, C.c_void_p, C.POINTER(CParams),
     C.POINTER(C.c_double), u64p, C.POINTER(CStats))
_sig("psc_batch_spec", C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_uint64,
     C.c_uint32, u32p, u64p)
_sig("psc_dev_gen_batch", C.c_int, C.c_uint64, u32p, u64p, C.c_uint64, u64p,
     C.POINTER(C.c_void_p))
_sig("psc_dev_batch_free", None, C.c_void_p)
_sig("psc_dev_batch_info", C.c_int, C.c_void_p, u64p, u64p, u64p, u32p)
_sig("psc_batch_create_dev", C.c_int, C.c_void_p, C.POINTER(C.c_void_p))
_sig("psc_lane_create", C.c_int, C.c_uint64, C.POINTER(C.c_void_p))
_sig("psc_batch_layout_bytes", C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64)
_sig("psc_handle_device_bytes", C.c_int, C.c_void_p, u64p, u64p)
_sig("psc_device_mem_info", C.c_int, u64p, u64p)
_sig("psc_pool_usage", C.c_int, u64p, u64p)
_sig("psc_pool_trim", C.c_int, C.c_uint64)
_sig("psc_dev_gen_batch_bytes", C.c_int, C.c_uint64, u32p, u64p, u64p, u64p)
_sig("psc_dev_batch_extract", C.c_int, C.c_void_p, C.c_uint64,
     C.POINTER(C.POINTER(CDevGraph)))


def _check(where, rc):
    if rc != PSC_OK:
        raise PscError(where, rc, lib.psc_last_error().decode(errors="replace"))


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None and a.size else None


class Graph:
    """Directed graph with out- and in-CSR (src/graph.hpp:35-43)."""

    def __init__(self, n, offset, edges, in_offset, in_edges):
        self.n = int(n)
        self.offset = np.ascontiguousarray(offset, dtype=np.uint64)
        self.edges = np.ascontiguousarray(edges, dtype=np.uint32)
        self.in_offset = np.ascontiguousarray(in_offset, dtype=np.uint64)
        self.in_edges = np.ascontiguousarray(in_edges, dtype=np.uint32)
        self.m = int(self.edges.size)

    def cgraph(self):
        return CGraph(self.n, self.m, _ptr(self.offset, u64p), _ptr(self.edges, u32p),
                      _ptr(self.in_offset, u64p), _ptr(self.in_edges, u32p))

    @staticmethod
    def _take(hgp):
        hg = hgp.contents
        n, m = hg.g.n, hg.g.m
        off = np.ctypeslib.as_array(hg.offset, (n + 1,)).copy()
        ioff = np.ctypeslib.as_array(hg.in_offset, (n + 1,)).copy()
        e = np.ctypeslib.as_array(hg.edges, (max(m, 1),))[:m].copy()
        ie = np.ctypeslib.as_array(hg.in_edges, (max(m, 1),))[:m].copy()
        lib.psc_host_graph_free(hgp)
        return Graph(n, off, e, ioff, ie)

    @staticmethod
    def from_out_csr(n, offset, edges):
        """In-CSR by transposition, as make_directed (src/graph.hpp:222-245)."""
        off = np.ascontiguousarray(offset, dtype=np.uint64)
        e = np.ascontiguousarray(edges, dtype=np.uint32)
        hgp = C.POINTER(CHostGraph)()
        _check("psc_make_directed",
               lib.psc_make_directed(n, e.size, _ptr(off, u64p), _ptr(e, u32p),
                                     C.byref(hgp)))
        return Graph._take(hgp)

    @staticmethod
    def from_edges(n, src, dst):
        """Edge list -> CSR (kept as given: duplicates and self-loops stay)."""
        src = np.asarray(src, dtype=np.uint64)
        dst = np.asarray(dst, dtype=np.uint32)
        order = np.lexsort((dst, src))
        src, dst = src[order], dst[order]
        off = np.zeros(n + 1, dtype=np.uint64)
        np.add.at(off, src.astype(np.int64) + 1, 1)
        off = np.cumsum(off).astype(np.uint64)
        return Graph.from_out_csr(n, off, dst)

    @staticmethod
    def read_bin(path, large=False):
        hgp = C.POINTER(CHostGraph)()
        _check("psc_read_bin", lib.psc_read_bin(os.fsencode(path), int(large),
                                                C.byref(hgp)))
        return Graph._take(hgp)

    def write_bin(self, path, large=True):
        g = self.cgraph()
        _check("psc_write_bin", lib.psc_write_bin(os.fsencode(path), C.byref(g),
                                                  int(large)))


def connectivity(g):
    """Connected components of the undirected graph under g's out-CSR
    (Connectivity, src/Connectivity.h:37-119): labels[v] = the minimum
    vertex id of v's component.  g is a Graph (host arrays) or a
    DeviceGraph; for a DeviceGraph the kernels' time is in connectivity.ms."""
    out = np.empty(g.n, dtype=np.uint32)
    cg = g.cgraph()
    if isinstance(g, DeviceGraph):
        d = C.c_void_p()
        _check("psc_device_alloc", lib.psc_device_alloc(max(g.n * 4, 4), C.byref(d)))
        try:
            ms = C.c_double()
            _check("psc_connectivity_device",
                   lib.psc_connectivity_device(C.byref(cg), d, C.byref(ms)))
            if g.n:
                _check("psc_device_copy", lib.psc_device_copy(out.ctypes.data, d, g.n * 4))
            connectivity.ms = ms.value
        finally:
            lib.psc_device_free(d)
    else:
        _check("psc_connectivity", lib.psc_connectivity(C.byref(cg), _ptr(out, u32p)))
    return out


def spanning_forest(g):
    """Connectivity(G, beta, pred, spanning_forest=True)
    (src/Connectivity.h:37-40,100-118): (labels, forest) with labels as in
    connectivity(g) and forest a (n - components, 2) uint32 array of arcs
    (u, v) of g that span its components."""
    labels = np.empty(g.n, dtype=np.uint32)
    forest = np.empty((max(g.n - 1, 1), 2), dtype=np.uint32)
    cnt = C.c_uint64()
    cg = g.cgraph()
    if isinstance(g, DeviceGraph):
        d, df = C.c_void_p(), C.c_void_p()
        _check("psc_device_alloc", lib.psc_device_alloc(max(g.n * 4, 4), C.byref(d)))
        try:
            _check("psc_device_alloc",
                   lib.psc_device_alloc(max((g.n - 1) * 8, 8), C.byref(df)))
            ms = C.c_double()
            _check("psc_spanning_forest_device",
                   lib.psc_spanning_forest_device(C.byref(cg), d, df, C.byref(cnt),
                                                  C.byref(ms)))
            if g.n:
                _check("psc_device_copy", lib.psc_device_copy(labels.ctypes.data, d, g.n * 4))
            if cnt.value:
                _check("psc_device_copy",
                       lib.psc_device_copy(forest.ctypes.data, df, cnt.value * 8))
            spanning_forest.ms = ms.value
        finally:
            lib.psc_device_free(d)
            if df.value:
                lib.psc_device_free(df)
    else:
        _check("psc_spanning_forest",
               lib.psc_spanning_forest(C.byref(cg), _ptr(labels, u32p), _ptr(forest, u32p),
                                       C.byref(cnt)))
    return labels, forest[:cnt.value].copy()


def random_order(n, seed=1):
    """A seeded random permutation of 0..n-1 (the role of the reference's
    parlay::random_permutation for LE_LIST::centers, src/LeList.hpp:64)."""
    return np.random.default_rng(seed).permutation(n).astype(np.uint32)


def le_lists(g, order=None, beta=1.5, seed=1):
    """LE-lists (LE_LIST::le_list, src/LeList.hpp:246-328) of g's out-arcs
    for the source order `order` (default: random_order(g.n, seed)): a
    (count, 3) uint32 array of rows (target, source vertex, distance),
    sorted by (target, position of the source in order).  g is a Graph or
    a DeviceGraph; for a DeviceGraph the kernels' time is in le_lists.ms."""
    if order is None:
        order = random_order(g.n, seed)
    order = np.ascontiguousarray(order, dtype=np.uint32)
    if order.shape != (g.n,):
        raise PscError(f"le_lists: order must hold n = {g.n} entries")
    cg = g.cgraph()
    cnt = C.c_uint64()
    if isinstance(g, DeviceGraph):
        d_ord, d_rows = C.c_void_p(), C.c_void_p()
        _check("psc_device_alloc", lib.psc_device_alloc(max(g.n * 4, 4), C.byref(d_ord)))
        try:
            if g.n:
                _check("psc_device_copy",
                       lib.psc_device_copy(d_ord, order.ctypes.data, g.n * 4))
            ms = C.c_double()
            _check("psc_le_lists_device",
                   lib.psc_le_lists_device(C.byref(cg), d_ord, float(beta), C.byref(d_rows),
                                           C.byref(cnt), C.byref(ms)))
            out = np.empty((cnt.value, 3), dtype=np.uint32)
            try:
                if cnt.value:
                    _check("psc_device_copy",
                           lib.psc_device_copy(out.ctypes.data, d_rows, cnt.value * 12))
            finally:
                lib.psc_le_free_device(d_rows)
            le_lists.ms = ms.value
        finally:
            lib.psc_device_free(d_ord)
        return out
    rows = u32p()
    _check("psc_le_lists", lib.psc_le_lists(C.byref(cg), _ptr(order, u32p), float(beta),
                                            C.byref(rows), C.byref(cnt)))
    try:
        if cnt.value == 0:
            return np.zeros((0, 3), dtype=np.uint32)
        return np.ctypeslib.as_array(rows, shape=(cnt.value, 3)).copy()
    finally:
        lib.psc_le_free(rows)


def gen_rmat(scale, samples, a=0.57, b=0.19, c=0.19, seed=1, permute=True):
    hgp = C.POINTER(CHostGraph)()
    _check("psc_gen_rmat", lib.psc_gen_rmat(scale, samples, a, b, c, seed,
                                            int(permute), C.byref(hgp)))
    return Graph._take(hgp)


def gen_grid(rows, cols, p_both=0.6, seed=2):
    hgp = C.POINTER(CHostGraph)()
    _check("psc_gen_grid", lib.psc_gen_grid(rows, cols, p_both, seed, C.byref(hgp)))
    return Graph._take(hgp)


def gen_knn3d(n, k=5, seed=4):
    hgp = C.POINTER(CHostGraph)()
    _check("psc_gen_knn3d", lib.psc_gen_knn3d(n, k, seed, C.byref(hgp)))
    return Graph._take(hgp)


class DeviceGraph:
    """A graph generated straight into HBM by the device generators."""

    def __init__(self, dgp):
        self._p = dgp
        self.n = dgp.contents.g.n
        self.m = dgp.contents.g.m

    @staticmethod
    def rmat(scale, samples, a=0.57, b=0.19, c=0.19, seed=1, permute=True):
        dgp = C.POINTER(CDevGraph)()
        _check("psc_dev_gen_rmat", lib.psc_dev_gen_rmat(
            scale, samples, a, b, c, seed, int(permute), C.byref(dgp)))
        return DeviceGraph(dgp)

    @staticmethod
    def grid(rows, cols, p_both=0.6, seed=2):
        dgp = C.POINTER(CDevGraph)()
        _check("psc_dev_gen_grid", lib.psc_dev_gen_grid(rows, cols, p_both, seed,
                                                        C.byref(dgp)))
        return DeviceGraph(dgp)

    @staticmethod
    def make_directed(n, offset, edges):
        """In-CSR built on the device (make_directed, src/graph.hpp:222-245):
        the out-CSR goes to device memory as given, the transpose runs there."""
        off = np.ascontiguousarray(offset, dtype=np.uint64)
        e = np.ascontiguousarray(edges, dtype=np.uint32)
        d_off, d_e = C.c_void_p(), C.c_void_p()
        _check("psc_device_alloc", lib.psc_device_alloc(max(off.nbytes, 8), C.byref(d_off)))
        try:
            _check("psc_device_alloc", lib.psc_device_alloc(max(e.nbytes, 4), C.byref(d_e)))
            try:
                _check("psc_device_copy", lib.psc_device_copy(d_off, off.ctypes.data, off.nbytes))
                if e.size:
                    _check("psc_device_copy", lib.psc_device_copy(d_e, e.ctypes.data, e.nbytes))
                dgp = C.POINTER(CDevGraph)()
                _check("psc_dev_make_directed",
                       lib.psc_dev_make_directed(int(n), int(e.size), d_off, d_e,
                                                 C.byref(dgp)))
                return DeviceGraph(dgp)
            finally:
                lib.psc_device_free(d_e)
        finally:
            lib.psc_device_free(d_off)

    def cgraph(self):
        return self._p.contents.g

    def to_host(self):
        hgp = C.POINTER(CHostGraph)()
        _check("psc_dev_graph_download",
               lib.psc_dev_graph_download(self._p, C.byref(hgp)))
        return Graph._take(hgp)

    def free(self):
        if self._p:
            lib.psc_dev_graph_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """A uint64 buffer in device memory (plumbing for psc_run_device)."""

    def __init__(self, count):
        self.count = int(count)
        p = C.c_void_p()
        _check("psc_device_alloc", lib.psc_device_alloc(max(self.count, 1) * 8,
                                                        C.byref(p)))
        self.ptr = p.value

    def numpy(self):
        out = np.empty(self.count, dtype=np.uint64)
        if self.count:
            _check("psc_device_copy", lib.psc_device_copy(
                out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.count * 8))
        return out

    def free(self):
        if self.ptr:
            lib.psc_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedLabels:
    """A growable page-locked uint64 host buffer; labels downloaded into it
    move at full link speed.  view(n) is valid until the next grow."""

    def __init__(self, count=0):
        self.ptr, self.cap = None, 0
        self.grow(count)

    def grow(self, count):
        count = int(count)
        if count <= self.cap:
            return
        self.free()
        cap = max(count, 1 << 20)
        p = C.c_void_p()
        _check("psc_host_alloc", lib.psc_host_alloc(cap * 8, C.byref(p)))
        self.ptr, self.cap = p.value, cap

    def view(self, count):
        self.grow(count)
        buf = (C.c_uint64 * self.cap).from_address(self.ptr)
        return np.ctypeslib.as_array(buf)[:count]

    def free(self):
        if self.ptr:
            lib.psc_host_free(C.c_void_p(self.ptr))
            self.ptr, self.cap = None, 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def default_params(**kw):
    p = CParams()
    lib.psc_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


class SCC:
    """SCC(Graph&) + scc(labels, beta, local_reach, local_scc)
    (src/scc.hpp:79-92,277).  The graph is uploaded once; every scc() call
    reuses the device working set."""

    def __init__(self, graph):
        self.graph = graph
        self.n = graph.n
        self._g = graph.cgraph()
        h = C.c_void_p()
        if isinstance(graph, DeviceGraph):
            _check("psc_create_device", lib.psc_create_device(C.byref(self._g),
                                                              C.byref(h)))
        else:
            _check("psc_create", lib.psc_create(C.byref(self._g), C.byref(h)))
        self._h = h
        self.stats = None

    def scc(self, labels=None, beta=1.5, local_reach=True, local_scc=True, **kw):
        p = default_params(beta=beta, local_reach=int(local_reach),
                           local_scc=int(local_scc), **kw)
        if labels is None:
            labels = np.empty(self.n, dtype=np.uint64)
        st = CStats()
        _check("psc_run", lib.psc_run(self._h, C.byref(p), _ptr(labels, u64p),
                                      C.byref(st)))
        self.stats = st.as_dict()
        return labels

    def scc_device(self, d_labels, beta=1.5, local_reach=True, local_scc=True, **kw):
        """Labels stay in HBM at device address d_labels (n uint64)."""
        p = default_params(beta=beta, local_reach=int(local_reach),
                           local_scc=int(local_scc), **kw)
        st = CStats()
        _check("psc_run_device", lib.psc_run_device(self._h, C.byref(p),
                                                    C.c_void_p(d_labels), C.byref(st)))
        self.stats = st.as_dict()
        return self.stats

    def run(self, beta=1.5, local_reach=True, local_scc=True, **kw):
        """One SCC pass with labels left in the handle's device buffer."""
        p = default_params(beta=beta, local_reach=int(local_reach),
                           local_scc=int(local_scc), **kw)
        st = CStats()
        _check("psc_run", lib.psc_run(self._h, C.byref(p), None, C.byref(st)))
        self.stats = st.as_dict()
        return self.stats

    def kernel_profile(self):
        """Per-kernel launches / time / algorithmic bytes of the last run
        made with timing=2."""
        buf = (CKernelProf * 16)()
        k = lib.psc_kernel_profile(self._h, buf, 16)
        if k < 0:
            _check("psc_kernel_profile", k)
        return [{"name": buf[i].name.decode(), "launches": buf[i].launches,
                 "total_ms": buf[i].total_ms, "bytes": buf[i].bytes}
                for i in range(k)]

    def status_device(self, d_labels):
        a, b = C.c_uint64(), C.c_uint64()
        _check("psc_status_device", lib.psc_status_device(
            self._h, C.c_void_p(d_labels), C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self._h:
            lib.psc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scc_noalloc(graph, d_labels, stream=None, beta=1.5, local_reach=True, local_scc=True,
                **kw):
    """psc_run_noalloc: SCC of a device graph (DeviceGraph or CGraph of
    device pointers) into d_labels (device address, n uint64) with every
    working buffer from the library's fixed scratch array; returns the
    stats."""
    g = graph.cgraph() if hasattr(graph, "cgraph") else graph
    p = default_params(beta=beta, local_reach=int(local_reach),
                       local_scc=int(local_scc), **kw)
    st = CStats()
    _check("psc_run_noalloc", lib.psc_run_noalloc(C.byref(g), C.byref(p),
                                                  C.c_void_p(d_labels),
                                                  C.c_void_p(stream), C.byref(st)))
    return st.as_dict()


def scc(graph, **kw):
    """One-shot SCC labels (uint64, equal iff same SCC)."""
    s = SCC(graph)
    try:
        return s.scc(**kw)
    finally:
        s.close()


def status(labels):
    """SCC_status (src/scc.cpp:13-29) on host labels: (n_scc, largest)."""
    labels = np.asarray(labels)
    if labels.size == 0:
        return 0, 0
    _, counts = np.unique(labels, return_counts=True)
    return int(counts.size), int(counts.max())


M64 = (1 << 64) - 1


def graph_seed(seed, g):
    """Seed of graph g of a generated batch: mix(seed * phi + g + 1)."""
    z = (seed * 0x9E3779B97F4A7C15 + g + 1) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def batch_spec(num_graphs, lo_log2=10, hi_log2=20, seed=4, kinds=(0, 1, 2)):
    """Seeded mix of graph kinds (0 RMAT, 1 grid, 2 3D k-NN) and sizes."""
    kind = np.zeros(num_graphs, dtype=np.uint32)
    n = np.zeros(num_graphs, dtype=np.uint64)
    mask = sum(1 << k for k in kinds)
    _check("psc_batch_spec", lib.psc_batch_spec(
        num_graphs, lo_log2, hi_log2, seed, mask, _ptr(kind, u32p), _ptr(n, u64p)))
    return kind, n


class DeviceBatch:
    """A batch of graphs generated straight into HBM as one disjoint union."""

    def __init__(self, kind, n, seed=4, seeds=None):
        kind = np.ascontiguousarray(kind, dtype=np.uint32)
        n = np.ascontiguousarray(n, dtype=np.uint64)
        sd = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.uint64)
        p = C.c_void_p()
        _check("psc_dev_gen_batch", lib.psc_dev_gen_batch(
            kind.size, _ptr(kind, u32p), _ptr(n, u64p), seed,
            _ptr(sd, u64p) if sd is not None else None, C.byref(p)))
        self._p = p
        self.seed = seed
        G = kind.size
        self.num_graphs = G
        self.vbase = np.zeros(G + 1, dtype=np.uint64)
        self.ebase = np.zeros(G + 1, dtype=np.uint64)
        self.kind = np.zeros(G, dtype=np.uint32)
        cnt = C.c_uint64()
        _check("psc_dev_batch_info", lib.psc_dev_batch_info(
            self._p, C.byref(cnt), _ptr(self.vbase, u64p), _ptr(self.ebase, u64p),
            _ptr(self.kind, u32p)))
        self.n = int(self.vbase[-1])
        self.m = int(self.ebase[-1])

    def union_graph(self):
        """The batch's union as a CGraph of device pointers."""
        g = CGraph()
        _check("psc_dev_batch_union", lib.psc_dev_batch_union(self._p, C.byref(g)))
        return g

    def view(self, g):
        """(CGraph view of graph g inside the union, id base, edge base)."""
        cg, vb, eb = CGraph(), C.c_uint64(), C.c_uint64()
        _check("psc_dev_batch_view", lib.psc_dev_batch_view(
            self._p, g, C.byref(cg), C.byref(vb), C.byref(eb)))
        return cg, vb.value, eb.value

    def extract(self, g):
        dgp = C.POINTER(CDevGraph)()
        _check("psc_dev_batch_extract", lib.psc_dev_batch_extract(self._p, g,
                                                                  C.byref(dgp)))
        return DeviceGraph(dgp)

    def free(self):
        if self._p:
            lib.psc_dev_batch_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BatchSCC:
    """SCC of many graphs in one handle; labels per graph equal a
    stand-alone SCC(graph).scc() with the same beta and seed."""

    def __init__(self, graphs=None, device_batch=None, device_graphs=None):
        h = C.c_void_p()
        if device_graphs is not None:
            arr = (CGraph * len(device_graphs))(*[d.cgraph() for d in device_graphs])
            _check("psc_batch_create_from_graphs", lib.psc_batch_create_from_graphs(
                arr, len(device_graphs), None, None, C.byref(h)))
            self._keep = None
        elif device_batch is not None:
            _check("psc_batch_create_dev", lib.psc_batch_create_dev(
                device_batch._p, C.byref(h)))
            self._keep = device_batch
        else:
            G = len(graphs)
            vbase = np.zeros(G + 1, dtype=np.uint64)
            ebase = np.zeros(G + 1, dtype=np.uint64)
            vbase[1:] = np.cumsum([g.n for g in graphs])
            ebase[1:] = np.cumsum([g.m for g in graphs])
            cat = lambda xs, dt: np.ascontiguousarray(
                np.concatenate(xs) if xs else np.zeros(0), dtype=dt)
            off = cat([g.offset for g in graphs], np.uint64)
            ioff = cat([g.in_offset for g in graphs], np.uint64)
            e = cat([g.edges for g in graphs], np.uint32)
            ie = cat([g.in_edges for g in graphs], np.uint32)
            d = CBatchDesc(G, _ptr(vbase, u64p), _ptr(ebase, u64p), _ptr(off, u64p),
                           _ptr(e, u32p), _ptr(ioff, u64p), _ptr(ie, u32p))
            _check("psc_batch_create", lib.psc_batch_create(C.byref(d), 0, C.byref(h)))
            self._keep = None
        self._h = h
        self._refresh()
        self.stats = None

    @classmethod
    def _wrap(cls, h):
        """Adopt a handle created through the C ABI."""
        obj = cls.__new__(cls)
        obj._h = h
        obj._keep = None
        obj.stats = None
        obj._refresh()
        return obj

    def _refresh(self):
        G = lib.psc_handle_graphs(self._h)
        self.num_graphs = int(G)
        self.vbase = np.zeros(G + 1, dtype=np.uint64)
        _check("psc_handle_vbase", lib.psc_handle_vbase(self._h, _ptr(self.vbase, u64p)))
        self.n = int(self.vbase[-1])

    def run(self, beta=1.5, beta_per_graph=None, labels=True, out=None, **kw):
        """labels: download the labels (into `out` when given, e.g. a
        PinnedLabels view of n entries)."""
        p = default_params(beta=beta, **kw)
        if labels and out is None:
            out = np.empty(self.n, dtype=np.uint64)
        if labels and (out.dtype != np.uint64 or out.size != self.n or
                       not out.flags.c_contiguous):
            raise ValueError("out must be a contiguous uint64 array of n entries")
        bp = None
        if beta_per_graph is not None:
            bpg = np.ascontiguousarray(beta_per_graph, dtype=np.float64)
            bp = bpg.ctypes.data_as(C.POINTER(C.c_double))
        st = CStats()
        _check("psc_batch_run", lib.psc_batch_run(
            self._h, C.byref(p), bp, _ptr(out, u64p) if labels else None, C.byref(st)))
        self.stats = st.as_dict()
        return out

    def split(self, labels):
        return [labels[self.vbase[g]:self.vbase[g + 1]] for g in range(self.num_graphs)]

    kernel_profile = SCC.kernel_profile

    def close(self):
        if self._h:
            lib.psc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
