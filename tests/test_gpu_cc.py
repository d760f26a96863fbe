"""GPU connectivity (src/Connectivity.h:37-119): labels must equal the
oracle's minimum-vertex-id labels of the components of the undirected graph
under the arcs, bit for bit, on every run."""
import numpy as np
import pytest

from conftest import random_digraph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,m,seed", [(1, 0, 1), (2, 1, 2), (300, 200, 3), (5000, 4000, 4),
                                      (20000, 60000, 5), (100, 5000, 6)])
def test_cc_random(psc, oracle, n, m, seed):
    g = random_digraph(psc, n, m, seed, self_loops=True, dups=True)
    print("cc random", n, m, seed, flush=True)
    assert np.array_equal(psc.connectivity(g), oracle.cc(g))


def test_cc_device_graphs(psc, oracle):
    for dg in (psc.DeviceGraph.rmat(16, 4 << 16, seed=9), psc.DeviceGraph.grid(300, 300, 0.1)):
        try:
            h = dg.to_host()
            a = psc.connectivity(dg)
            b = psc.connectivity(dg)
        finally:
            dg.free()
        expect = oracle.cc(h)
        assert np.array_equal(a, expect)
        assert np.array_equal(b, expect)   # same labels whatever the CAS order


def test_cc_long_path_and_heavy_star(psc, oracle):
    n = 200000
    # path n-1 -> n-2 -> ... -> 0: every union meets an earlier root
    path = psc.Graph.from_edges(n, np.arange(1, n), np.arange(0, n - 1))
    assert np.array_equal(psc.connectivity(path), np.zeros(n, dtype=np.uint32))
    # a hub with 100 000 arcs (the wave-shared path) plus isolated vertices
    k = 100000
    star = psc.Graph.from_edges(n, np.full(k, 7), np.arange(n - k, n))
    lab = psc.connectivity(star)
    assert np.array_equal(lab, oracle.cc(star))
    assert lab[7] == 7 and lab[n - 1] == 7 and lab[0] == 0


def test_cc_rejects_bad_input(psc):
    g = psc.Graph.from_edges(3, [0, 1], [1, 2])
    g.edges[1] = 9   # endpoint >= n, caught before any kernel indexes with it
    with pytest.raises(psc.PscError):
        psc.connectivity(g)


def test_cc_without_in_csr(psc, oracle):
    # out-CSR only: no vertex may skip its arcs, every arc is linked
    import ctypes as C
    g = random_digraph(psc, 20000, 30000, 8, self_loops=True, dups=True)
    cg = g.cgraph()
    cg.in_offset = None
    cg.in_edges = None
    out = np.empty(g.n, dtype=np.uint32)
    rc = psc.lib.psc_connectivity(C.byref(cg), out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert rc == 0
    assert np.array_equal(out, oracle.cc(g))


def test_cc_device_transpose_random(psc, oracle):
    # device graphs built by make_directed carry the in-CSR: the skip path
    for n, m, seed in ((3000, 2500, 11), (5000, 40000, 12), (200, 150, 13)):
        h = random_digraph(psc, n, m, seed, self_loops=True, dups=True)
        d = psc.DeviceGraph.make_directed(h.n, h.offset, h.edges)
        try:
            got = psc.connectivity(d)
        finally:
            d.free()
        assert np.array_equal(got, oracle.cc(h))


def _check_forest(psc, oracle, g, labels, forest):
    expect = oracle.cc(g)
    assert np.array_equal(labels, expect)
    n_comp = np.unique(expect).size
    assert forest.shape == (g.n - n_comp, 2)
    # every forest arc is an arc of g
    src = np.repeat(np.arange(g.n, dtype=np.uint64), np.diff(g.offset.astype(np.int64)))
    arcs = set(zip(src.tolist(), g.edges.tolist()))
    assert all((int(u), int(v)) in arcs for u, v in forest)
    # n - c arcs that join every component: a spanning forest (no cycles)
    f = psc.Graph.from_edges(g.n, forest[:, 0], forest[:, 1])
    assert np.array_equal(oracle.cc(f), expect)


@pytest.mark.parametrize("n,m,seed", [(1, 0, 1), (2, 1, 2), (5000, 4000, 4), (20000, 60000, 5)])
def test_spanning_forest_host(psc, oracle, n, m, seed):
    g = random_digraph(psc, n, m, seed, self_loops=True, dups=True)
    labels, forest = psc.spanning_forest(g)
    _check_forest(psc, oracle, g, labels, forest)


def test_spanning_forest_device(psc, oracle):
    dg = psc.DeviceGraph.rmat(16, 4 << 16, seed=9)
    try:
        h = dg.to_host()
        labels, forest = psc.spanning_forest(dg)
    finally:
        dg.free()
    _check_forest(psc, oracle, h, labels, forest)
