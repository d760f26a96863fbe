"""GPU LE-lists (LE_LIST::le_list, src/LeList.hpp:246-328) against the
oracle's sequential pruned BFS (seq_le_list, :76-115): for a given source
order the rows are unique, so they must agree entry for entry, whatever
beta batches the sources -- the reference's own -check compares exactly
these two (src/LeList.cpp:59-97)."""
import numpy as np
import pytest

from conftest import random_digraph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,m,seed,beta", [
    (1, 0, 1, 1.5), (2, 1, 2, 1.5), (60, 200, 3, 1.5), (3000, 9000, 4, 1.5),
    (3000, 2500, 5, 1.5), (2000, 30000, 6, 1.5), (500, 2000, 7, 1.0),
    (4000, 16000, 8, 4.0), (1500, 6000, 9, 1e6),
])
def test_le_lists_random(psc, oracle, n, m, seed, beta):
    g = random_digraph(psc, n, m, seed, self_loops=True, dups=True)
    order = psc.random_order(n, seed)
    print("le_lists random", n, m, seed, beta, flush=True)
    got = psc.le_lists(g, order, beta=beta)
    assert np.array_equal(got, oracle.le_lists(g, order))


def test_le_lists_device_graphs(psc, oracle):
    for dg in (psc.DeviceGraph.rmat(14, 8 << 14, seed=9), psc.DeviceGraph.grid(120, 120, 0.3)):
        try:
            h = dg.to_host()
            order = psc.random_order(h.n, 21)
            print("le_lists device graph", h.n, h.m, flush=True)
            a = psc.le_lists(dg, order)
            b = psc.le_lists(dg, order, beta=2.0)
        finally:
            dg.free()
        expect = oracle.le_lists(h, order)
        assert np.array_equal(a, expect)
        assert np.array_equal(b, expect)


def test_le_lists_path_and_heavy_star(psc, oracle):
    n = 400
    path = psc.Graph.from_edges(n, np.arange(0, n - 1), np.arange(1, n))
    got = psc.le_lists(path, np.arange(n, dtype=np.uint32))   # n - 1 sub-rounds
    expect = np.array([(t, s, t - s) for t in range(n) for s in range(t + 1)], dtype=np.uint32)
    assert np.array_equal(got, expect)
    # a hub whose 100 000 out-arcs are shared by a wave, reached late
    n = 120000
    k = 100000
    src = np.concatenate([np.full(k, 7), np.arange(n - k, n - 1)])
    dst = np.concatenate([np.arange(n - k, n), np.full(k - 1, 7)])
    star = psc.Graph.from_edges(n, src, dst)
    order = psc.random_order(n, 3)
    assert np.array_equal(psc.le_lists(star, order), oracle.le_lists(star, order))


def test_le_lists_rejects_bad_input(psc):
    g = psc.Graph.from_edges(3, [0, 1], [1, 2])
    with pytest.raises(psc.PscError):   # not a permutation
        psc.le_lists(g, np.array([0, 0, 2], dtype=np.uint32))
    with pytest.raises(psc.PscError):   # entry >= n
        psc.le_lists(g, np.array([0, 1, 3], dtype=np.uint32))
    with pytest.raises(psc.PscError):   # beta below 1
        psc.le_lists(g, np.arange(3, dtype=np.uint32), beta=0.5)
    g.edges[1] = 9   # endpoint >= n, caught before any kernel indexes with it
    with pytest.raises(psc.PscError):
        psc.le_lists(g, np.arange(3, dtype=np.uint32))


@pytest.mark.parametrize("n,m,seed,beta", [(3000, 9000, 4, 1.5), (2000, 30000, 6, 4.0)])
def test_le_lists_counted_bound(psc, oracle, monkeypatch, n, m, seed, beta):
    """With the growth limit at one entry, every sub-round sizes its buffers
    by the counting pass (k_le_count) instead of the degree bound; the rows
    must not change."""
    monkeypatch.setenv("PSC_LE_GROW", "1")
    g = random_digraph(psc, n, m, seed, self_loops=True, dups=True)
    order = psc.random_order(n, seed)
    print("le_lists counted bound", n, m, seed, beta, flush=True)
    got = psc.le_lists(g, order, beta=beta)
    assert np.array_equal(got, oracle.le_lists(g, order))
