"""GPU ingest: make_directed on the device (src/graph.hpp:222-245).

The device transpose must give exactly the in-CSR of the host
make_directed (psc_make_directed: a stable counting transpose, in-lists
sorted by source, duplicates and self-loops kept), and the SCC of the graph
it builds must equal Tarjan's.
"""
import numpy as np
import pytest

from conftest import random_digraph

pytestmark = pytest.mark.gpu


def _same_graph(h, g):
    assert g.n == h.n and g.m == h.m
    for a, b in ((g.offset, h.offset), (g.edges, h.edges),
                 (g.in_offset, h.in_offset), (g.in_edges, h.in_edges)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n,m,seed,loops,dups", [
    (1, 0, 1, False, False),
    (7, 0, 2, False, False),
    (50, 400, 3, True, True),
    (1000, 20000, 4, True, True),
    (5000, 4000, 5, False, False),
    (3, 3000, 6, True, True),
])
def test_device_transpose_equals_host(psc, n, m, seed, loops, dups):
    h = random_digraph(psc, n, m, seed, self_loops=loops, dups=dups)
    d = psc.DeviceGraph.make_directed(h.n, h.offset, h.edges)
    try:
        _same_graph(h, d.to_host())
    finally:
        d.free()


def test_device_transpose_rmat_and_scc(psc, oracle):
    h = psc.gen_rmat(16, 16 << 16, seed=12)
    d = psc.DeviceGraph.make_directed(h.n, h.offset, h.edges)
    try:
        _same_graph(h, d.to_host())
        s = psc.SCC(d)
        labels = s.scc(canonical=1)
        s.close()
    finally:
        d.free()
    assert np.array_equal(np.asarray(labels).astype(np.uint32),
                          oracle.tarjan_canonical(h))


def test_device_transpose_rejects_bad_input(psc):
    off = np.array([0, 2, 3], dtype=np.uint64)
    with pytest.raises(psc.PscError):   # endpoint 5 >= n = 2
        psc.DeviceGraph.make_directed(2, off, np.array([1, 5, 0], dtype=np.uint32))
    with pytest.raises(psc.PscError):   # offsets end at 3, m = 2
        psc.DeviceGraph.make_directed(2, off, np.array([1, 0], dtype=np.uint32))
    with pytest.raises(psc.PscError):   # offsets fall
        psc.DeviceGraph.make_directed(3, np.array([0, 2, 1, 3], dtype=np.uint64),
                                      np.array([1, 0, 2], dtype=np.uint32))
