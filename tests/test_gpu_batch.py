"""Batched SCC parity: every graph of a batch gets exactly the labels of a
stand-alone run (bit-exact raw labels), and canonical labels equal Tarjan.
Also pins the device batch generator against the host generators."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz")


@pytest.fixture(scope="module")
def fixture_graphs(psc):
    d = np.load(FIX)
    out = []
    for name in d["names"]:
        name = str(name)
        n = int(d[f"{name}__n"][0])
        out.append(psc.Graph.from_out_csr(n, d[f"{name}__offset"], d[f"{name}__edges"]))
    return out


def test_host_batch_equals_standalone(psc, oracle, fixture_graphs):
    graphs = list(fixture_graphs) + [psc.gen_rmat(14, 16 << 14, seed=21),
                                     psc.gen_grid(200, 150, 0.6, seed=22),
                                     psc.gen_knn3d(5000, 5, seed=23)]
    b = psc.BatchSCC(graphs)
    labels = b.split(b.run())
    for g, lab in zip(graphs, labels):
        alone = psc.scc(g)
        assert np.array_equal(lab, alone)
        assert np.array_equal(oracle.canonical(lab), oracle.tarjan_canonical(g))
    # per-graph beta, and the canonical output
    rng = np.random.default_rng(0)
    betas = rng.choice([1.2, 1.5, 2.0, 3.0], len(graphs))
    labels = b.split(b.run(beta_per_graph=betas))
    for g, lab, be in zip(graphs, labels, betas):
        assert np.array_equal(lab, psc.scc(g, beta=float(be)))
    labels = b.split(b.run(canonical=1))
    for g, lab in zip(graphs, labels):
        assert np.array_equal(lab.astype(np.uint32), oracle.tarjan_canonical(g))
    # schedule knobs do not change results
    labels2 = b.split(b.run(tau=1, local_reach=False, local_scc=False, table_init=16))
    for g, lab in zip(graphs, labels2):
        assert np.array_equal(lab, psc.scc(g))
    # host-driven rounds for every graph == single-launch remainders
    labels3 = b.split(b.run(tail_max=0))
    for g, lab in zip(graphs, labels3):
        assert np.array_equal(lab, psc.scc(g))
    b.close()


def test_device_batch_generator_matches_host(psc):
    kind = np.array([0, 1, 2, 0, 2, 1], dtype=np.uint32)
    n = np.array([1024, 900, 3000, 4096, 20000, 10000], dtype=np.uint64)
    db = psc.DeviceBatch(kind, n, seed=9)
    for g in range(len(kind)):
        sg = psc.graph_seed(9, g)
        dg = db.extract(g)
        h = dg.to_host()
        dg.free()
        if kind[g] == 0:
            sc = int(np.log2(int(n[g])))
            ref = psc.gen_rmat(sc, 16 << sc, 0.57, 0.19, 0.19, seed=sg)
        elif kind[g] == 1:
            side = int(round(np.sqrt(int(n[g]))))
            ref = psc.gen_grid(side, side, 0.6, seed=sg)
        else:
            ref = psc.gen_knn3d(int(n[g]), 5, seed=sg)
        for f in ("offset", "edges", "in_offset", "in_edges"):
            assert np.array_equal(getattr(h, f), getattr(ref, f)), (g, f)
    db.free()


def test_device_batch_equals_standalone(psc, oracle):
    kind, n = psc.batch_spec(48, 10, 15, seed=5)
    db = psc.DeviceBatch(kind, n, seed=5)
    b = psc.BatchSCC(device_batch=db)
    labels = b.split(b.run())
    for g in range(db.num_graphs):
        dg = db.extract(g)
        s = psc.SCC(dg)
        alone = s.scc()
        s.close()
        h = dg.to_host()
        dg.free()
        assert np.array_equal(labels[g], alone), g
        assert np.array_equal(oracle.canonical(labels[g]), oracle.tarjan_canonical(h)), g
    b.close()
    db.free()
