"""Writes tests/golden/fixtures.npz: small directed graphs with their
canonical SCC labels (minimum vertex id per component).

Expected labels come from the oracle's iterative Tarjan and are accepted
only when networkx.strongly_connected_components and
scipy.sparse.csgraph.connected_components(connection='strong') -- two
independent implementations -- give the same partition.  The reference
ships no fixtures of its own (its data/ and result/ are empty).

    python tests/golden/make_fixtures.py
"""
import os
import sys

import networkx as nx
import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import load_psc  # noqa: E402
import oracle  # noqa: E402


def canon_from_groups(n, groups):
    out = np.zeros(n, dtype=np.uint32)
    for grp in groups:
        g = sorted(grp)
        out[g] = g[0]
    return out


def nx_canon(n, src, dst):
    G = nx.DiGraph()
    G.add_nodes_from(range(n))
    G.add_edges_from(zip(src.tolist(), dst.tolist()))
    return canon_from_groups(n, nx.strongly_connected_components(G))


def scipy_canon(n, src, dst):
    if n == 0:
        return np.zeros(0, dtype=np.uint32)
    A = sp.csr_matrix((np.ones(len(src)), (src, dst)), shape=(n, n))
    _, lab = connected_components(A, directed=True, connection="strong")
    groups = {}
    for v, l in enumerate(lab):
        groups.setdefault(l, []).append(v)
    return canon_from_groups(n, groups.values())


def edges_of(g):
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.offset).astype(np.int64))
    return src, g.edges.astype(np.int64)


def families(psc):
    rng = np.random.default_rng(12345)
    out = []

    def add(name, g):
        out.append((name, g))

    def from_pairs(n, pairs):
        if len(pairs) == 0:
            return psc.Graph.from_edges(n, np.zeros(0, np.int64), np.zeros(0, np.int64))
        a = np.asarray(pairs, dtype=np.int64)
        return psc.Graph.from_edges(n, a[:, 0], a[:, 1])

    add("empty_n0", from_pairs(0, []))
    add("single", from_pairs(1, []))
    add("isolated_10", from_pairs(10, []))
    add("self_loop", from_pairs(3, [(0, 0), (1, 1), (1, 2)]))
    add("dups", from_pairs(4, [(0, 1), (0, 1), (1, 0), (1, 0), (2, 3), (2, 3)]))
    add("cycle_50", from_pairs(50, [(i, (i + 1) % 50) for i in range(50)]))
    add("path_40", from_pairs(40, [(i, i + 1) for i in range(39)]))
    cyc = []
    base = 0
    for k in (1, 2, 3, 5, 8, 13):
        cyc += [(base + i, base + (i + 1) % k) for i in range(k)]
        base += k
    add("disjoint_cycles", from_pairs(base, cyc))
    add("two_cycles_joined",
        from_pairs(10, [(i, (i + 1) % 5) for i in range(5)] +
                   [(5 + i, 5 + (i + 1) % 5) for i in range(5)] + [(2, 7)]))
    add("star_out", from_pairs(20, [(0, i) for i in range(1, 20)]))
    add("star_in", from_pairs(20, [(i, 0) for i in range(1, 20)]))
    add("star_both", from_pairs(20, [(0, i) for i in range(1, 20)] +
                                [(i, 0) for i in range(1, 20, 2)]))
    # random DAG over a random topological order
    perm = rng.permutation(300)
    pairs = []
    for _ in range(1200):
        a, b = sorted(rng.integers(0, 300, 2))
        if a != b:
            pairs.append((perm[a], perm[b]))
    add("dag_300", from_pairs(300, pairs))
    # bidirected and one-way grids
    R = Cc = 12
    bi, ow = [], []
    for r in range(R):
        for c in range(Cc):
            v = r * Cc + c
            if c + 1 < Cc:
                bi += [(v, v + 1), (v + 1, v)]
                ow.append((v, v + 1))
            if r + 1 < R:
                bi += [(v, v + Cc), (v + Cc, v)]
                ow.append((v, v + Cc))
    add("grid_bidir_12", from_pairs(R * Cc, bi))
    add("grid_oneway_12", from_pairs(R * Cc, ow))
    # nested structure: chain of cliques with back edges
    pairs = []
    for blk in range(10):
        b0 = blk * 6
        pairs += [(b0 + i, b0 + (i + 1) % 6) for i in range(6)]
        if blk + 1 < 10:
            pairs.append((b0, b0 + 6))
        if blk % 3 == 2:
            pairs.append((b0 + 3, b0 - 12))
    add("clique_chain", from_pairs(60, pairs))
    for i, (n, m) in enumerate([(30, 40), (100, 150), (200, 260), (500, 900),
                                (1000, 1500), (2000, 2600), (2000, 6000),
                                (5000, 6000), (4000, 16000)]):
        src = rng.integers(0, n, m)
        dst = rng.integers(0, n, m)
        add(f"random_{n}_{m}", psc.Graph.from_edges(n, src, dst))
    for sc, seed in ((8, 1), (10, 2), (12, 3)):
        add(f"rmat_{sc}", psc.gen_rmat(sc, 8 << sc, seed=seed))
    add("rmat_10_noperm", psc.gen_rmat(10, 4 << 10, seed=9, permute=False))
    add("grid_32x32", psc.gen_grid(32, 32, 0.6, seed=2))
    add("grid_20x50", psc.gen_grid(20, 50, 0.6, seed=7))
    add("grid_40x40_p03", psc.gen_grid(40, 40, 0.3, seed=5))
    add("knn_500", psc.gen_knn3d(500, 5, seed=4))
    add("knn_3000", psc.gen_knn3d(3000, 5, seed=11))
    return out


def main():
    psc = load_psc()
    data = {}
    names = []
    for name, g in families(psc):
        src, dst = edges_of(g)
        comp, _ = oracle.tarjan(g)
        ours = oracle.canonical(comp.astype(np.uint64))
        a = nx_canon(g.n, src, dst)
        b = scipy_canon(g.n, src, dst)
        assert np.array_equal(ours, a), f"{name}: tarjan != networkx"
        assert np.array_equal(ours, b), f"{name}: tarjan != scipy"
        names.append(name)
        data[f"{name}__n"] = np.array([g.n], dtype=np.uint64)
        data[f"{name}__offset"] = g.offset
        data[f"{name}__edges"] = g.edges
        data[f"{name}__canon"] = ours
        print(f"{name:24s} n={g.n:6d} m={g.m:7d} n_scc={len(np.unique(ours))}")
    data["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **data)


if __name__ == "__main__":
    main()
