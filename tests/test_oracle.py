"""Pins the oracle before it is trusted as the checker.

* Tarjan (oracle/tarjan.c) against the committed fixtures, whose labels were
  accepted only where networkx and scipy agreed, and live against networkx
  on random digraphs and brute-force transitive closure on tiny ones;
* closed-form families (DAGs, cycles, grids, stars);
* the CPU restatement of the reference algorithm (oracle/scc_ref.cpp)
  against Tarjan over every fixture and parameter corner, including the
  forced table-overflow path;
* the per-function pieces: label table vs. a Python set, hash bag vs. the
  set of inserted ids, single reach vs. BFS closure;
* planted mistakes: corrupted labellings must be rejected.
"""
import os

import networkx as nx
import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from conftest import random_digraph

FIX = os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz")


@pytest.fixture(scope="module")
def fixtures(psc):
    d = np.load(FIX)
    out = []
    for name in d["names"]:
        name = str(name)
        n = int(d[f"{name}__n"][0])
        g = psc.Graph.from_out_csr(n, d[f"{name}__offset"], d[f"{name}__edges"])
        out.append((name, g, d[f"{name}__canon"]))
    return out


def nx_canon(g):
    G = nx.DiGraph()
    G.add_nodes_from(range(g.n))
    src = np.repeat(np.arange(g.n), np.diff(g.offset).astype(np.int64))
    G.add_edges_from(zip(src.tolist(), g.edges.tolist()))
    out = np.zeros(g.n, dtype=np.uint32)
    for comp in nx.strongly_connected_components(G):
        c = sorted(comp)
        out[c] = c[0]
    return out


def closure_canon(g):
    """SCC = mutual reachability, by boolean transitive closure (n <= 64)."""
    n = g.n
    R = np.eye(n, dtype=bool)
    for u in range(n):
        R[u, g.edges[g.offset[u]:g.offset[u + 1]]] = True
    for k in range(n):
        R |= R[:, [k]] & R[[k], :]
    M = R & R.T
    return np.array([np.flatnonzero(M[v])[0] for v in range(n)], dtype=np.uint32)


def test_fixture_count(fixtures):
    assert len(fixtures) >= 30


def test_tarjan_matches_fixtures(oracle, fixtures):
    for name, g, canon in fixtures:
        assert np.array_equal(oracle.tarjan_canonical(g), canon), name


def test_tarjan_completion_order(oracle, psc):
    # labels are the completion order 1..cnt (tarjan_scc.hpp:53-57):
    # a path 0->1->2 completes 2 first
    g = psc.Graph.from_edges(3, [0, 1], [1, 2])
    comp, cnt = oracle.tarjan(g)
    assert cnt == 3 and comp.tolist() == [3, 2, 1]


@settings(max_examples=60, deadline=None)
@given(n=st.integers(1, 60), density=st.floats(0.0, 3.0), seed=st.integers(0, 10**6))
def test_tarjan_vs_networkx_random(psc, oracle, n, density, seed):
    g = random_digraph(psc, n, int(n * density) + 1, seed, self_loops=True, dups=True)
    assert np.array_equal(oracle.tarjan_canonical(g), nx_canon(g))


@pytest.mark.parametrize("seed", range(8))
def test_tarjan_vs_closure(psc, oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 48))
    g = random_digraph(psc, n, int(n * rng.uniform(0.5, 2.5)), seed)
    assert np.array_equal(oracle.tarjan_canonical(g), closure_canon(g))


def test_closed_forms(psc, oracle):
    n = 500
    # directed cycle: one SCC
    g = psc.Graph.from_edges(n, np.arange(n), (np.arange(n) + 1) % n)
    assert np.all(oracle.tarjan_canonical(g) == 0)
    # path and random DAG: n singletons
    g = psc.Graph.from_edges(n, np.arange(n - 1), np.arange(1, n))
    assert np.array_equal(oracle.tarjan_canonical(g), np.arange(n))
    rng = np.random.default_rng(3)
    a = rng.integers(0, n, 3000)
    b = rng.integers(0, n, 3000)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    keep = lo != hi
    perm = rng.permutation(n)
    g = psc.Graph.from_edges(n, perm[lo[keep]], perm[hi[keep]])
    assert np.array_equal(oracle.tarjan_canonical(g), np.arange(n))
    # grids: all one-way -> singletons; both directions -> one SCC
    r, c = np.divmod(np.arange(900), 30)
    right = c + 1 < 30
    down = r + 1 < 30
    src = np.concatenate([np.arange(900)[right], np.arange(900)[down]])
    dst = np.concatenate([np.arange(900)[right] + 1, np.arange(900)[down] + 30])
    g = psc.Graph.from_edges(900, src, dst)
    assert np.array_equal(oracle.tarjan_canonical(g), np.arange(900))
    assert np.all(oracle.tarjan_canonical(psc.gen_grid(30, 30, 1.0, 1)) == 0)
    # deep grid: recursion depth ~n would overflow a recursive Tarjan
    g = psc.gen_grid(300, 300, 1.0, 1)
    assert np.all(oracle.tarjan_canonical(g) == 0)


def test_canonical_and_status(oracle):
    lab = np.array([7, 3, 7, 9, 3, 3], dtype=np.uint64)
    assert oracle.canonical(lab).tolist() == [0, 1, 0, 3, 1, 1]
    assert oracle.status(oracle.canonical(lab)) == (3, 3)


def test_planted_mistakes_are_caught(psc, oracle, fixtures):
    """A labelling that merges two SCCs or splits one must not pass."""
    name, g, canon = [f for f in fixtures if f[0] == "rmat_12"][0]
    lab, _ = oracle.scc_ref(g)
    good = oracle.canonical(lab)
    assert np.array_equal(good, canon)
    bad = lab.copy()
    a, b = np.unique(lab)[:2]
    bad[lab == b] = a                       # merge two SCCs
    assert not np.array_equal(oracle.canonical(bad), canon)
    big = np.bincount(canon).argmax()
    bad = lab.copy()
    v = np.flatnonzero(canon == big)[-1]
    bad[v] = g.n + 5                        # split the giant SCC
    assert not np.array_equal(oracle.canonical(bad), canon)


REF_PARAMS = [
    {},
    {"local_reach": False, "local_scc": False},
    {"tau": 1},
    {"tau": 1 << 17},
    {"beta": 1.0},
    {"beta": 3.0},
    {"table_init": 16},
    {"min_bag": 16},
    {"dense_div": 1},
    {"seed": 99},
]


@pytest.mark.parametrize("kw", REF_PARAMS, ids=lambda k: ",".join(
    f"{a}={b}" for a, b in k.items()) or "default")
def test_restatement_matches_tarjan(oracle, fixtures, kw):
    for name, g, canon in fixtures:
        lab, st_ = oracle.scc_ref(g, **kw)
        assert np.array_equal(oracle.canonical(lab), canon), (name, kw)
        if g.n:
            assert lab.max() <= g.n


def test_restatement_resize_path_runs(oracle, fixtures):
    name, g, canon = [f for f in fixtures if f[0] == "random_2000_2600"][0]
    lab, st_ = oracle.scc_ref(g, table_init=16)
    assert st_["resizes"] > 0
    assert np.array_equal(oracle.canonical(lab), canon)


def test_restatement_deterministic(oracle, fixtures):
    name, g, canon = [f for f in fixtures if f[0] == "grid_40x40_p03"][0]
    a, _ = oracle.scc_ref(g, threads=1)
    b, _ = oracle.scc_ref(g, threads=4)
    assert np.array_equal(a, b)


def test_restatement_c1(psc, oracle):
    g = psc.gen_rmat(16, 1 << 20, seed=1)
    lab, st_ = oracle.scc_ref(g)
    assert np.array_equal(oracle.canonical(lab), oracle.tarjan_canonical(g))
    assert st_["first_scc_size"] == oracle.status(oracle.tarjan_canonical(g))[1]


def test_table_semantics(oracle):
    rng = np.random.default_rng(0)
    keys = rng.integers(0, 300, 5000).astype(np.uint32)
    vals = rng.integers(0, 40, 5000).astype(np.uint32)
    t = oracle.Table(1 << 14)
    fresh, full = t.insert(keys, vals)
    assert not full
    pairs = set(zip(keys.tolist(), vals.tolist()))
    assert int(fresh.sum()) == len(pairs) == t.size()
    for k in range(300):
        want = sorted(v for (kk, v) in pairs if kk == k)
        assert sorted(t.values(k).tolist()) == want
    for k, v in list(pairs)[:200]:
        assert t.contains(k, v)
    assert not t.contains(1000, 1)
    # re-inserting reports nothing new
    fresh, _ = t.insert(keys, vals)
    assert not fresh.any()
    # a too-small table reports full
    small = oracle.Table(64)
    _, full = small.insert(np.arange(200, dtype=np.uint32), np.zeros(200, np.uint32))
    assert full


@pytest.mark.parametrize("stamped", [False, True])
def test_hash_bag_is_a_set(oracle, stamped):
    rng = np.random.default_rng(1)
    n = 50000
    items = rng.permutation(n)[:30000].astype(np.uint32)
    out = oracle.bag_roundtrip(n, items, min_bag=64, stamped=stamped)
    assert sorted(out.tolist()) == sorted(items.tolist())


@pytest.mark.parametrize("local,tau", [(False, 512), (True, 1), (True, 8), (True, 512)])
def test_reach_is_bfs_closure(psc, oracle, local, tau):
    g = psc.gen_rmat(11, 6 << 11, seed=4)
    G = nx.DiGraph()
    src = np.repeat(np.arange(g.n), np.diff(g.offset).astype(np.int64))
    G.add_edges_from(zip(src.tolist(), g.edges.tolist()))
    s = int(src[0])
    for direction, gg in ((0, G), (1, G.reverse())):
        got, _ = oracle.reach(g, s, direction, local, tau)
        want = np.zeros(g.n, dtype=bool)
        want[list(nx.descendants(gg, s)) + [s]] = True
        assert np.array_equal(got, want)


def test_pivot_perm_is_bijection(oracle):
    for N in (1, 2, 3, 7, 64, 1000, 4097):
        vals = [oracle.pivot_perm(N, 5, i) for i in range(N)]
        assert sorted(vals) == list(range(N))


# ---------------------------------------------------------------- connectivity
def _bfs_components(n, off, adj):
    """Min vertex id per connected component of the undirected graph under
    the arcs, by breadth-first search over an explicit neighbour list."""
    nb = [[] for _ in range(n)]
    for u in range(n):
        for j in range(int(off[u]), int(off[u + 1])):
            v = int(adj[j])
            nb[u].append(v)
            nb[v].append(u)
    lab = [-1] * n
    for s in range(n):
        if lab[s] >= 0:
            continue
        lab[s] = s
        q = [s]
        for x in q:
            for y in nb[x]:
                if lab[y] < 0:
                    lab[y] = s
                    q.append(y)
    return np.array(lab, dtype=np.uint32)


@pytest.mark.parametrize("n,m,seed", [(1, 0, 1), (2, 1, 2), (40, 30, 3), (200, 150, 4),
                                      (300, 900, 5)])
def test_cc_oracle_vs_bfs(psc, oracle, n, m, seed):
    g = random_digraph(psc, n, m, seed, self_loops=True, dups=True)
    assert np.array_equal(oracle.cc(g), _bfs_components(g.n, g.offset, g.edges))


def test_cc_oracle_vs_scipy(psc, oracle):
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    for n, m, seed in ((5000, 4000, 11), (20000, 30000, 12)):
        g = random_digraph(psc, n, m, seed)
        a = csr_matrix((np.ones(g.m), g.edges.astype(np.int64), g.offset.astype(np.int64)),
                       shape=(n, n))
        _, comp = connected_components(a, directed=True, connection="weak")
        first = np.full(comp.max() + 1, n, dtype=np.int64)
        np.minimum.at(first, comp, np.arange(n))
        assert np.array_equal(oracle.cc(g), first[comp].astype(np.uint32))


def test_cc_oracle_planted(oracle):
    # two triangles {0, 2, 4} and {1, 3, 5} joined by nothing, 6 isolated;
    # arcs point towards larger ids and away, so direction must not matter
    src = [4, 2, 0, 5, 1]
    dst = [2, 0, 4, 3, 5]
    n = 7
    order = np.argsort(src, kind="stable")
    s, d = np.array(src)[order], np.array(dst)[order]
    off = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(off, s + 1, 1)
    off = np.cumsum(off).astype(np.uint64)
    assert oracle.cc_csr(n, off, d).tolist() == [0, 1, 0, 1, 0, 1, 6]


# ------------------------------------------------------------ LE-lists

def _le_lists_brute(n, off, adj, order):
    """The LE-list definition, computed without pruning: a full BFS from
    every source (scipy's unweighted shortest paths), then for each target
    the sources in order whose distance is below every earlier source's."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import shortest_path
    a = csr_matrix((np.ones(len(adj)), np.asarray(adj, dtype=np.int64),
                    np.asarray(off, dtype=np.int64)), shape=(n, n))
    d = shortest_path(a, directed=True, unweighted=True, indices=np.asarray(order))
    rows = []
    for t in range(n):
        best = np.inf
        for i, s in enumerate(order):
            if d[i, t] < best:
                best = d[i, t]
                rows.append((t, s, int(best)))
    return np.array(rows, dtype=np.uint32).reshape(-1, 3)


@pytest.mark.parametrize("n,m,seed", [(1, 0, 1), (6, 9, 2), (40, 80, 3), (200, 600, 4),
                                      (300, 250, 5), (150, 3000, 6)])
def test_le_lists_oracle_vs_brute_force(psc, oracle, n, m, seed):
    g = random_digraph(psc, n, m, seed, self_loops=True, dups=True)
    order = np.random.default_rng(seed).permutation(n).astype(np.uint32)
    assert np.array_equal(oracle.le_lists(g, order), _le_lists_brute(g.n, g.offset, g.edges, order))


def test_le_lists_oracle_closed_forms(psc, oracle):
    n = 50
    path = psc.Graph.from_edges(n, np.arange(0, n - 1), np.arange(1, n))   # 0 -> 1 -> ...
    # order 0, 1, ...: every earlier source reaches t, each one step closer
    got = oracle.le_lists(path, np.arange(n, dtype=np.uint32))
    expect = [(t, s, t - s) for t in range(n) for s in range(t + 1)]
    assert np.array_equal(got, np.array(expect, dtype=np.uint32))
    # order n-1, ..., 0: the first source that reaches t is t itself
    got = oracle.le_lists(path, np.arange(n - 1, -1, -1).astype(np.uint32))
    assert np.array_equal(got, np.array([(t, t, 0) for t in range(n)], dtype=np.uint32))
    # a star with arcs out of the hub 0, hub first: one entry per vertex
    # from the hub, plus each leaf's own entry at distance 0
    star = psc.Graph.from_edges(n, np.zeros(n - 1, dtype=np.int64), np.arange(1, n))
    got = oracle.le_lists(star, np.arange(n, dtype=np.uint32))
    expect = [(0, 0, 0)] + [r for t in range(1, n) for r in ((t, 0, 1), (t, t, 0))]
    assert np.array_equal(got, np.array(expect, dtype=np.uint32))


def test_le_lists_oracle_rejects_bad_order(psc, oracle):
    g = psc.Graph.from_edges(3, [0, 1], [1, 2])
    with pytest.raises(ValueError):
        oracle.le_lists(g, np.array([0, 0, 1], dtype=np.uint32))


def test_le_lists_brute_force_catches_planted_mistake(psc, oracle):
    # a plausible slip (keeping entries whose distance only ties an earlier
    # source's) must differ from the definition on a graph with ties
    g = psc.Graph.from_edges(4, [0, 1, 0, 2], [1, 3, 2, 3])
    order = np.array([0, 1, 2, 3], dtype=np.uint32)
    good = _le_lists_brute(4, g.offset, g.edges, order)
    assert [3, 2, 1] not in good.tolist()   # d(2,3) = 1 ties d(1,3) = 1
    assert np.array_equal(oracle.le_lists(g, order), good)
