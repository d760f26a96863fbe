"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point of include/psc.h, the .bin reader/writer, the in-CSR
transposition, the seeded generators, and the `scc` command line."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "psc.h")
CLI = os.path.join(ROOT, "parallel-strong-connectivity_amd", "bin", "scc")


def header_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(psc_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(psc):
    names = header_functions(HEADER)
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(psc.lib, n)]
    assert not missing, missing


def test_nm_lists_symbols(psc):
    out = subprocess.run(["nm", "-D", "--defined-only", psc.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    for n in header_functions(HEADER):
        assert re.search(rf"\bT {n}\b", out), n


def test_params_default(psc):
    p = psc.default_params()
    assert p.beta == 1.5 and p.tau == 512 and p.min_bag == 1024
    assert p.dense_div == 20 and p.local_reach == 1 and p.local_scc == 1


def check_csr(g, sorted_lists=True, simple=True):
    n = g.n
    assert g.offset[0] == 0 and g.offset[-1] == g.m
    assert g.in_offset[0] == 0 and g.in_offset[-1] == g.m
    assert np.all(np.diff(g.offset.astype(np.int64)) >= 0)
    src = np.repeat(np.arange(n), np.diff(g.offset).astype(np.int64))
    isrc = np.repeat(np.arange(n), np.diff(g.in_offset).astype(np.int64))
    fwd = np.sort(src.astype(np.int64) * n + g.edges)
    bwd = np.sort(g.in_edges.astype(np.int64) * n + isrc)
    assert np.array_equal(fwd, bwd), "in-CSR is not the transpose"
    if sorted_lists:
        key = src.astype(np.int64) * n + g.edges
        assert np.all(np.diff(key) > 0) if simple else np.all(np.diff(key) >= 0)
        ikey = isrc.astype(np.int64) * n + g.in_edges
        assert np.all(np.diff(ikey) > 0) if simple else np.all(np.diff(ikey) >= 0)
    if simple:
        assert not np.any(src == g.edges), "self-loop"


def test_rmat_generator(psc):
    g = psc.gen_rmat(16, 1 << 20, 0.57, 0.19, 0.19, seed=1)
    check_csr(g)
    assert g.n == 1 << 16
    assert 0.85 * (1 << 20) < g.m < (1 << 20)
    g2 = psc.gen_rmat(16, 1 << 20, 0.57, 0.19, 0.19, seed=1)
    assert np.array_equal(g.edges, g2.edges) and np.array_equal(g.offset, g2.offset)
    g3 = psc.gen_rmat(16, 1 << 20, 0.57, 0.19, 0.19, seed=2)
    assert not np.array_equal(g.edges[:1000], g3.edges[:1000])
    # skew: with a=0.57 the unpermuted vertex 0 is the largest hub
    h = psc.gen_rmat(12, 1 << 16, seed=3, permute=False)
    deg = np.diff(h.offset.astype(np.int64))
    assert deg[0] == deg.max()
    hp = psc.gen_rmat(12, 1 << 16, seed=3, permute=True)
    assert sorted(deg.tolist()) == sorted(np.diff(hp.offset.astype(np.int64)).tolist())


def test_rmat_quadrant_frequencies(psc):
    # scale 1: P(edge 0->1) ~ b, (1->0) ~ c; (0,0),(1,1) are self-loops
    # and dropped; with many samples both off-diagonal edges exist
    g = psc.gen_rmat(1, 1000, 0.57, 0.19, 0.19, seed=5, permute=False)
    assert g.m == 2


def test_grid_generator(psc):
    R, Cc = 40, 30
    g = psc.gen_grid(R, Cc, 0.6, seed=2)
    check_csr(g)
    pairs = R * (Cc - 1) + (R - 1) * Cc
    src = np.repeat(np.arange(g.n), np.diff(g.offset).astype(np.int64))
    d = np.abs(src - g.edges.astype(np.int64))
    assert np.all((d == 1) | (d == Cc))
    # every 4-neighbour pair present in at least one direction
    und = set(zip(np.minimum(src, g.edges).tolist(), np.maximum(src, g.edges).tolist()))
    assert len(und) == pairs
    both = g.m - pairs
    assert 0.45 < both / pairs < 0.75
    assert psc.gen_grid(20, 20, 1.0, 1).m == 2 * (2 * 20 * 19)
    assert psc.gen_grid(20, 20, 0.0, 1).m == 2 * 20 * 19


def test_knn_generator_exact(psc):
    n, k = 700, 5
    g = psc.gen_knn3d(n, k, seed=4)
    check_csr(g)
    assert g.m == n * k
    assert np.all(np.diff(g.offset) == k)
    # brute force with the same integer coordinates
    import sys
    sys.path.insert(0, ROOT)
    pts = np.array([[knn_coord(4, i, ax) for ax in range(3)] for i in range(n)],
                   dtype=np.int64)
    d = ((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d, np.iinfo(np.int64).max)
    for i in range(0, n, 7):
        order = np.lexsort((np.arange(n), d[i]))[:k]
        assert sorted(order.tolist()) == g.edges[g.offset[i]:g.offset[i + 1]].tolist()


M64 = (1 << 64) - 1


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def knn_coord(seed, i, ax):
    """Python restatement of gen_common.hpp knn_coord/draw64."""
    stream = 0x6B6E6E0000000000 + ax
    z = (seed * 0xD1B54A32D192ED03 + stream * 0x9E3779B97F4A7C15 +
         i * 0xA0761D6478BD642F + 0x2545F4914F6CDD1D) & M64
    return _mix64(z) >> 43


def test_bin_roundtrip(psc, tmp_path):
    g = psc.gen_rmat(10, 8 << 10, seed=6)
    p_large = str(tmp_path / "g.bin")
    g.write_bin(p_large, large=True)
    h = psc.Graph.read_bin(p_large, large=True)
    for f in ("offset", "edges", "in_offset", "in_edges"):
        assert np.array_equal(getattr(g, f), getattr(h, f))
    size = os.path.getsize(p_large)
    assert size == 2 * (24 + (g.n + 1) * 8 + g.m * 4)
    p_plain = str(tmp_path / "plain.bin")
    g.write_bin(p_plain, large=False)
    h = psc.Graph.read_bin(p_plain, large=False)
    for f in ("offset", "edges", "in_offset", "in_edges"):
        assert np.array_equal(getattr(g, f), getattr(h, f))
    # header layout: n, m, sizes
    hdr = np.fromfile(p_plain, dtype=np.uint64, count=3)
    assert hdr.tolist() == [g.n, g.m, (g.n + 1) * 8 + g.m * 4 + 24]


def test_bin_sym_name_mirrors_out_csr(psc, tmp_path):
    g = psc.gen_rmat(8, 4 << 8, seed=6)
    p = str(tmp_path / "x_sym.bin")
    g.write_bin(p, large=False)
    h = psc.Graph.read_bin(p, large=False)
    assert np.array_equal(h.in_edges, g.edges)


def test_bin_rejects_bad_files(psc, tmp_path):
    p = str(tmp_path / "bad.bin")
    np.array([4, 2, 12345], dtype=np.uint64).tofile(p)
    with pytest.raises(psc.PscError):
        psc.Graph.read_bin(p)
    g = psc.gen_rmat(6, 100, seed=1)
    g.write_bin(p, large=False)
    with open(p, "ab") as f:
        f.write(b"junk")
    with pytest.raises(psc.PscError):
        psc.Graph.read_bin(p)
    with pytest.raises(psc.PscError):
        psc.Graph.read_bin(str(tmp_path / "missing.bin"))


def test_make_directed_keeps_dups_and_loops(psc):
    g = psc.Graph.from_edges(4, [0, 0, 1, 2, 2], [1, 1, 1, 3, 0])
    check_csr(g, simple=False)
    assert g.m == 5
    assert g.in_edges[g.in_offset[1]:g.in_offset[2]].tolist() == [0, 0, 1]


def test_make_directed_rejects_out_of_range(psc):
    with pytest.raises(psc.PscError):
        psc.Graph.from_out_csr(2, np.array([0, 1, 1], np.uint64), np.array([5], np.uint32))


def test_cli_usage_without_gpu():
    out = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    assert "usage: ./scc fileName" in out.stdout


def test_cli_bad_file_fails_cleanly(tmp_path):
    out = subprocess.run([CLI, str(tmp_path / "nope.bin")], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 1
    assert "cannot open" in out.stderr


def test_serve_rejects_over_budget_without_gpu(psc):
    """Admission control runs on the host: a request whose binding does not
    fit a lane's arena (sized by the library's own layout count) is refused
    before any device work."""
    import psc_amd.serve as sv
    reqs = sv.poisson_requests(psc, 3, 0.0, lo_log2=12, hi_log2=13, seed=1)
    srv = sv.Server(psc, budget_bytes=1 << 30, lanes=2, lane_bytes=1 << 20)
    with pytest.raises(RuntimeError, match="does not fit a lane"):
        srv.serve(reqs)


def test_autotune_grouping(psc):
    import psc_amd.autotune as at
    kind, n = psc.batch_spec(500, 10, 20, seed=4)
    table = {c: {"tau": 64, "beta": 2.0 if c < 15 else 1.5, "dense_div": 20}
             for c in range(10, 21)}
    groups = at.group_by_class(n, table, max_groups=4)
    idx = np.sort(np.concatenate([g[0] for g in groups]))
    assert np.array_equal(idx, np.arange(500))
    assert len(groups) <= 4
    for ix, prm, cs in groups:
        cls = [at.size_class(x) for x in n[ix]]
        assert set(cls) <= set(cs)
        assert np.allclose(prm["betas"], [2.0 if c < 15 else 1.5 for c in cls])


def test_autotune_grouping_by_family(psc):
    import psc_amd.autotune as at
    kind, n = psc.batch_spec(600, 10, 20, seed=5)
    table = {c: {"tau": 64, "beta": 1.5, "dense_div": 20} for c in range(10, 21)}
    groups = at.group_by_class(n, table, max_groups=6, family=kind)
    idx = np.sort(np.concatenate([g[0] for g in groups]))
    assert np.array_equal(idx, np.arange(600))
    assert len(groups) <= 6
    for ix, prm, cs in groups:
        assert len(set(kind[ix].tolist())) == 1          # one family per group
        assert set(at.size_class(x) for x in n[ix]) <= set(cs)
        assert len(prm["betas"]) == len(ix)


def test_batch_spec_distribution(psc):
    kind, n = psc.batch_spec(4000, 10, 20, seed=4)
    assert set(np.unique(kind).tolist()) == {0, 1, 2}
    assert n.min() >= 2 ** 10 * 0.9 and n.max() <= 2 ** 20 * 1.01
    ln = np.log2(n.astype(np.float64))
    assert abs(np.median(ln) - 15) < 0.5     # log-uniform in [10, 20]
    rm = n[kind == 0]
    assert np.all((rm & (rm - 1)) == 0)      # RMAT sizes are powers of two
    side = np.sqrt(n[kind == 1].astype(np.float64))
    assert np.all(side == np.round(side))    # grids are squares


def test_pivot_perm_matches_oracle_and_inverts(psc, oracle):
    """The device path's pivot order (element i of P_g) is the oracle's, and
    its inverse (used to key the remaining vertices by P position without
    materialising P) undoes it."""
    for N in (1, 2, 3, 17, 1000, 4097, 65536):
        for seed in (1, 3, 99):
            fwd = [psc.lib.psc_pivot_perm(N, seed, i, 0) for i in range(N)]
            assert fwd == [oracle.pivot_perm(N, seed, i) for i in range(N)]
            assert sorted(fwd) == list(range(N))
            assert [psc.lib.psc_pivot_perm(N, seed, x, 1) for x in fwd] == list(range(N))
