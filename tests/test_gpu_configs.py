"""GPU parity at the sizes of the benchmark workloads.

* RMAT n = 2^25, 2^29 samples (a,b,c,d) = (.57,.19,.19,.05), seed 3 (the
  headline graph, m = 528 691 493): canonical labels (host and device) equal
  Tarjan's (baselines/tarjan_scc/tarjan_scc.hpp:39-75) and raw labels equal
  the CPU restatement of src/scc.hpp:277-479.
* 2048 x 2048 directed grid (p_both 0.6, seed 2) at tau 1, 512 and 2^17:
  raw labels equal the restatement at the same tau, canonical labels equal
  Tarjan's.
* 2048 graphs in one batch (RMAT / grid / 3D 5-NN by seed, n log-uniform in
  [2^10, 2^20], seed 4): every graph's labels equal its stand-alone run bit
  for bit, and a seeded 64 of them equal Tarjan's partition.
* Serving: 10 000 requests of the same mix with n log-uniform in
  [2^10, 2^24] under a 200 GB budget: a seeded 256 of them equal their
  stand-alone runs and 32 of those equal Tarjan's partition.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_rmat25_full(psc, oracle):
    dg = psc.DeviceGraph.rmat(25, 1 << 29, 0.57, 0.19, 0.19, seed=3, permute=True)
    print(f"RMAT 2^25: n={dg.n} m={dg.m}", flush=True)
    assert dg.n == 1 << 25 and dg.m == 528691493
    s = psc.SCC(dg)
    labels = s.scc().copy()
    out = psc.DeviceArray(dg.n)
    s.scc_device(out.ptr, canonical=1)
    nscc_dev, big_dev = s.status_device(out.ptr)
    dcanon = out.numpy()
    out.free()
    s.close()
    g = dg.to_host()
    dg.free()
    print("downloaded; Tarjan", flush=True)
    expect = oracle.tarjan_canonical(g)
    assert np.array_equal(oracle.canonical(labels), expect)
    assert np.array_equal(dcanon.astype(np.uint32), expect)
    assert (nscc_dev, big_dev) == oracle.status(expect)
    print("restatement", flush=True)
    ref, rst = oracle.scc_ref(g, threads=16)
    assert np.array_equal(labels, ref), "raw labels differ from the restatement"


def test_grid2048_tau_sweep(psc, oracle):
    g = psc.gen_grid(2048, 2048, 0.6, seed=2)
    expect = oracle.tarjan_canonical(g)
    s = psc.SCC(g)
    for tau in (1, 512, 1 << 17):
        print("grid 2048^2 tau", tau, flush=True)
        labels = s.scc(tau=tau).copy()
        assert np.array_equal(oracle.canonical(labels), expect), tau
        ref, _ = oracle.scc_ref(g, tau=tau, threads=16)
        assert np.array_equal(labels, ref), tau
    s.close()


def test_batch_2048_graphs(psc, oracle):
    G = 2048
    kind, n = psc.batch_spec(G, 10, 20, seed=4)
    db = psc.DeviceBatch(kind, n, seed=4)
    print(f"batch: {G} graphs n={db.n} m={db.m}", flush=True)
    b = psc.BatchSCC(device_batch=db)
    parts = [p.copy() for p in b.split(b.run())]
    b.close()
    rng = np.random.default_rng(4)
    tarjan = set(rng.choice(G, 64, replace=False).tolist())
    for g in range(G):
        dg = db.extract(g)
        s = psc.SCC(dg)
        alone = s.scc()
        s.close()
        assert np.array_equal(parts[g], alone), g
        if g in tarjan:
            h = dg.to_host()
            assert np.array_equal(oracle.canonical(parts[g]), oracle.tarjan_canonical(h)), g
        dg.free()
        if g % 256 == 255:
            print(f"  {g + 1} graphs equal their stand-alone runs", flush=True)
    db.free()


def test_serve_10000_requests(psc, oracle):
    import psc_amd.serve as sv
    reqs = sv.poisson_requests(psc, 10000, 2000.0, lo_log2=10, hi_log2=24, seed=11)
    rng = np.random.default_rng(11)
    sample = set(rng.choice(len(reqs), 256, replace=False).tolist())
    srv = sv.Server(psc, budget_bytes=200 * 10**9, lanes=4, mode="batched",
                    keep_labels=sample)
    summary = srv.serve(reqs)
    print(summary, flush=True)
    assert summary["requests"] == 10000
    assert summary["peak_reserved_GB"] <= summary["budget_GB"]
    assert (summary["peak_device_used_GB"] - summary["device_used_before_GB"]
            <= summary["budget_GB"])
    assert summary["mean_batch"] > 1.0
    tarjan = sorted(sample)[::8]
    for r in reqs:
        if r.rid not in sample:
            assert r.labels is None
            continue
        db = psc.DeviceBatch(np.array([r.kind], np.uint32), np.array([r.n], np.uint64),
                             seeds=np.array([r.seed], np.uint64))
        dg = db.extract(0)
        db.free()
        s = psc.SCC(dg)
        alone = s.scc(beta=r.beta)
        s.close()
        assert np.array_equal(r.labels, alone), r.rid
        if r.rid in tarjan:
            h = dg.to_host()
            assert np.array_equal(oracle.canonical(r.labels), oracle.tarjan_canonical(h)), r.rid
        dg.free()
