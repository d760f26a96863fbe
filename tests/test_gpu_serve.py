"""Serving: every request's labels are the same alone and in the mix.

A burst of requests is served by the continuous-batching server (several
lanes, an HBM budget that forces admission to wait) and by the plain
one-request-per-launch server; each request's labels must equal a
stand-alone SCC of its graph bit for bit and match Tarjan."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host_graph(psc, r):
    if r.kind == 0:
        sc = int(np.log2(r.n))
        return psc.gen_rmat(sc, 16 << sc, 0.57, 0.19, 0.19, seed=r.seed)
    if r.kind == 1:
        side = int(round(np.sqrt(r.n)))
        return psc.gen_grid(side, side, 0.6, seed=r.seed)
    return psc.gen_knn3d(r.n, 5, seed=r.seed)


@pytest.mark.parametrize("mode,rate,capture", [("batched", 0.0, 0), ("batched", 400.0, 0),
                                               ("plain", 0.0, 0), ("batched", 0.0, 1),
                                               ("batched", 400.0, 1), ("plain", 0.0, 1)])
def test_serve_alone_equals_mixed(psc, oracle, mode, rate, capture):
    import psc_amd.serve as sv
    reqs = sv.poisson_requests(psc, 40, rate, lo_log2=10, hi_log2=15, seed=3)
    for i, r in enumerate(reqs):
        r.beta = (1.5, 2.0, 3.0)[i % 3]
    # three lanes of 24 MiB and 96 MiB for the graph pool: admission waits
    # for memory, and a lane's arena holds only part of the ready requests
    srv = sv.Server(psc, budget_bytes=168 << 20, lanes=3, lane_bytes=24 << 20, mode=mode,
                    keep_labels=True, target_batch=8, max_wait_s=0.05, capture=capture)
    summary = srv.serve(reqs)
    print(summary, flush=True)
    assert summary["peak_reserved_GB"] <= summary["budget_GB"]
    # the driver's figure also counts the runtime's own memory (queues, code
    # objects, the pool's granularity), up to a few hundred MiB in a fresh
    # process: small next to the serving budget, not next to this one
    assert (summary["peak_device_used_GB"] - summary["device_used_before_GB"]
            <= summary["budget_GB"] + 1.0)
    if mode == "batched" and rate == 0.0:
        assert summary["mean_batch"] > 1.5
    if mode == "plain":
        assert summary["mean_batch"] == 1.0
    for r in reqs:
        g = host_graph(psc, r)
        alone = psc.scc(g, beta=r.beta)
        assert np.array_equal(r.labels, alone), r.rid
        assert np.array_equal(oracle.canonical(r.labels), oracle.tarjan_canonical(g)), r.rid
