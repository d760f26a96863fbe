#!/usr/bin/env python3
"""SCC throughput on one MI355X: RMAT n=2^25, 2^29 edge samples.

A step is one full SCC pass (trim, first-SCC reaches, multi-search rounds,
partition, finalize) over the device-resident graph, i.e. the window the
reference times around SCC::scc (src/scc.cpp:62-67).  The graph is
generated on the device (seed 3, (a,b,c,d) = (.57,.19,.19,.05), ids
scrambled, self-loops and duplicates removed) and is resident in HBM before
the timed region.  Defaults follow the reference experiments: beta 1.5,
tau 512, both VGC flags (scripts/Experiment1.py:29).

Prints one JSON line: Gedges/s (m / t_scc), the roofline of the dominant
kernel (algorithmic bytes per launch / average launch time, HIP events on
the handle's stream), and the CPU baseline: the OpenMP restatement of the
reference algorithm (oracle/scc_ref.cpp) on this host's cores, on the same
graph, whose raw labels must equal the GPU's bit for bit.

    python bench.py [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_pmc(kernel, workload):
    """HBM traffic per launch from committed rocprofv3 --pmc summaries."""
    path = os.path.join(ROOT, "profiles", f"{workload}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    if not k:
        return None
    return k.get("hbm_bytes_per_launch")


def bench_grid(args, psc):
    """C2: 2048x2048 directed grid (p_both 0.6, seed 2), tau swept
    2^0..2^17 as scripts/Experiment4.py:13 does; labels checked vs Tarjan."""
    import numpy as np
    import oracle
    side = args.grid_side
    dg = psc.DeviceGraph.grid(side, side, 0.6, seed=2)
    s = psc.SCC(dg)
    rows = []
    taus = [1 << k for k in range(18)] if args.taus is None else args.taus
    for tau in taus:
        prm = dict(beta=1.5, local_reach=True, local_scc=True, tau=tau, timing=1)
        if args.deep_vgc:   # a wave may spend its whole tau budget serially
            prm["vgc_batches"] = max(64, 4 * tau // 64)
        s.run(**prm)
        ts = []
        for _ in range(args.steps):
            t = time.perf_counter()
            st = s.run(**prm)
            ts.append(time.perf_counter() - t)
        ms = float(np.median(ts)) * 1e3
        rows.append({"tau": tau, "ms": round(ms, 3),
                     "vgc_batches": prm.get("vgc_batches", 64),
                     "gedges_s": round(dg.m / (ms * 1e-3) / 1e9, 4),
                     "levels": [st["levels_fwd"], st["levels_bwd"]],
                     "ms_first": round(st["ms_first"], 3),
                     "ms_multi": round(st["ms_multi"], 3)})
        log(f"tau {tau:7d}: {ms:9.3f} ms  levels {st['levels_fwd']}/{st['levels_bwd']} "
            f"first {st['ms_first']:.2f} multi {st['ms_multi']:.2f} init {st['ms_init']:.2f} "
            f"rounds {st['rounds']} sub {st['sub_rounds']} syncs {st['host_syncs']}")
    labels = s.scc(beta=1.5, tau=512)
    g = dg.to_host()
    ok = bool(np.array_equal(oracle.canonical(labels), oracle.tarjan_canonical(g)))
    best = min(rows, key=lambda r: r["ms"])
    print(json.dumps({"metric": "SCC time on the 2048x2048 directed grid vs tau",
                      "value": best["ms"], "unit": "ms", "higher_is_better": False,
                      "best_tau": best["tau"], "n": dg.n, "m": dg.m,
                      "labels_equal_tarjan": ok, "sweep": rows}), flush=True)
    s.close()
    dg.free()


def bench_batch(args, psc):
    """C4: 2048 graphs (RMAT avg out-degree 16 / directed grid / 3D 5-NN,
    chosen by seed), n log-uniform in [2^10, 2^20]; batched graphs/s against
    the one-graph-per-launch path on the same GPU, identical labels."""
    import numpy as np
    import oracle
    import psc_amd.autotune as at
    G = args.graphs
    kind, n = psc.batch_spec(G, 10, args.hi_log2, seed=4)
    seeds = np.array([psc.graph_seed(4, g) for g in range(G)], dtype=np.uint64)
    table = at.load_table()
    groups = at.group_by_class(n, table, max_groups=args.lanes, family=kind)
    t = time.perf_counter()
    dbs = [psc.DeviceBatch(kind[ix], n[ix], seeds=seeds[ix]) for ix, _, _ in groups]
    gen_s = time.perf_counter() - t
    n_tot = sum(db.n for db in dbs)
    m_tot = sum(db.m for db in dbs)
    log(f"generated {G} graphs in {len(groups)} class groups: n={n_tot} m={m_tot} "
        f"in {gen_s:.1f}s")
    bss = [psc.BatchSCC(device_batch=db) for db in dbs]
    prms = [p for _, p, _ in groups]
    if args.tail_max is not None:
        for prm in prms:
            prm["tail_max"] = args.tail_max
    at.run_groups(bss, prms)                       # warm-up (allocations)
    walls = [at.run_groups(bss, prms)[0] for _ in range(args.steps)]
    wall = float(np.median(walls))
    _, labs = at.run_groups(bss, prms, labels=True, timing=1)
    log(f"batched: {wall * 1e3:.2f} ms for {G} graphs ({G / wall:.1f} graphs/s)")
    if args.levels:   # per-level and remainder-round reports on stderr, group by group
        for b, prm in zip(bss, prms):
            at.run_groups([b], [prm], timing=3)
    for (ix, prm, cs), b in zip(groups, bss):
        st = b.stats
        log(f"  group {cs[0]}..{cs[-1]}: {len(ix)} graphs n={b.n} "
            f"first {st['ms_first']:.1f} ms multi {st['ms_multi']:.1f} ms "
            f"init {st['ms_init']:.1f} rounds {st['rounds']} sub-rounds "
            f"{st['sub_rounds']} levels {st['levels_fwd']}/{st['levels_bwd']} "
            f"syncs {st['host_syncs']} resizes {st['resizes']}")
    if args.no_plain:
        return
    # one graph per launch, same parameters; only psc_run is timed
    plain_s, mism, checked = 0.0, 0, 0
    rng = np.random.default_rng(0)
    check = set(rng.choice(G, min(G, args.tarjan_checks), replace=False).tolist())
    for gi, (ix, prm, _) in enumerate(groups):
        parts = bss[gi].split(labs[gi])
        for j, g in enumerate(ix):
            dg = dbs[gi].extract(j)
            s = psc.SCC(dg)
            kw = dict(beta=float(prm["betas"][j]), tau=prm["tau"],
                      dense_div=prm["dense_div"], timing=0)
            if args.tail_max is not None:
                kw["tail_max"] = args.tail_max
            s.run(**kw)                            # warm-up (allocations)
            t = time.perf_counter()
            s.run(**kw)
            plain_s += time.perf_counter() - t
            lab = s.scc(**kw)
            if not np.array_equal(lab, parts[j]):
                mism += 1
            if int(g) in check:
                h = dg.to_host()
                if not np.array_equal(oracle.canonical(lab), oracle.tarjan_canonical(h)):
                    mism += 1
                checked += 1
            s.close()
            dg.free()
    log(f"plain: {plain_s * 1e3:.2f} ms ({G / plain_s:.1f} graphs/s); "
        f"mismatches {mism}; Tarjan-checked {checked}")
    out = {"metric": "batched SCC graphs/s vs one-graph-per-launch",
           "value": round(G / wall, 3), "unit": "graphs/s", "higher_is_better": True,
           "plain_graphs_per_s": round(G / plain_s, 3),
           "speedup_vs_plain": round(plain_s / wall, 3),
           "batched_ms": round(wall * 1e3, 3), "plain_ms": round(plain_s * 1e3, 3),
           "graphs": G, "n_total": n_tot, "m_total": m_tot,
           "batched_gedges_s": round(m_tot / wall / 1e9, 3),
           "labels_identical": mism == 0, "tarjan_checked": checked,
           "groups": [{"classes": cs, "family": int(kind[ix[0]]), "graphs": int(len(ix)),
                       "tau": p["tau"], "dense_div": p["dense_div"]} for ix, p, cs in groups],
           "gen_seconds": round(gen_s, 2)}
    print(json.dumps(out), flush=True)
    for b in bss:
        b.close()
    for db in dbs:
        db.free()


def bench_serve(args, psc):
    """C5: Poisson SCC requests (config-4 graph mix, n log-uniform in
    [2^10, 2^hi]) under a device budget, continuous batching vs one request
    per lane run; p50/p99 latency and graphs/s per arrival rate (0 = every
    request at t = 0).  Each request's graph is generated on the device at
    admission (its ingest), which is part of the served time."""
    import psc_amd.autotune as at
    import psc_amd.serve as sv
    table = at.load_table()
    rows = []
    for rate in args.rates:
        for mode in ("batched", "plain"):
            reqs = sv.poisson_requests(psc, args.requests, rate, 10, args.hi_log2,
                                       seed=11, table=table)
            srv = sv.Server(psc, budget_bytes=int(args.budget_gb * 1e9),
                            lanes=args.lanes, mode=mode,
                            max_batch_vertices=1 << args.max_batch_log2, progress_s=20.0)
            summ = srv.serve(reqs)
            summ["rate"] = rate
            rows.append(summ)
            log(json.dumps(summ))
    best = max((r for r in rows if r["mode"] == "batched"), key=lambda r: r["graphs_per_s"])
    plain = [r for r in rows if r["mode"] == "plain" and r["rate"] == best["rate"]][0]
    print(json.dumps({"metric": "SCC serving graphs/s and p50/p99 latency",
                      "value": best["graphs_per_s"], "unit": "graphs/s",
                      "higher_is_better": True, "rate": best["rate"],
                      "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"],
                      "plain_graphs_per_s": plain["graphs_per_s"],
                      "plain_p50_ms": plain["p50_ms"], "plain_p99_ms": plain["p99_ms"],
                      "requests": args.requests, "hi_log2": args.hi_log2,
                      "budget_GB": args.budget_gb, "sweep": rows}), flush=True)


def bench_copy(args, psc):
    """Calibration for the HBM counters: device-to-device copies of a known
    size (the runtime's plain copy kernel); FETCH_SIZE / WRITE_SIZE of these
    dispatches against the bytes copied give the counters' scale."""
    nbytes = args.copy_mb << 20
    a = psc.DeviceArray(nbytes // 8)
    b = psc.DeviceArray(nbytes // 8)
    for _ in range(args.warmup):
        psc.lib.psc_device_copy(psc.C.c_void_p(b.ptr), psc.C.c_void_p(a.ptr), nbytes)
    t = time.perf_counter()
    for _ in range(args.steps):
        psc._check("psc_device_copy", psc.lib.psc_device_copy(
            psc.C.c_void_p(b.ptr), psc.C.c_void_p(a.ptr), nbytes))
    dt = (time.perf_counter() - t) / args.steps
    print(json.dumps({"metric": "device copy", "bytes_per_copy": nbytes,
                      "ms_per_copy": round(dt * 1e3, 4),
                      "GBs_read_plus_write": round(2 * nbytes / dt / 1e9, 1)}), flush=True)
    a.free()
    b.free()


def bench_cc(args, psc):
    """Connectivity (src/Connectivity.h:37-119) of the undirected graph under
    the RMAT arcs (2^scale vertices, 2^samples_log2 samples, seed): kernel
    time from events, labels checked against the oracle's union-find."""
    import numpy as np
    import oracle
    scale, samples = args.scale, 1 << args.samples_log2
    dg = psc.DeviceGraph.rmat(scale, samples, 0.57, 0.19, 0.19, seed=args.seed, permute=True)
    for _ in range(args.warmup):
        psc.connectivity(dg)
    ms = []
    for _ in range(args.steps):
        labels = psc.connectivity(dg)
        ms.append(psc.connectivity.ms)
    dev_ms = float(np.median(ms))
    log(f"connectivity: {dev_ms:.3f} ms (kernels), {dg.m / (dev_ms * 1e-3) / 1e9:.1f} Gedges/s")
    ok = None
    cpu_s = None
    if not args.no_cpu:
        g = dg.to_host()
        t = time.perf_counter()
        expect = oracle.cc(g)
        cpu_s = time.perf_counter() - t
        ok = bool(np.array_equal(labels, expect))
        log(f"oracle union-find (1 thread): {cpu_s:.2f} s; labels identical: {ok}")
    print(json.dumps({"metric": "connectivity Gedges/s on RMAT",
                      "value": round(dg.m / (dev_ms * 1e-3) / 1e9, 2), "unit": "Gedges/s",
                      "ms": round(dev_ms, 3), "n": dg.n, "m": dg.m,
                      "components": int(np.unique(labels).size),
                      "labels_equal_oracle": ok,
                      "oracle_seconds_1_thread": None if cpu_s is None else round(cpu_s, 2)}),
          flush=True)
    dg.free()


def bench_le(args, psc):
    """LE-lists (src/LeList.hpp:246-328) of the RMAT arcs (2^scale vertices,
    2^samples_log2 samples, seed) for a seeded random source order, beta
    1.5 (the reference's default, src/LeList.cpp:21): kernel time from
    events, rows checked against the oracle's seq_le_list."""
    import numpy as np
    import oracle
    scale, samples = args.scale, 1 << args.samples_log2
    dg = psc.DeviceGraph.rmat(scale, samples, 0.57, 0.19, 0.19, seed=args.seed, permute=True)
    order = psc.random_order(dg.n, args.seed)
    for _ in range(args.warmup):
        psc.le_lists(dg, order)
    ms = []
    for _ in range(args.steps):
        rows = psc.le_lists(dg, order)
        ms.append(psc.le_lists.ms)
    dev_ms = float(np.median(ms))
    log(f"le_lists: {dev_ms:.3f} ms, {rows.shape[0]} rows")
    ok = None
    cpu_s = None
    if not args.no_cpu:
        g = dg.to_host()
        t = time.perf_counter()
        expect = oracle.le_lists(g, order)
        cpu_s = time.perf_counter() - t
        ok = bool(np.array_equal(rows, expect))
        log(f"oracle seq_le_list (1 thread): {cpu_s:.2f} s; rows identical: {ok}")
    print(json.dumps({"metric": "LE-lists ms on RMAT", "value": round(dev_ms, 3), "unit": "ms",
                      "higher_is_better": False, "n": dg.n, "m": dg.m,
                      "rows": int(rows.shape[0]), "rows_equal_oracle": ok,
                      "oracle_seconds_1_thread": None if cpu_s is None else round(cpu_s, 2)}),
          flush=True)
    dg.free()


def numa_interleave():
    """The CPU baseline's memory policy: `numactl -i all` on a host with more
    than one NUMA node (the reference's setting, README.md:191), applied to
    this process's later allocations through set_mempolicy."""
    import ctypes
    import glob
    nodes = sorted(int(p.rsplit("node", 1)[1])
                   for p in glob.glob("/sys/devices/system/node/node[0-9]*"))
    if len(nodes) <= 1:
        return {"nodes": max(len(nodes), 1), "policy": "single node, none needed"}
    mask = ctypes.c_ulong(sum(1 << k for k in nodes if k < 64))
    libc = ctypes.CDLL(None, use_errno=True)
    MPOL_INTERLEAVE, SYS_set_mempolicy = 3, 238
    rc = libc.syscall(SYS_set_mempolicy, MPOL_INTERLEAVE, ctypes.byref(mask), 65)
    return {"nodes": len(nodes), "policy": "interleave all" if rc == 0 else
            f"set_mempolicy failed (errno {ctypes.get_errno()})"}


def main():
    # OpenMP threads of the CPU baseline spread over the cores (the
    # reference's experiments, README.md:191); read when libgomp loads
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=25)
    ap.add_argument("--samples-log2", type=int, default=29)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--tau", type=int, default=512)
    ap.add_argument("--dense-div", type=int, default=20)
    ap.add_argument("--tail-max", type=int, default=None,
                    help="largest remainder run in one launch (0: host rounds)")
    ap.add_argument("--no-vgc", action="store_true")
    ap.add_argument("--levels", action="store_true",
                    help="report the first reach level by level (stderr)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile-only", action="store_true",
                    help="warm-up + steps only (for rocprofv3 runs)")
    ap.add_argument("--mode", choices=("c3", "grid", "batch", "serve", "copy", "cc", "le"),
                    default="c3")
    ap.add_argument("--copy-mb", type=int, default=1024)
    ap.add_argument("--grid-side", type=int, default=2048)
    ap.add_argument("--taus", type=int, nargs="*", default=None)
    ap.add_argument("--deep-vgc", action="store_true",
                    help="grid sweep: let a wave's local search run its whole tau "
                         "budget (vgc_batches = 4 tau / 64) instead of 64 batches")
    ap.add_argument("--graphs", type=int, default=2048)
    ap.add_argument("--hi-log2", type=float, default=None)
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--tarjan-checks", type=int, default=64)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--requests", type=int, default=10000)
    ap.add_argument("--rates", type=float, nargs="*", default=[100.0, 150.0, 200.0, 0.0])
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--max-batch-log2", type=int, default=27,
                    help="serving: vertex cap of one lane's batch")
    ap.add_argument("--cpu-runs", type=int, default=10,
                    help="timed runs of the CPU baseline after one warm-up (src/scc.cpp:55-68)")
    ap.add_argument("--no-seq", action="store_true",
                    help="skip the sequential Tarjan time of the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not "
              "supported (the multi-GPU sharding is a design, not a build)",
              file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge
    psc = ge.load_package()
    if args.mode == "le":
        return bench_le(args, psc)
    if args.mode == "cc":
        return bench_cc(args, psc)
    if args.mode == "copy":
        return bench_copy(args, psc)
    if args.mode == "grid":
        return bench_grid(args, psc)
    if args.mode == "batch":
        args.hi_log2 = args.hi_log2 or 20
        return bench_batch(args, psc)
    if args.mode == "serve":
        args.hi_log2 = args.hi_log2 or 24
        return bench_serve(args, psc)

    scale, samples = args.scale, 1 << args.samples_log2
    workload = f"rmat{scale}_s{args.samples_log2}"
    # process-level HIP set-up (code objects, memory pool, a first stream)
    # on a tiny graph, so that create_ms below is the per-graph set-up alone
    t = time.perf_counter()
    tiny = psc.DeviceGraph.rmat(10, 1 << 14, seed=1)
    ts = psc.SCC(tiny)
    ts.run()
    ts.close()
    tiny.free()
    process_init_s = time.perf_counter() - t
    t = time.perf_counter()
    dg = psc.DeviceGraph.rmat(scale, samples, 0.57, 0.19, 0.19, seed=args.seed,
                              permute=True)
    gen_s = time.perf_counter() - t
    log(f"generated {workload}: n={dg.n} m={dg.m} in {gen_s:.2f}s")
    t = time.perf_counter()
    s = psc.SCC(dg)   # per-graph set-up (SCC::SCC), outside the timed steps
    create_s = time.perf_counter() - t
    log(f"handle created in {create_s * 1e3:.2f} ms (buffers, pull heads)")
    prm = dict(beta=1.5, local_reach=not args.no_vgc, local_scc=not args.no_vgc,
               tau=args.tau, dense_div=args.dense_div, timing=1)
    if args.tail_max is not None:
        prm["tail_max"] = args.tail_max
    # the first run on a new handle also sizes the round buffers: with the
    # set-up it is the cold cost of one SCC on a graph already in HBM
    t = time.perf_counter()
    s.run(**prm)
    first_run_s = time.perf_counter() - t
    for _ in range(max(args.warmup - 1, 0)):
        s.run(**prm)
    times = []
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        st = s.run(**prm)          # returns after the stream is synchronised
        times.append(time.perf_counter() - a)
        dev_ms.append(st["ms_total"])
    t1 = time.perf_counter()
    total = t1 - t0
    if args.profile_only:
        log(f"profile-only: {args.steps} steps, {total / args.steps * 1e3:.3f} ms/step")
        return
    ms_step = total / args.steps * 1e3
    gedges = dg.m * args.steps / total / 1e9
    log(f"{args.steps} steps: {ms_step:.3f} ms/step (min {min(times)*1e3:.3f}), "
        f"device {sum(dev_ms)/len(dev_ms):.3f} ms, {gedges:.3f} Gedges/s")
    log("stats:", json.dumps(st))

    if args.levels:   # per-level report of the first reach on stderr
        s.run(**dict(prm, timing=3))
    # per-kernel profile: one extra pass with an event pair per launch
    s.run(**dict(prm, timing=2))
    prof_st = s.stats
    kp = [k for k in s.kernel_profile() if k["launches"]]
    for k in kp:
        log(f"  {k['name']:22s} launches {k['launches']:6d}  {k['total_ms']:9.3f} ms  "
            f"{k['bytes'] / 1e9:8.3f} GB  "
            f"{k['bytes'] / max(k['total_ms'], 1e-9) / 1e6:8.1f} GB/s")
    dom = max(kp, key=lambda k: k["total_ms"])
    avg_ms = dom["total_ms"] / dom["launches"]
    bpl = dom["bytes"] / dom["launches"]
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    traffic = load_pmc(dom["name"], workload)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2),
                "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": traffic, "kernel": dom["name"],
                "launches": dom["launches"], "avg_launch_us": round(avg_ms * 1e3, 3),
                "bytes_per_launch": round(bpl)}
    path_bytes = sum(prof_st[k] for k in ("bytes_csr", "bytes_label",
                                          "bytes_table", "bytes_front"))
    mean_step_s = total / args.steps
    path = {"algorithmic_bytes": path_bytes,
            "achieved_GBs": round(path_bytes / mean_step_s / 1e9, 2),
            "frac": round(path_bytes / mean_step_s / 1e9 / PEAK_HBM_GBS, 4),
            "floor_bytes": 2 * (8 * (dg.n + 1) + 4 * dg.m) + 16 * dg.n}

    # set-up of further handles on the same graph once the memory pool holds
    # their buffers (a re-bound or re-created handle's cost)
    warm = []
    for _ in range(4):
        t = time.perf_counter()
        s2 = psc.SCC(dg)
        warm.append(time.perf_counter() - t)
        s2.close()
    create_warm_s = sorted(warm[1:])[1]
    log(f"handle re-created in {create_warm_s * 1e3:.2f} ms (median of 3, warm pool)")

    cpu = None
    parity = None
    if not args.no_cpu:
        import numpy as np
        import oracle
        labels = s.scc(**prm)
        numa = numa_interleave()
        t = time.perf_counter()
        g = dg.to_host()
        log(f"graph downloaded in {time.perf_counter() - t:.1f}s")
        nproc = os.cpu_count() or 1
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0")) \
            or len(os.sched_getaffinity(0))
        # the reference's protocol (src/scc.cpp:55-68, scripts/Experiment1.py:
        # 108-109): construct once, one warm-up run, then timed runs
        ref = oracle.RefSCC(g, beta=1.5, local_reach=not args.no_vgc,
                            local_scc=not args.no_vgc, tau=args.tau, threads=threads)
        rl, rst = ref.run()
        cpu_times = []
        for _ in range(args.cpu_runs):
            t = time.perf_counter()
            rl, rst = ref.run()
            cpu_times.append(time.perf_counter() - t)
        ref.close()
        cpu_mean, cpu_min = sum(cpu_times) / len(cpu_times), min(cpu_times)
        parity = bool(np.array_equal(rl, labels))
        log(f"cpu restatement: mean {cpu_mean:.3f}s min {cpu_min:.3f}s over "
            f"{len(cpu_times)} runs on {threads} threads; raw labels "
            f"{'identical' if parity else 'DIFFER'}")
        seq_s = None
        if not args.no_seq:   # the "SEQ" column (scripts/Experiment1.py:51)
            t = time.perf_counter()
            oracle.tarjan(g)
            seq_s = time.perf_counter() - t
            log(f"sequential Tarjan: {seq_s:.2f}s")
        cpu = {"value": round(dg.m / cpu_mean / 1e9, 4), "unit": "Gedges/s",
               "cores": threads, "kind": "port",
               "sample": f"the OpenMP restatement of the reference (oracle/scc_ref.cpp) "
                         f"on the full {workload} graph, constructed once, 1 warm-up + "
                         f"{len(cpu_times)} timed scc() runs (value = m / mean)",
               "mean_s": round(cpu_mean, 4), "min_s": round(cpu_min, 4),
               "runs": len(cpu_times), "warmup": 1, "nproc": nproc,
               "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
               "numa": numa,
               "seq_tarjan_s": None if seq_s is None else round(seq_s, 3),
               "seq_tarjan_gedges_s": None if seq_s is None else round(dg.m / seq_s / 1e9, 4)}
        if not parity:
            log("ERROR: GPU labels differ from the oracle restatement")
    out = {
        "metric": "SCC Gedges/s + %HBM roofline on RMAT 2^25",
        "value": round(gedges, 4),
        "unit": "Gedges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64",
        "data": "synthetic",
        "config": {"workload": f"RMAT n=2^{scale}, 2^{args.samples_log2} samples, "
                               f"(a,b,c,d)=(.57,.19,.19,.05), seed {args.seed}, "
                               "dedup, ids scrambled",
                   "n": dg.n, "m": dg.m, "beta": 1.5, "tau": args.tau,
                   "vgc": not args.no_vgc},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "path_roofline": path,
        "parity_vs_oracle": parity,
        "phases_ms": {k: round(prof_st[k], 3) for k in
                      ("ms_init", "ms_first", "ms_multi", "ms_resize", "ms_finalize")},
        "gen_seconds": round(gen_s, 2),
        "process_init_ms": round(process_init_s * 1e3, 2),
        "create_ms": round(create_s * 1e3, 2),
        "create_warm_ms": round(create_warm_s * 1e3, 2),
        "first_run_ms": round(first_run_s * 1e3, 3),
        "cold": {"ms": round((create_s + first_run_s) * 1e3, 3),
                 "gedges_s": round(dg.m / (create_s + first_run_s) / 1e9, 2),
                 "note": "handle set-up (SCC::SCC) + the first scc() on it, graph "
                         "resident in HBM"},
    }
    print(json.dumps(out), flush=True)
    s.close()
    dg.free()


if __name__ == "__main__":
    main()
