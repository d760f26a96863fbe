"""Per-size-class parameters for batched SCC, and the class-batch runner.

A size class is floor(log2 n).  For each class the tuner times the batched
SCC of a sample batch of that class's graphs over a small grid of the three
schedule knobs -- tau (VGC budget), beta (pivot-batch growth) and dense_div
(the sparse/dense switch m / dense_div of src/reach.hpp:49,103) -- by
coordinate descent, and keeps the fastest.  tau and dense_div never change
labels; beta changes raw label values but not the partition, and the
one-graph-per-launch path uses the same per-class beta, so both paths
produce identical labels.

The runner executes one batch per class group on its own HIP stream from a
small pool of host threads (ctypes releases the GIL), so the groups' work
overlaps on the device.
"""
import json
import os
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(HERE, "tuned_params.json")

DEFAULT = {"tau": 512, "beta": 1.5, "dense_div": 20}
TAUS = (16, 128, 512, 4096)
BETAS = (1.5, 2.0, 3.0)
DENSE_DIVS = (5, 20, 80)


def size_class(n):
    return int(np.floor(np.log2(max(int(n), 2))))


def load_table(path=TABLE_PATH):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(k): v for k, v in json.load(f)["classes"].items()}


def params_for(table, cls):
    if cls in table:
        return {k: table[cls][k] for k in DEFAULT}
    if table:   # nearest tuned class
        near = min(table, key=lambda c: abs(c - cls))
        return {k: table[near][k] for k in DEFAULT}
    return dict(DEFAULT)


def _time_batch(b, prm, reps=2):
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        b.run(labels=False, beta=prm["beta"], tau=prm["tau"],
              dense_div=prm["dense_div"], timing=0)
        ts.append(time.perf_counter() - t)
    return min(ts)


def tune_class(psc, kind, n, seeds, log=print):
    """Coordinate descent over (beta, tau, dense_div) for one class sample."""
    db = psc.DeviceBatch(kind, n, seeds=seeds)
    b = psc.BatchSCC(device_batch=db)
    best = dict(DEFAULT)
    _time_batch(b, best, reps=1)   # warm-up
    best_t = _time_batch(b, best)
    for key, grid in (("beta", BETAS), ("tau", TAUS), ("dense_div", DENSE_DIVS)):
        for val in grid:
            if val == best[key]:
                continue
            cand = dict(best, **{key: val})
            t = _time_batch(b, cand)
            if t < best_t * 0.97:
                best, best_t = cand, t
    b.close()
    db.free()
    log(f"  tuned {best} {best_t * 1e3:.2f} ms for {len(kind)} graphs")
    return dict(best, ms=round(best_t * 1e3, 3), graphs=int(len(kind)))


def tune(psc, lo=10, hi=24, per_class=24, seed=77, log=print, path=TABLE_PATH):
    """Tune every class in [lo, hi] on a sample of per_class graphs (the
    kind mix of the batch generator, sizes inside the class)."""
    table = {}
    for cls in range(lo, hi + 1):
        cnt = per_class if cls < 20 else max(4, per_class // (2 ** (cls - 19)))
        kind, n = psc.batch_spec(cnt, cls, cls + 0.999, seed=seed + cls)
        seeds = np.array([psc.graph_seed(seed + cls, g) for g in range(cnt)],
                         dtype=np.uint64)
        log(f"class 2^{cls}: {cnt} graphs, n_total={int(n.sum())}")
        table[cls] = tune_class(psc, kind, n, seeds, log=log)
    with open(path, "w") as f:
        json.dump({"classes": {str(k): v for k, v in table.items()},
                   "grid": {"tau": TAUS, "beta": BETAS, "dense_div": DENSE_DIVS}},
                  f, indent=1)
    return table


def group_by_class(n, table, max_groups=4, family=None):
    """Partition graph indices into at most max_groups groups of adjacent
    size classes, balancing total vertices; each group runs with the
    parameters of its dominant class.

    family (optional, one id per graph, e.g. the generator of a batch
    graph): graphs of different families go to different groups, at most
    max(1, max_groups // families) class groups per family.  A handle
    switches between push and pull per level for all its graphs together,
    and runs as many levels as its deepest graph: a skewed graph (RMAT) that
    is done in a few pull-heavy levels would otherwise ride along the many
    levels of a lattice or k-NN graph."""
    n = np.asarray(n)
    if family is not None:
        family = np.asarray(family)
        fams = sorted(set(family.tolist()))
        per = max(1, max_groups // len(fams))
        out = []
        for f in fams:
            fi = np.flatnonzero(family == f)
            for ix, prm, cs in group_by_class(n[fi], table, max_groups=per):
                out.append((fi[ix], prm, cs))
        return out
    cls = np.array([size_class(x) for x in n])
    classes = sorted(set(cls.tolist()))
    weight = {c: float(np.sum(n[cls == c])) for c in classes}
    total = sum(weight.values())
    groups, cur, acc = [], [], 0.0
    for c in classes:
        cur.append(c)
        acc += weight[c]
        if acc >= total / max_groups and len(groups) < max_groups - 1:
            groups.append(cur)
            cur, acc = [], 0.0
    if cur:
        groups.append(cur)
    out = []
    for cs in groups:
        idx = np.flatnonzero(np.isin(cls, cs))
        dom = max(cs, key=lambda c: weight[c])
        prm = params_for(table, dom)
        # beta stays per graph (its own class), tau / dense_div per group
        prm["betas"] = np.array([params_for(table, c)["beta"] for c in cls[idx]],
                                dtype=np.float64)
        out.append((idx, prm, cs))
    return out


def run_groups(batches, params, labels=False, timing=0):
    """Run BatchSCC objects concurrently, one host thread each.  Returns
    (wall seconds, per-batch label arrays or None)."""
    results = [None] * len(batches)
    errors = []

    def work(i):
        try:
            p = params[i]
            extra = {"tail_max": p["tail_max"]} if "tail_max" in p else {}
            results[i] = batches[i].run(labels=labels, beta=p["beta"],
                                        beta_per_graph=p.get("betas"), tau=p["tau"],
                                        dense_div=p["dense_div"], timing=timing,
                                        **extra)
        except Exception as e:   # surfaced to the caller below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(batches))]
    t = time.perf_counter()
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    wall = time.perf_counter() - t
    if errors:
        raise errors[0]
    return wall, results


def main():
    """python parallel-strong-connectivity_amd/autotune.py [lo hi [out.json]]"""
    import importlib.util
    import sys
    spec = importlib.util.spec_from_file_location(
        "psc_amd", os.path.join(HERE, "__init__.py"),
        submodule_search_locations=[HERE])
    psc = importlib.util.module_from_spec(spec)
    sys.modules["psc_amd"] = psc
    spec.loader.exec_module(psc)
    lo = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    hi = int(sys.argv[2]) if len(sys.argv) > 2 else 24
    path = sys.argv[3] if len(sys.argv) > 3 else TABLE_PATH
    tune(psc, lo, hi, log=lambda s: print(s, flush=True), path=path)


if __name__ == "__main__":
    main()
