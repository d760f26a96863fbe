"""SCC serving with continuous batching.

Requests (a graph kind, a vertex count and a seed) arrive over time.  The
admission thread generates each admitted request's graph straight into HBM
(device generator, the request's own seed) as soon as the device pool's
byte budget allows it; admitted graphs wait in the ready queue.  A fixed
number of lanes -- host threads, each owning one re-bindable batch handle
and its HIP stream -- repeatedly take everything that is ready (up to a
vertex cap), fold it into one batch (device gather into the handle's union
buffers), run the batched SCC and hand each request its labels.  Requests
that arrive while lanes are busy join the next batch of the first free
lane, so the work of several batches overlaps on the device.  Memory is
released, and waiting requests admitted, as batches complete.

mode="plain" serves the same stream one request per launch (batch size 1)
with the same lanes: the one-graph-per-launch baseline.

Every request's labels equal a stand-alone SCC of its graph with the same
beta and seed (the batched engine keeps per-graph label namespaces and
schedules), which tests/test_gpu_serve.py checks.
"""
import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

AVG_DEG = {0: 16.0, 1: 3.2, 2: 5.0}   # generator edges per vertex by kind


@dataclass
class Request:
    rid: int
    kind: int
    n: int
    seed: int
    arrival: float            # seconds after the start of the run
    beta: float = 1.5
    # filled by the server
    admitted: float = 0.0
    started: float = 0.0
    done: float = 0.0
    batch: int = 0            # size of the batch it ran in
    labels: np.ndarray = field(default=None, repr=False)


def poisson_requests(psc, count, rate, lo_log2=10, hi_log2=24, seed=11,
                     kinds=(0, 1, 2), table=None):
    """count requests with exponential inter-arrival times of mean 1/rate;
    graphs from the batch generator's mix (kinds, log-uniform sizes)."""
    from . import autotune as at
    kind, n = psc.batch_spec(count, lo_log2, hi_log2, seed=seed, kinds=kinds)
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count)) if rate > 0 else np.zeros(count)
    out = []
    for i in range(count):
        beta = at.params_for(table or {}, at.size_class(n[i]))["beta"]
        out.append(Request(i, int(kind[i]), int(n[i]), int(psc.graph_seed(seed, i)),
                           float(t[i]), beta))
    return out


def est_bytes(kind, n):
    """Device bytes one request holds while admitted: its generated CSR in
    both directions plus its share of a lane's union copy and working set."""
    m = AVG_DEG.get(kind, 8.0) * n
    csr = 2 * (8 * (n + 1) + 4 * m)
    return int(2 * csr + 96 * n + (1 << 20))


class _Pool:
    """A device batch shared by the requests admitted together; freed when
    the last of them completes."""

    def __init__(self, db, users):
        self.db = db
        self.users = users
        self.lock = threading.Lock()

    def release(self):
        with self.lock:
            self.users -= 1
            last = self.users == 0
        if last:
            self.db.free()


class Server:
    def __init__(self, psc, budget_bytes=200 << 30, lanes=4, mode="batched",
                 max_batch_vertices=1 << 27, tau=512, dense_div=20,
                 keep_labels=False, target_batch=16, max_wait_s=0.002,
                 max_admit=256):
        self.psc = psc
        self.max_admit = max_admit
        self.target_batch = target_batch
        self.max_wait_s = max_wait_s
        self.budget = int(budget_bytes)
        self.lanes = lanes
        self.mode = mode
        self.max_batch_vertices = max_batch_vertices
        self.tau = tau
        self.dense_div = dense_div
        self.keep_labels = keep_labels
        self.cv = threading.Condition()
        self.ready = []          # (request, _Pool, index in the pool, bytes)
        self.in_use = 0
        self.peak_in_use = 0
        self.finished = 0
        self.errors = []
        # seconds spent per phase (summed over threads)
        self.phase = {"generate": 0.0, "gather": 0.0, "scc": 0.0, "labels": 0.0}
        self.calls = {"generate": 0, "batches": 0}

    # ---------------------------------------------------------- admission
    def _admit_loop(self, reqs, t0):
        """Admit every request that has arrived and fits the budget; the
        graphs of one admission tick are generated together as one device
        batch (one generator call), and lanes gather per-request views."""
        i = 0
        total = len(reqs)
        while i < total:
            if self.errors:
                return
            now = time.perf_counter() - t0
            if reqs[i].arrival > now:
                time.sleep(reqs[i].arrival - now)
                continue
            group = []
            err = None
            with self.cv:
                while i < total and reqs[i].arrival <= time.perf_counter() - t0:
                    r = reqs[i]
                    need = est_bytes(r.kind, r.n)
                    if need > self.budget:
                        err = RuntimeError(f"request {r.rid} needs {need} B, over the "
                                           f"{self.budget} B budget")
                        break
                    if self.in_use + need > self.budget:
                        if group:
                            break
                        # wait for memory to free up; a failed lane ends
                        # the run (its error is raised by serve())
                        self.cv.wait(timeout=0.5)
                        if self.errors:
                            return
                        continue
                    self.in_use += need
                    self.peak_in_use = max(self.peak_in_use, self.in_use)
                    group.append((r, need))
                    i += 1
                    if len(group) >= self.max_admit:
                        break
            if err is not None:
                with self.cv:
                    for _, need in group:
                        self.in_use -= need
                    self.finished += len(group)
                self._abort(err, total)
                return
            if not group:
                continue
            tg = time.perf_counter()
            try:
                db = self.psc.DeviceBatch(
                    np.array([r.kind for r, _ in group], np.uint32),
                    np.array([r.n for r, _ in group], np.uint64),
                    seeds=np.array([r.seed for r, _ in group], np.uint64))
            except Exception as e:
                with self.cv:
                    for _, need in group:
                        self.in_use -= need
                    self.finished += len(group)
                self._abort(e, total)
                return
            shared = _Pool(db, len(group))
            adm = time.perf_counter() - t0
            self.phase["generate"] += adm + t0 - tg
            self.calls["generate"] += 1
            with self.cv:
                for j, (r, need) in enumerate(group):
                    r.admitted = adm
                    self.ready.append((r, shared, j, need))
                self.cv.notify_all()

    # --------------------------------------------------------------- lanes
    def _take(self):
        """Everything ready (batched) or the oldest request (plain)."""
        if self.mode == "plain":
            return [self.ready.pop(0)]
        take, nv = [], 0
        while self.ready and (not take or
                              nv + self.ready[0][0].n <= self.max_batch_vertices):
            item = self.ready.pop(0)
            nv += item[0].n
            take.append(item)
        return take

    def _release(self, items):
        """Give back the pool references and the budget of finished (or
        failed) items."""
        for _, pool, _, _ in items:
            pool.release()
        with self.cv:
            for _, _, _, need in items:
                self.in_use -= need
            self.finished += len(items)
            self.cv.notify_all()

    def _abort(self, err, total):
        """A failed lane ends the run: the error is kept for serve(), every
        waiting request gives its memory back, the other threads stop."""
        with self.cv:
            self.errors.append(err)
            waiting, self.ready = self.ready, []
        self._release(waiting)
        with self.cv:
            self.finished = total
            self.cv.notify_all()

    def _keep(self, r):
        k = self.keep_labels
        return bool(k) if isinstance(k, bool) else r.rid in k

    def _lane_loop(self, total, t0):
        psc = self.psc
        handle = None
        pinned = None
        items = []
        try:
            pinned = psc.PinnedLabels()
            while True:
                with self.cv:
                    while not self.ready and self.finished < total and not self.errors:
                        self.cv.wait(timeout=0.5)
                    if self.errors or (not self.ready and self.finished >= total):
                        return
                    if not self.ready:
                        continue
                    # dynamic batching: a short wait for more requests unless
                    # the batch is already large or the oldest has waited
                    if self.mode == "batched":
                        while (len(self.ready) < self.target_batch and
                               self.finished + len(self.ready) < total):
                            waited = time.perf_counter() - t0 - self.ready[0][0].admitted
                            if waited >= self.max_wait_s:
                                break
                            self.cv.wait(timeout=self.max_wait_s - waited)
                            if not self.ready:
                                break
                        if not self.ready:
                            continue
                    items = self._take()
                handle = self._run_batch(handle, items, pinned, t0)
                done, items = items, []
                self._release(done)
        except Exception as e:   # surfaced by serve()
            failed, items = items, []
            self._release(failed)
            self._abort(e, total)
        finally:
            if handle is not None:
                handle.close()
            if pinned is not None:
                pinned.free()

    def _run_batch(self, handle, items, pinned, t0):
        """Bind the lane's handle to the items' graphs (views into their
        admission batches), run the batched SCC, hand out the labels."""
        psc = self.psc
        start = time.perf_counter() - t0
        views = [pool.db.view(j) for _, pool, j, _ in items]
        arr = (psc.CGraph * len(views))(*[v[0] for v in views])
        vb = np.array([v[1] for v in views], dtype=np.uint64)
        eb = np.array([v[2] for v in views], dtype=np.uint64)
        vbp, ebp = psc._ptr(vb, psc.u64p), psc._ptr(eb, psc.u64p)
        if handle is None:
            h = C.c_void_p()
            psc._check("psc_batch_create_from_graphs",
                       psc.lib.psc_batch_create_from_graphs(
                           arr, len(views), vbp, ebp, C.byref(h)))
            handle = psc.BatchSCC._wrap(h)
        else:
            psc._check("psc_batch_rebind",
                       psc.lib.psc_batch_rebind(handle._h, arr, len(views), vbp, ebp))
            handle._refresh()
        t1 = time.perf_counter()
        betas = np.array([r.beta for r, _, _, _ in items], dtype=np.float64)
        labels = handle.run(beta_per_graph=betas, tau=self.tau,
                            dense_div=self.dense_div, timing=0,
                            out=pinned.view(handle.n))
        done = time.perf_counter() - t0
        scc_s = handle.stats["ms_total"] / 1e3
        self.phase["gather"] += t1 - t0 - start
        self.calls["batches"] += 1
        self.phase["scc"] += scc_s
        self.phase["labels"] += done + t0 - t1 - scc_s
        parts = handle.split(labels)
        for (r, _, _, _), lab in zip(items, parts):
            r.started, r.done, r.batch = start, done, len(items)
            if self._keep(r):
                r.labels = lab.copy()
        return handle

    def serve(self, reqs):
        """Run the request stream; returns a summary dict."""
        reqs = sorted(reqs, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        adm = threading.Thread(target=self._admit_loop, args=(reqs, t0))
        lanes = [threading.Thread(target=self._lane_loop, args=(len(reqs), t0))
                 for _ in range(self.lanes)]
        adm.start()
        for th in lanes:
            th.start()
        adm.join()
        for th in lanes:
            th.join()
        wall = time.perf_counter() - t0
        if self.errors:
            raise self.errors[0]
        lat = np.array([r.done - r.arrival for r in reqs])
        return {
            "mode": self.mode, "requests": len(reqs), "wall_s": round(wall, 4),
            "graphs_per_s": round(len(reqs) / wall, 3),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "mean_batch": round(float(np.mean([r.batch for r in reqs])), 2),
            "peak_pool_GB": round(self.peak_in_use / 2**30, 2),
            "budget_GB": round(self.budget / 2**30, 2),
            "phase_s": {k: round(v, 3) for k, v in self.phase.items()},
            "calls": dict(self.calls),
        }
