"""SCC serving with continuous batching under an exact device-memory budget.

Requests (a graph kind, a vertex count and a seed) arrive over time.  The
admission thread generates each admitted request's graph straight into HBM
(device generator, the request's own seed); the requests that have arrived
by one admission tick are generated together as one device batch (the
graph pool).  A fixed number of lanes -- host threads, each owning one
serving lane handle (psc_lane_create: its own HIP stream and one fixed
device arena that every buffer of its bindings is carved from) -- take what
is ready, fold it into one batch (device gather into the arena), run the
batched SCC and hand each request its labels.  Requests that arrive while
lanes are busy join the next batch of the first free lane, so several
batches overlap on the device.

Memory is accounted exactly, never estimated:
  * the lanes' arenas are allocated once and reserved from the budget up
    front;
  * an admission group reserves what the library reports the generator will
    allocate at its peak (psc_dev_gen_batch_bytes), gives back the transient
    part as soon as the graphs exist and the rest when its last request is
    done;
  * the graph pool lives in the device's stream-ordered memory pool, which
    keeps freed memory (and the gaps between live allocations) reserved
    from the driver: a group is admitted only when the lanes' arenas plus
    what that pool really holds (psc_pool_usage) plus the group's peak fit
    the budget, and the pool's unused memory is given back (psc_pool_trim)
    when it alone is in the way;
  * a lane takes ready requests while their binding (psc_batch_layout_bytes)
    fits a share of its arena; a batch whose rounds outgrow the arena
    (PSC_ERR_NOMEM) is split in halves and run again, with the same labels.
So the device memory held for serving stays within the budget; the
driver's own figure (hipMemGetInfo) is sampled after every generation and
batch and reported beside it.

mode="plain" serves the same stream one request per lane run (batch size
1): the one-graph-per-launch baseline.

Every request's labels equal a stand-alone SCC of its graph with the same
beta and seed (the batched engine keeps per-graph label namespaces and
schedules), which tests/test_gpu_serve.py and tests/test_gpu_configs.py
check.
"""
import ctypes as C
import threading
import time
from dataclasses import dataclass, field

import numpy as np

# share of a lane's arena a binding may take before its first run; the rest
# holds the round buffers (pivot bitmap, tables, logs, remainder copy)
LANE_FILL = 0.6


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
    m: int = 0                # edges of its generated graph
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


def gen_bytes(psc, reqs):
    """(keep, peak) device bytes of generating these requests' graphs as one
    device batch (the library's own count)."""
    kind = np.array([r.kind for r in reqs], np.uint32)
    n = np.array([r.n for r in reqs], np.uint64)
    keep, peak = C.c_uint64(), C.c_uint64()
    psc._check("psc_dev_gen_batch_bytes", psc.lib.psc_dev_gen_batch_bytes(
        len(reqs), psc._ptr(kind, psc.u32p), psc._ptr(n, psc.u64p),
        C.byref(keep), C.byref(peak)))
    return keep.value, peak.value


def gen_bound_m(kind, n):
    """Most edges the generator can give a request (its key count)."""
    if kind == 0:
        return 16 * (1 << max(1, int(np.ceil(np.log2(max(n, 2))))))
    if kind == 1:
        side = max(2, int(round(np.sqrt(n))))
        return 2 * 2 * side * (side - 1)
    return 5 * max(n, 6)


class _Pool:
    """A device batch shared by the requests admitted together; freed, and
    its bytes given back, when the last of them completes."""

    def __init__(self, server, db, users, keep):
        self.server = server
        self.db = db
        self.users = users
        self.keep = keep
        self.lock = threading.Lock()

    def release(self):
        with self.lock:
            self.users -= 1
            last = self.users == 0
        if last:
            self.db.free()
            self.server._unreserve(self.keep)


class Server:
    def __init__(self, psc, budget_bytes=200 << 30, lanes=4, mode="batched",
                 lane_bytes=None, max_batch_vertices=1 << 27, tau=512, dense_div=20,
                 keep_labels=False, target_batch=16, max_wait_s=0.002,
                 max_admit=256, capture=1, progress_s=0.0):
        self.psc = psc
        self.max_admit = max_admit
        self.target_batch = target_batch
        self.max_wait_s = max_wait_s
        self.budget = int(budget_bytes)
        self.lanes = lanes
        # a quarter of the budget for the lanes' arenas by default
        self.lane_bytes = int(lane_bytes or min(self.budget // (4 * lanes), 16 << 30))
        self.mode = mode
        self.max_batch_vertices = max_batch_vertices
        self.tau = tau
        self.dense_div = dense_div
        self.capture = capture
        self.keep_labels = keep_labels
        self.progress_s = progress_s   # > 0: a progress line on stderr this often
        self.cv = threading.Condition()
        self.ready = []          # (request, _Pool, index in the pool)
        self.reserved = 0        # bytes of the budget in use (exact)
        self.peak_reserved = 0
        self.finished = 0
        self.errors = []
        self.splits = 0
        self.trims = 0
        self.dev_base = None     # driver-reported use before serving
        self.peak_dev = 0
        # seconds spent per phase (summed over threads)
        self.phase = {"generate": 0.0, "gather": 0.0, "scc": 0.0, "labels": 0.0}
        self.calls = {"generate": 0, "batches": 0}

    # ---------------------------------------------------------- accounting
    def _reserve_now(self, b):
        """Take b bytes of the budget (the caller holds self.cv)."""
        self.reserved += b
        self.peak_reserved = max(self.peak_reserved, self.reserved)

    def _unreserve(self, b):
        with self.cv:
            self.reserved -= b
            self.cv.notify_all()

    def _sample_device(self):
        f, t = C.c_uint64(), C.c_uint64()
        if self.psc.lib.psc_device_mem_info(C.byref(f), C.byref(t)) == 0:
            used = t.value - f.value
            with self.cv:
                self.peak_dev = max(self.peak_dev, used)
            return used
        return None

    def _pool_usage(self):
        r, u = C.c_uint64(), C.c_uint64()
        self.psc._check("psc_pool_usage", self.psc.lib.psc_pool_usage(C.byref(r), C.byref(u)))
        return r.value, u.value

    def _fits_device(self, peak):
        """The lanes' arenas, what the memory pool holds and the group's
        peak fit the budget (after giving the pool's unused memory back if
        that is what stands in the way)."""
        arenas = self.lanes * self.lane_bytes
        res, used = self._pool_usage()
        if arenas + res + peak <= self.budget:
            return True
        if arenas + used + peak > self.budget:
            return False
        self.psc._check("psc_pool_trim", self.psc.lib.psc_pool_trim(0))
        self.trims += 1
        res, used = self._pool_usage()
        return arenas + res + peak <= self.budget

    def _layout(self, reqs):
        n = sum(r.n for r in reqs)
        m = sum(r.m or gen_bound_m(r.kind, r.n) for r in reqs)
        return int(self.psc.lib.psc_batch_layout_bytes(n, m, len(reqs)))

    # ---------------------------------------------------------- admission
    def _admit_loop(self, reqs, t0):
        try:
            self._admit(reqs, t0)
        except Exception as e:   # surfaced by serve(); the lanes stop
            self._abort(e, len(reqs))

    def _admit(self, reqs, t0):
        """Admit every request that has arrived while its generation's peak
        fits the budget; one generator call per admission tick."""
        i = 0
        total = len(reqs)
        lane_cap = LANE_FILL * self.lane_bytes
        while i < total:
            if self.errors:
                return
            now = time.perf_counter() - t0
            if reqs[i].arrival > now:
                time.sleep(reqs[i].arrival - now)
                continue
            group, peak, keep = [], 0, 0
            err = None
            with self.cv:
                while i < total and reqs[i].arrival <= time.perf_counter() - t0:
                    r = reqs[i]
                    if self._layout([r]) > lane_cap:
                        err = RuntimeError(f"request {r.rid} (n={r.n}) does not fit a lane "
                                           f"arena of {self.lane_bytes} B")
                        break
                    k2, p2 = gen_bytes(self.psc, group + [r])
                    if self.reserved + p2 > self.budget or not self._fits_device(p2):
                        if group:
                            break
                        if p2 > self.budget - self.lanes * self.lane_bytes:
                            err = RuntimeError(f"request {r.rid} needs {p2} B to generate, "
                                               f"over the budget")
                            break
                        if self.reserved == self.lanes * self.lane_bytes:
                            # nothing of ours is left to free: the memory
                            # pool is held by allocations outside serving
                            err = RuntimeError(f"request {r.rid} needs {p2} B to generate; the "
                                               f"device memory pool holds "
                                               f"{self._pool_usage()[1]} B of other "
                                               f"allocations")
                            break
                        # wait for memory to free up; a failed lane ends
                        # the run (its error is raised by serve())
                        self.cv.wait(timeout=0.5)
                        if self.errors:
                            return
                        continue
                    group.append(r)
                    keep, peak = k2, p2
                    i += 1
                    if len(group) >= self.max_admit:
                        break
                if group:
                    self._reserve_now(peak)
            if err is not None:
                if group:
                    self._unreserve(peak)
                with self.cv:
                    self.finished += len(group)
                self._abort(err, total)
                return
            if not group:
                continue
            tg = time.perf_counter()
            try:
                db = self.psc.DeviceBatch(
                    np.array([r.kind for r in group], np.uint32),
                    np.array([r.n for r in group], np.uint64),
                    seeds=np.array([r.seed for r in group], np.uint64))
            except Exception as e:
                self._unreserve(peak)
                with self.cv:
                    self.finished += len(group)
                self._abort(e, total)
                return
            self._sample_device()
            self._unreserve(peak - keep)   # the generator's transients are gone
            shared = _Pool(self, db, len(group), keep)
            adm = time.perf_counter() - t0
            self.phase["generate"] += adm + t0 - tg
            self.calls["generate"] += 1
            with self.cv:
                for j, r in enumerate(group):
                    r.admitted = adm
                    r.m = int(db.ebase[j + 1] - db.ebase[j])
                    self.ready.append((r, shared, j))
                self.cv.notify_all()

    # --------------------------------------------------------------- lanes
    def _take(self):
        """The oldest request (plain), or the ready requests in order while
        their binding fits the lane's arena share (batched)."""
        if self.mode == "plain":
            return [self.ready.pop(0)]
        take, nv = [], 0
        cap = LANE_FILL * self.lane_bytes
        while self.ready:
            r = self.ready[0][0]
            if take and (nv + r.n > self.max_batch_vertices or
                         self._layout([x[0] for x in take] + [r]) > cap):
                break
            nv += r.n
            take.append(self.ready.pop(0))
        return take

    def _release(self, items):
        """Give back the pool references of finished (or failed) items."""
        for _, pool, _ in items:
            pool.release()
        with self.cv:
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
            h = C.c_void_p()
            psc._check("psc_lane_create", psc.lib.psc_lane_create(self.lane_bytes, C.byref(h)))
            handle = psc.BatchSCC._wrap(h)
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
                self._run_split(handle, items, pinned, t0)
                done, items = items, []
                self._release(done)
                self._sample_device()
        except Exception as e:   # surfaced by serve()
            failed, items = items, []
            self._release(failed)
            self._abort(e, total)
        finally:
            if handle is not None:
                handle.close()
            if pinned is not None:
                pinned.free()

    def _run_split(self, handle, items, pinned, t0):
        """Run the items as one batch; one that outgrows the lane's arena is
        split in halves (each graph's labels do not depend on its batch)."""
        try:
            self._run_batch(handle, items, pinned, t0)
        except self.psc.PscError as e:
            if e.code != -3 or len(items) == 1:
                raise
            self.splits += 1
            half = len(items) // 2
            self._run_split(handle, items[:half], pinned, t0)
            self._run_split(handle, items[half:], pinned, t0)

    def _run_batch(self, handle, items, pinned, t0):
        """Bind the lane to the items' graphs (views into their admission
        batches), run the batched SCC, hand out the labels."""
        psc = self.psc
        start = time.perf_counter() - t0
        views = [pool.db.view(j) for _, pool, j in items]
        arr = (psc.CGraph * len(views))(*[v[0] for v in views])
        vb = np.array([v[1] for v in views], dtype=np.uint64)
        eb = np.array([v[2] for v in views], dtype=np.uint64)
        vbp, ebp = psc._ptr(vb, psc.u64p), psc._ptr(eb, psc.u64p)
        psc._check("psc_batch_rebind",
                   psc.lib.psc_batch_rebind(handle._h, arr, len(views), vbp, ebp))
        handle._refresh()
        t1 = time.perf_counter()
        betas = np.array([r.beta for r, _, _ in items], dtype=np.float64)
        labels = handle.run(beta_per_graph=betas, tau=self.tau,
                            dense_div=self.dense_div, timing=0, capture=self.capture,
                            out=pinned.view(handle.n))
        done = time.perf_counter() - t0
        scc_s = handle.stats["ms_total"] / 1e3
        self.phase["gather"] += t1 - t0 - start
        self.calls["batches"] += 1
        self.phase["scc"] += scc_s
        self.phase["labels"] += done + t0 - t1 - scc_s
        parts = handle.split(labels)
        for (r, _, _), lab in zip(items, parts):
            r.started, r.done, r.batch = start, done, len(items)
            if self._keep(r):
                r.labels = lab.copy()

    def serve(self, reqs):
        """Run the request stream; returns a summary dict."""
        reqs = sorted(reqs, key=lambda r: r.arrival)
        if self.lanes * self.lane_bytes >= self.budget:
            raise ValueError("the lanes' arenas take the whole budget")
        # host-side admission control before any device work: every request
        # must fit a lane's arena share on its own
        cap = LANE_FILL * self.lane_bytes
        for r in reqs:
            if self._layout([r]) > cap:
                raise RuntimeError(f"request {r.rid} (n={r.n}) does not fit a lane "
                                   f"arena of {self.lane_bytes} B")
        self.dev_base = self._sample_device()
        self.peak_dev = self.dev_base or 0
        with self.cv:
            self._reserve_now(self.lanes * self.lane_bytes)
        t0 = time.perf_counter()
        adm = threading.Thread(target=self._admit_loop, args=(reqs, t0))
        lanes = [threading.Thread(target=self._lane_loop, args=(len(reqs), t0))
                 for _ in range(self.lanes)]
        adm.start()
        for th in lanes:
            th.start()
        if self.progress_s > 0:   # a progress line every progress_s seconds
            import sys
            alive = [adm] + lanes
            while alive:
                deadline = time.perf_counter() + self.progress_s
                for th in alive:
                    th.join(timeout=max(0.0, deadline - time.perf_counter()))
                alive = [th for th in alive if th.is_alive()]
                print(f"serve[{self.mode}]: {self.finished}/{len(reqs)} done after "
                      f"{time.perf_counter() - t0:.0f}s", file=sys.stderr, flush=True)
        adm.join()
        for th in lanes:
            th.join()
        wall = time.perf_counter() - t0
        self._unreserve(self.lanes * self.lane_bytes)
        if self.errors:
            raise self.errors[0]
        lat = np.array([r.done - r.arrival for r in reqs])
        gb = 2**30
        return {
            "mode": self.mode, "requests": len(reqs), "wall_s": round(wall, 4),
            "graphs_per_s": round(len(reqs) / wall, 3),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "mean_batch": round(float(np.mean([r.batch for r in reqs])), 2),
            "budget_GB": round(self.budget / gb, 2),
            "lane_arena_GB": round(self.lane_bytes / gb, 2),
            "peak_reserved_GB": round(self.peak_reserved / gb, 2),
            "device_used_before_GB": round((self.dev_base or 0) / gb, 2),
            "peak_device_used_GB": round(self.peak_dev / gb, 2),
            "splits": self.splits,
            "pool_trims": self.trims,
            "phase_s": {k: round(v, 3) for k, v in self.phase.items()},
            "calls": dict(self.calls),
        }
