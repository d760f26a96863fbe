"""Serving control flow on the CPU (no device work): a lane that fails
ends the run with its error, gives back the memory of its batch and of
every waiting request, and no thread is left waiting.  The device calls
are replaced by stand-ins that count pool releases."""
import threading
import types

import numpy as np
import pytest


class FakeBatch:
    live = 0
    lock = threading.Lock()

    def __init__(self, kind, n, seeds=None):
        self.n = list(n)
        self.ebase = [0] * (len(self.n) + 1)
        with FakeBatch.lock:
            FakeBatch.live += 1

    def view(self, j):
        raise RuntimeError("lane failure injected")

    def free(self):
        with FakeBatch.lock:
            FakeBatch.live -= 1


class FakePinned:
    def free(self):
        pass


class FakeLib:
    """The library calls the server makes, without a device: sizes come from
    simple formulas, a lane handle is a token."""

    def psc_lane_create(self, nbytes, out):
        return 0

    def psc_batch_layout_bytes(self, n, m, g):
        return 64 * n + 8 * m

    def psc_dev_gen_batch_bytes(self, g, kind, n, keep, peak):
        keep._obj.value = 1 << 20
        peak._obj.value = 2 << 20
        return 0

    def psc_device_mem_info(self, free, total):
        return -2   # no driver figure on the host

    def psc_pool_usage(self, reserved, used):
        reserved._obj.value = used._obj.value = 0
        return 0

    def psc_pool_trim(self, keep):
        return 0

    def psc_destroy(self, h):
        pass


class FakeHandle:
    _h = None

    def close(self):
        pass


def fake_psc(real):
    return types.SimpleNamespace(
        DeviceBatch=FakeBatch, PinnedLabels=FakePinned, lib=FakeLib(),
        _check=real._check, _ptr=real._ptr, u32p=real.u32p, u64p=real.u64p,
        PscError=real.PscError,
        BatchSCC=types.SimpleNamespace(_wrap=lambda h: FakeHandle()))


@pytest.mark.parametrize("lanes", [1, 3])
def test_lane_error_ends_run_and_frees_memory(psc, lanes):
    import psc_amd.serve as sv
    FakeBatch.live = 0
    reqs = [sv.Request(i, 0, 1 << 12, i, 0.0) for i in range(40)]
    # lanes of 4 MiB; room for three admission groups at a time (2 MiB peak
    # each): admission must wait for memory
    lane = 4 << 20
    srv = sv.Server(fake_psc(psc), budget_bytes=lanes * lane + (6 << 20) + 1, lanes=lanes,
                    lane_bytes=lane, mode="batched", max_wait_s=0.0, max_admit=1)
    done = {}

    def run():
        try:
            srv.serve(reqs)
        except RuntimeError as e:
            done["err"] = e

    th = threading.Thread(target=run)
    th.start()
    th.join(timeout=30)
    assert not th.is_alive(), "serve() did not return after a lane error"
    assert "lane failure injected" in str(done.get("err"))
    assert srv.reserved == 0
    assert FakeBatch.live == 0
    assert srv.finished == len(reqs)


def test_request_over_budget_raises(psc):
    import psc_amd.serve as sv
    FakeBatch.live = 0
    reqs = [sv.Request(0, 0, 1 << 10, 1, 0.0)]
    # the lanes leave 1 MiB: a 2 MiB generation can never be admitted
    srv = sv.Server(fake_psc(psc), budget_bytes=(9 << 20), lanes=2, lane_bytes=4 << 20)
    with pytest.raises(RuntimeError, match="over the budget"):
        srv.serve(reqs)
    assert srv.reserved == 0 and FakeBatch.live == 0


def test_request_over_lane_raises(psc):
    import psc_amd.serve as sv
    FakeBatch.live = 0
    reqs = [sv.Request(0, 0, 1 << 20, 1, 0.0)]
    srv = sv.Server(fake_psc(psc), budget_bytes=(64 << 20), lanes=2, lane_bytes=4 << 20)
    with pytest.raises(RuntimeError, match="does not fit a lane"):
        srv.serve(reqs)
    assert srv.reserved == 0 and FakeBatch.live == 0
