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


def fake_psc():
    return types.SimpleNamespace(DeviceBatch=FakeBatch, PinnedLabels=FakePinned)


@pytest.mark.parametrize("lanes", [1, 3])
def test_lane_error_ends_run_and_frees_memory(psc, lanes):
    import psc_amd.serve as sv
    FakeBatch.live = 0
    reqs = [sv.Request(i, 0, 1 << 12, i, 0.0) for i in range(40)]
    need = sv.est_bytes(0, 1 << 12)
    # room for three requests at a time: admission must wait for memory
    srv = sv.Server(fake_psc(), budget_bytes=3 * need + 1, lanes=lanes,
                    mode="batched", max_wait_s=0.0)
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
    assert srv.in_use == 0
    assert FakeBatch.live == 0
    assert srv.finished == len(reqs)


def test_request_over_budget_raises(psc):
    import psc_amd.serve as sv
    FakeBatch.live = 0
    reqs = [sv.Request(0, 0, 1 << 20, 1, 0.0)]
    srv = sv.Server(fake_psc(), budget_bytes=1 << 20, lanes=2)
    with pytest.raises(RuntimeError, match="over the"):
        srv.serve(reqs)
    assert srv.in_use == 0 and FakeBatch.live == 0
