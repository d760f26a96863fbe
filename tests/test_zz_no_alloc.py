"""The allocation-free entry point (psc_run_noalloc) through the C ABI:
its labels equal psc_run_device's on the same device graph, bit for bit, on
two calls in a row, the second with a smaller graph (the scratch array
keeps the first call's contents); limits and overflow fail with a message."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def labels_device(psc, dg, **kw):
    s = psc.SCC(dg)
    out = psc.DeviceArray(dg.n)
    s.scc_device(out.ptr, **kw)
    lab = out.numpy()
    out.free()
    s.close()
    return lab


def labels_noalloc(psc, dg, **kw):
    out = psc.DeviceArray(dg.n)
    st = psc.scc_noalloc(dg, out.ptr, **kw)
    lab = out.numpy()
    out.free()
    return lab, st


@pytest.mark.parametrize("kw", [{}, {"tail_max": 0}, {"canonical": 1, "tau": 1}])
def test_noalloc_equals_run_device(psc, oracle, kw):
    big = psc.DeviceGraph.rmat(18, 16 << 18, seed=31)
    small = psc.DeviceGraph.grid(200, 150, 0.6, seed=32)
    for dg in (big, small):
        print("noalloc", dg.n, dg.m, kw, flush=True)
        want = labels_device(psc, dg, **kw)
        got, st = labels_noalloc(psc, dg, **kw)
        assert np.array_equal(got, want)
        if not kw.get("canonical"):
            h = dg.to_host()
            assert np.array_equal(oracle.canonical(got), oracle.tarjan_canonical(h))
    assert 0 < psc.lib.psc_noalloc_scratch_peak() <= psc.lib.psc_noalloc_scratch_bytes()
    big.free()
    small.free()


def test_noalloc_limits(psc):
    dg = psc.DeviceGraph.rmat(21, 4 << 21, seed=33)   # n = 2^21 > 2^20
    out = psc.DeviceArray(dg.n)
    with pytest.raises(psc.PscError, match="exceed the limits"):
        psc.scc_noalloc(dg, out.ptr)
    out.free()
    dg.free()
    # forced tiny tables double until the scratch no longer holds them:
    # a clean PSC_ERR_NOMEM, and the next call works again
    dg = psc.DeviceGraph.rmat(16, 16 << 16, seed=34)
    out = psc.DeviceArray(dg.n)
    with pytest.raises(psc.PscError, match="scratch exhausted"):
        psc.scc_noalloc(dg, out.ptr, tail_max=0, table_init=1 << 27)
    st = psc.scc_noalloc(dg, out.ptr)
    lab = out.numpy()
    out.free()
    assert np.array_equal(lab, labels_device(psc, dg))
    dg.free()
