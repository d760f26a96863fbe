"""GPU: the scc command line on .bin files (src/scc.cpp), read into device
memory by psc_read_bin_device: a one-CSR file is transposed on the GPU, a
-large file uploads both CSRs.  Canonical labels must equal Tarjan and the
-status lines the oracle's summary."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "parallel-strong-connectivity_amd", "bin", "scc")


@pytest.mark.parametrize("large", [False, True])
def test_cli_reads_bin_into_device_memory(psc, oracle, tmp_path, large):
    g = psc.gen_rmat(12, 8 << 12, seed=5)
    path = tmp_path / f"rmat12_{int(large)}.bin"
    g.write_bin(str(path), large=large)
    out = tmp_path / "labels.bin"
    args = [CLI, str(path), "-t", "1", "-canonical", "-status", "-out", str(out)]
    if large:
        args.append("-large")
    print("running", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    labels = np.fromfile(out, dtype=np.uint64)
    expect = oracle.tarjan_canonical(g)
    assert np.array_equal(labels.astype(np.uint32), expect)
    nscc, big = oracle.status(expect)
    assert f"n_scc = {nscc}" in r.stdout
    assert f"Largest StronglyConnectedComponents has {big}" in r.stdout
    assert f"number of vertices: {g.n}, number of edges: {g.m}" in r.stdout
