"""The reference-side binding (include/scc_gpu.hpp): it compiles against a
Graph with the members of src/graph.hpp:35-43 and links against libpsc.so
(CPU); on the GPU, the program built from it gives Tarjan's partition."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "parallel-strong-connectivity_amd")
SRC = os.path.join(ROOT, "tests", "cpp", "shim_driver.cpp")


def build_driver(out_dir):
    exe = os.path.join(str(out_dir), "shim_driver")
    cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror",
           f"-I{os.path.join(ROOT, 'include')}", SRC, "-o", exe,
           f"-L{PKG}", "-lpsc", f"-Wl,-rpath,{PKG}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_shim_compiles_and_links(tmp_path):
    assert os.path.exists(build_driver(tmp_path))


@pytest.mark.gpu
def test_shim_runs_scc(psc, oracle, tmp_path):
    exe = build_driver(tmp_path)
    g = psc.gen_rmat(13, 8 << 13, seed=12)
    path = tmp_path / "g.bin"
    g.write_bin(str(path), large=True)
    out = tmp_path / "labels.u64"
    print("running", exe, flush=True)
    r = subprocess.run([exe, str(path), str(out), "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "average first_round_time" in r.stdout
    labels = np.fromfile(out, dtype=np.uint64)
    assert np.array_equal(labels, psc.scc(g))
    assert np.array_equal(oracle.canonical(labels), oracle.tarjan_canonical(g))
