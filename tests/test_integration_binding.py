"""The reference-side binding shown in the integration notes
(tests/integration/lightne_runner_binding.cpp) compiles and links against the
C-ABI header and library (host compiler only, no GPU needed)."""
import os
import subprocess

import pytest

from tests import helpers as H

SRC = os.path.join(H.ROOT, "tests", "integration", "lightne_runner_binding.cpp")
LIB = os.path.join(H.ROOT, "lightne_amd")


def build_binding(out):
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(H.ROOT, "include"),
           "-I/opt/rocm/include", SRC, "-o", out, "-L" + LIB, "-llightne_amd", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath," + LIB + ":/opt/rocm/lib"]
    return subprocess.run(cmd, capture_output=True, text=True, timeout=300)


@pytest.mark.skipif(not os.path.exists(os.path.join(LIB, "liblightne_amd.so")), reason="library not built")
def test_binding_compiles_and_links(tmp_path):
    r = build_binding(str(tmp_path / "runner"))
    assert r.returncode == 0, r.stderr
    assert "warning" not in r.stderr


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(os.path.join(LIB, "liblightne_amd.so")), reason="library not built")
def test_binding_runs_like_the_api(gpu, tmp_path):
    """The binding's -ne_out / -pro_out files hold the same bits as the Python
    wrapper's NetSMF + chebyshev_expansion with the same flags and seeds."""
    import numpy as np

    import lightne_amd as L

    exe = str(tmp_path / "runner")
    r = build_binding(exe)
    assert r.returncode == 0, r.stderr
    hg = L.chung_lu(3000, 2.1, 12.0, 300.0, 5)
    path = str(tmp_path / "g.adj")
    L.write_adjacency_graph(path, hg)
    ne, pro = str(tmp_path / "ne.bin"), str(tmp_path / "pro.bin")
    ratio = H.c1_sample_ratio(hg.n, hg.m, 10, 20.0)
    r = subprocess.run([exe, "-walklen", "10", "-rank", "64", "-dim", "32", "-order", "10", "-sample_ratio",
                        repr(ratio), "-seed", "11", "-ne_out", ne, "-pro_out", pro, path],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "### Running Time:" in r.stdout
    g = hg.to_device()
    emb, _, _ = L.NetSMF(g, walk_len=10, rank=64, dim=32, sample_ratio=ratio, seed=11, ctx=gpu)
    p, _ = L.chebyshev_expansion(emb, g, 32, 32, ctx=gpu)
    got_ne = np.fromfile(ne, np.float32).reshape(hg.n, 32)
    got_pro = np.fromfile(pro, np.float32).reshape(hg.n, 32)
    assert H.same_bits(got_ne, emb.cpu().numpy())
    assert H.same_bits(got_pro, p.cpu().numpy())
