"""The C-ABI library: loads, exports what include/lightne.h declares, host-side
logic (graph IO, generators, sampling scalars).  No device compute here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import lightne_amd as L
from lightne_amd import _lib
from oracle import oracle as O
from tests import helpers as H

HEADER = os.path.join(H.ROOT, "include", "lightne.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lightne_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    # the ctypes signature table covers the whole header
    assert set(names) <= set(_lib.exported_symbols())
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (lightne_\w+)", out))
    assert set(names) <= exported


_STRUCTS = [("lightne_graph", _lib.Graph), ("lightne_sparsify_params", _lib.SparsifyParams),
            ("lightne_sparsify_stats", _lib.SparsifyStats), ("lightne_rsvd_params", _lib.RsvdParams),
            ("lightne_rsvd_stats", _lib.RsvdStats), ("lightne_prop_params", _lib.PropParams),
            ("lightne_prop_stats", _lib.PropStats), ("lightne_netsmf_stats", _lib.NetsmfStats)]


def test_ctypes_structs_match_the_header(tmp_path):
    """Every ctypes mirror has the C struct's size and field offsets (a C
    program built against include/lightne.h prints them)."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "lightne.h"', "int main(void) {"]
    for cname, py in _STRUCTS:
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(H.ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                          check=True).stdout.splitlines())
    for cname, py in _STRUCTS:
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, f"{cname}.{f}"


def test_library_is_a_gfx950_hip_binary():
    out = subprocess.run(["strings", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_no_device_is_an_error_not_an_exit():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is visible; covered by the GPU tests")
    lib = _lib.load()
    ctx = C.c_void_p()
    rc = lib.lightne_ctx_create(0, None, C.byref(ctx))
    assert rc == _lib.ERR_NO_DEVICE
    assert "no HIP device" in _lib.global_error()


def test_adjacency_graph_round_trip(tmp_path):
    off, edg = H.golden_graph("toy")
    g = L.HostGraph(off, edg)
    p = str(tmp_path / "toy.adj")
    L.write_adjacency_graph(p, g)
    txt = open(p).read().split()
    assert txt[0] == "AdjacencyGraph" and int(txt[1]) == g.n and int(txt[2]) == g.m
    g2 = L.read_unweighted_symmetric_graph(p)
    np.testing.assert_array_equal(g2.offsets, off)
    np.testing.assert_array_equal(g2.edges, edg)


def test_adjacency_reader_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.adj"
    p.write_text("AdjacencyGraph\n3\n4\n0\n2\n")
    with pytest.raises(L.LightNEError) as e:
        L.read_unweighted_symmetric_graph(str(p))
    assert e.value.code == _lib.ERR_IO
    p.write_text("EdgeList\n")
    with pytest.raises(L.LightNEError):
        L.read_unweighted_symmetric_graph(str(p))
    with pytest.raises(L.LightNEError):
        L.read_unweighted_symmetric_graph(str(tmp_path / "missing.adj"))


def _check_simple_symmetric(g):
    n = g.n
    rows = np.repeat(np.arange(n), np.diff(g.offsets))
    assert np.all(rows != g.edges)  # no self loops
    key = rows.astype(np.int64) * n + g.edges
    assert np.all(np.diff(key) > 0)  # sorted, no duplicates
    rkey = np.sort(g.edges.astype(np.int64) * n + rows)
    np.testing.assert_array_equal(np.sort(key), rkey)  # symmetric


def test_chung_lu_generator_shape_and_determinism():
    g = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    _check_simple_symmetric(g)
    d = g.degrees()
    assert 55 < d.mean() < 70 and 2500 < d.max() < 4500
    g2 = L.chung_lu(10312, 2.1, 65.0, 3900.0, 1)
    np.testing.assert_array_equal(g.edges, g2.edges)
    g3 = L.chung_lu(10312, 2.1, 65.0, 3900.0, 2)
    assert not np.array_equal(g.offsets, g3.offsets)


def test_rmat_generator_shape_and_determinism():
    g = L.rmat(12, 8.0, 0.57, 0.19, 0.19, 3)
    _check_simple_symmetric(g)
    assert g.n == 4096 and 0.5 * 2 * 8 * 4096 < g.m <= 2 * 8 * 4096
    g2 = L.rmat(12, 8.0, 0.57, 0.19, 0.19, 3)
    np.testing.assert_array_equal(g.edges, g2.edges)


@pytest.mark.parametrize("n,m,T,ratio,mem,tsz", [(10312, 667966, 10, 1.0, 1.0, 0), (1140000, 5980000, 10, 0.37, 1.0, 0),
                                                 (200, 1234, 5, 2.0, 0.5, 5000000), (7, 12, 1, 3.3, 1.0, 0)])
def test_sample_scalars_match_oracle(n, m, T, ratio, mem, tsz):
    got = L.sample_scalars(n, m, T, ratio, mem, tsz)
    ref = O.sample_scalars(n, m, T, ratio, mem, tsz)
    for k in ["logn", "M", "one_over_lambda", "fractional_bits", "exp_rv_rounded", "table_size"]:
        assert got[k] == ref[k], k


def test_bessel_coefficients_match_oracle():
    for k in range(10):
        assert L.bessel_i(k, 0.5) == O.bessel_iv(k, 0.5)


CLI = os.path.join(H.ROOT, "lightne_amd", "lightne")


def test_cli_usage_and_errors(tmp_path):
    out = subprocess.run([CLI], capture_output=True, text=True)
    assert out.returncode == 2 and "usage" in out.stdout
    out = subprocess.run([CLI, "-rounds", "1", "-s", str(tmp_path / "missing.adj")], capture_output=True, text=True)
    assert out.returncode == 1 and "cannot open" in out.stderr
    out = subprocess.run([CLI, "-rounds", "1", "-c", "-s", str(tmp_path / "missing.cmp")], capture_output=True,
                         text=True)
    assert out.returncode == 1 and "cannot open" in out.stderr
    (tmp_path / "short.cmp").write_bytes(b"\x01" * 10)
    out = subprocess.run([CLI, "-rounds", "1", "-c", "-s", str(tmp_path / "short.cmp")], capture_output=True,
                         text=True)
    assert out.returncode == 1 and "not a compressed graph" in out.stderr
    out = subprocess.run([CLI, "-ith_rule", "nope", "-s", str(tmp_path / "x.adj")], capture_output=True, text=True)
    assert out.returncode == 2
    out = subprocess.run([CLI, "-ne_method", "nope", "-s", str(tmp_path / "x.adj")], capture_output=True, text=True)
    assert out.returncode == 2


def test_sample_scalars_without_edges_and_out_of_range():
    """m = 0 draws nothing (the reference divides by zero); an absurd budget is an
    error, not a runaway table size."""
    st = L.sample_scalars(300, 0, walk_len=10, sample_ratio=1.0)
    assert st["one_over_lambda"] == 0.0 and st["exp_rv_rounded"] == 0 and st["fractional_bits"] == 0.0
    with pytest.raises(L.LightNEError):
        L.sample_scalars(10, 2, walk_len=10, sample_ratio=1e30)
