"""Host-side checks that need no GPU: the native libraries load and export
every declared entry point, the FE generator reproduces the oracle's
cell-by-cell assembly bit for bit, and the GPU path fails loudly when no
device is present."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from ba_amd import hflmp
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hflmp_\w+)\s*\(", text)))


def exported(lib_path):
    out = subprocess.check_output(["nm", "-D", "--defined-only", lib_path], text=True)
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.mark.parametrize("header,lib", [("hflmp_hip.h", hflmp.HIP_LIB),
                                        ("hflmp_gen.h", hflmp.GEN_LIB)])
def test_library_exports_every_declared_symbol(header, lib):
    names = declared_functions(header)
    assert len(names) >= 5
    missing = [n for n in names if n not in exported(lib)]
    assert not missing, missing
    L = C.CDLL(lib)
    for n in names:
        getattr(L, n)


def test_lmp_library_exports():
    path = os.path.join(ROOT, "ba_amd", "lib", "libhflmp_lmp.so")
    if not os.path.exists(path):
        pytest.skip("libhflmp_lmp.so not built")
    names = declared_functions("hflmp_lmp.h")
    missing = [n for n in names if n not in exported(path)]
    assert not missing, missing


@pytest.mark.parametrize("order,dim,n,sigma", [(1, 2, 9, 0.0), (1, 3, 6, 0.0), (2, 2, 7, 0.0),
                                               (1, 2, 11, 1.0), (1, 3, 7, 1.0), (2, 2, 5, 1.0),
                                               (1, 3, 1, 0.0)])
def test_generator_bit_identical_to_oracle_assembly(order, dim, n, sigma):
    kap = hflmp.fe_kappa(dim, n, sigma, 123) if sigma else None
    if sigma:
        assert np.array_equal(kap, O.fe_kappa(dim, n, sigma, 123))
    got = hflmp.fe_system(order, dim, n, kap)
    ref = O.fe_system(order, dim, n, kap)
    for a, b in zip(got, ref):
        assert a.dtype == b.dtype and np.array_equal(a, b)
    assert np.array_equal(hflmp.fe_rhs(order, dim, n, 77), O.fe_rhs(order, dim, n, 77))


def test_generator_sizes_of_benchmark_configs():
    assert hflmp.fe_sizes(1, 2, 256) == (66049, 591361)
    assert hflmp.fe_sizes(1, 3, 128) == (2146689, 57066625)
    assert hflmp.fe_sizes(2, 2, 512) == (1050625, 16785409)
    assert hflmp.fe_sizes(1, 3, 96) == (912673, 24137569)
    with pytest.raises(ValueError):
        hflmp.fe_sizes(2, 3, 4)


def test_no_gpu_fails_loudly():
    if hflmp.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(hflmp.HflmpError):
        hflmp.Context(0)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "single" in r.stderr


def test_serve_tune_choose_window():
    from ba_amd import serve_tune as T
    w64, big, all_ = (0, 0), (1 << 20, 0), (1, 0)
    # the fastest window, when it beats FP64 by more than the margin
    assert T.choose_window({w64: 10.0, big: 7.9, all_: 7.5}) == all_
    # a gain within the margin keeps FP64
    assert T.choose_window({w64: 10.0, big: 9.5, all_: 9.2}) == w64
    # a window whose sample did not converge is never chosen
    assert T.choose_window({w64: 10.0, big: None, all_: 8.5}) == all_
    assert T.choose_window({w64: 10.0, all_: None}) == w64
    # FP64 failing: the best converged window, whatever the margin
    assert T.choose_window({w64: None, big: 9.9}) == big
    assert T.choose_window({w64: None, big: None}) == w64
    assert T.WINDOWS[0] == w64 and set(T.WINDOWS[1:]) == {
        (1, 0), (1 << 16, 0), (1 << 20, 0), (8 << 20, 0)}
    assert [T.size_class(n) for n in (0, 65535, 65536, 1 << 20, 9 << 20)] == [0, 0, 1, 2, 3]
