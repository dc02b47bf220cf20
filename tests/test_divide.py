"""The kernels' exact-division helpers (square-wheel_amd/csrc/sw_divide.h) against integer division.

The kernels replace integer divisions by double-precision quotients or reciprocal products with one
correction step (fixed16_div for the edge and triangle slopes, rasterizer.rs fixed16 arithmetic;
div_rcp / mod_rcp for the surface texel's row, column and texture wrap; z_from_inv_z for the
perspective span's z). The same header is compiled here by the host compiler and compared with
Python's exact integers on random and edge inputs, and swept over whole divisor ranges in C++.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "divide_check.cpp")
HDR = os.path.join(os.path.dirname(HERE), "square-wheel_amd", "csrc", "sw_divide.h")
LIB = os.path.join(HERE, "native", "libdivide_check.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(SRC), os.path.getmtime(HDR)):
        subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fwrapv", "-ffp-contract=off", "-o", LIB, SRC],
                       check=True)
    lib = C.CDLL(LIB)
    u64 = C.c_uint64
    lib.sweep_rcp.restype = u64
    lib.sweep_rcp.argtypes = [C.c_uint32, C.c_uint32]
    lib.sweep_fixed16_div.restype = u64
    lib.sweep_fixed16_div.argtypes = [C.c_int32, C.c_int32, C.c_void_p, C.c_uint32]
    lib.sweep_z.restype = u64
    lib.sweep_z.argtypes = [C.c_uint32, C.c_uint32]
    lib.sweep_z_rcp_perturbed.restype = u64
    lib.sweep_z_rcp_perturbed.argtypes = [C.c_uint32, C.c_uint32, C.c_int32]
    for n in ("check_fixed16_div", "check_div_rcp", "check_mod_rcp", "check_z_from_inv_z"):
        getattr(lib, n).restype = None
    return lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _trunc_div(a, b):
    """Integer quotient rounded toward zero (Rust's i64 `/`), in exact Python integers."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def test_fixed16_div_random_and_edges(lib):
    rng = np.random.default_rng(31)
    edges = np.array([0, 1, -1, 2, -2, 32767, -32768, 32768, 65535, 65536, -65536, 2**31 - 1, -2**31, 2**30, -2**30,
                      12345678, -12345678], np.int64)
    x = np.concatenate([rng.integers(-2**31, 2**31, 400000), rng.integers(-2**20, 2**20, 200000),
                        np.repeat(edges, edges.size)]).astype(np.int32)
    y = np.concatenate([rng.integers(-2**31, 2**31, 400000), rng.integers(-2**12, 2**12, 200000),
                        np.tile(edges, edges.size)]).astype(np.int32)
    y[y == 0] = 1
    out = np.zeros_like(x)
    lib.check_fixed16_div(_p(x), _p(y), _p(out), x.size)
    # the result is the i64 quotient cast to i32 (wrapping, as the reference's `as i32`)
    expect = np.array([_trunc_div(int(a) * 65536, int(b)) for a, b in zip(x, y)], dtype=object)
    expect = np.array([((int(e) + 2**31) % 2**32) - 2**31 for e in expect], np.int64).astype(np.int32)
    assert np.array_equal(out, expect)


def test_fixed16_div_dense_sweep(lib):
    ys = np.array([1, -1, 3, 7, -7, 255, 256, 4095, 65535, 65536, 65537, -99991, 2**24 + 1, 2**31 - 1, -2**31], np.int32)
    assert lib.sweep_fixed16_div(-2**21, 2**21, _p(ys), ys.size) == 0


def test_div_and_mod_rcp_random(lib):
    rng = np.random.default_rng(32)
    n = np.concatenate([rng.integers(0, 2**32, 300000, dtype=np.uint64), np.full(1000, 2**32 - 1, np.uint64),
                        np.arange(1000, dtype=np.uint64)]).astype(np.uint32)
    d = np.concatenate([rng.integers(1, 2**30 + 1, 150000, dtype=np.uint64), rng.integers(1, 4096, 150000, dtype=np.uint64),
                        rng.integers(1, 2**30 + 1, 2000, dtype=np.uint64)]).astype(np.uint32)
    q = np.zeros_like(n)
    r = np.zeros_like(n)
    lib.check_div_rcp(_p(n), _p(d), _p(q), _p(r), n.size)
    assert np.array_equal(q, n // d) and np.array_equal(r, n % d)
    r2 = np.zeros_like(n)
    lib.check_mod_rcp(_p(n), _p(d), _p(r2), n.size)
    assert np.array_equal(r2, n % d)


def test_div_and_mod_rcp_every_divisor_to_70000(lib):
    """Every divisor up to 70,000 (texture and surface widths stay far below) plus the top of the
    range, at numerators next to multiples near every k * 2^24 and 2^32 - 1."""
    assert lib.sweep_rcp(1, 70000) == 0
    assert lib.sweep_rcp(2**30 - 3000, 2**30) == 0


def test_z_from_inv_z_every_divisor(lib):
    """inv_z is clamped to [2^8, 2^29] before the shift by 8, so the divisor is in [1, 2^21]."""
    assert lib.sweep_z(1, 1 << 21) == 0
    d = np.array([1, 2, 3, (1 << 21) - 1, 1 << 21, (1 << 24) - 1, 2**31, 2**32 - 1], np.uint32)
    z = np.zeros_like(d)
    lib.check_z_from_inv_z(_p(d), _p(z), d.size)
    assert np.array_equal(z, ((1 << 31) // d.astype(np.uint64)).astype(np.uint32))


def test_z_fast_branch_tolerates_an_inexact_reciprocal(lib):
    """The fast branch (1024 <= div < 2^24) starts from an approximate f32 reciprocal (the device's
    v_rcp_f32, within 1 ulp); with the reciprocal off by up to 2 ulp either way its one correction
    step still lands on 2^31 / div for every divisor."""
    assert lib.sweep_z_rcp_perturbed(1024, (1 << 24) - 1, 2) == 0
