"""Exact float64 summation of the join accumulators (vexq_amd/csrc/fixpoint.h).

The join kernels sum float64 aggregate arguments as three fixed-point integer
limbs and round once per group.  These tests build the same header for the
host and check it against exact rational arithmetic (fractions.Fraction, whose
conversion to float is correctly rounded): the result must be the float64
nearest to mult * (exact sum), ties to even, bit for bit, whatever the order
of the addends.  The scale is chosen as the planner chooses it
(vq_api.hip, vq_join_agg): E with bound < 2^E, L = ceil(log2(n + 1)),
w = min(40, 62 - L), lsb = E - 3w.
"""
from __future__ import annotations

import ctypes
import math
import os
import random
import shutil
import subprocess
from fractions import Fraction

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fx(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("fx") / "libfx.so"
    subprocess.run([gxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(out),
                    os.path.join(HERE, "cpp", "fixpoint_check.cpp")], check=True)
    lib = ctypes.CDLL(str(out))
    lib.vqfx_chunks.argtypes = [ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                ctypes.POINTER(ctypes.c_int64)]
    lib.vqfx_to_f64.argtypes = [ctypes.c_int64] * 3 + [ctypes.c_int32, ctypes.c_int32,
                                                       ctypes.c_uint64]
    lib.vqfx_to_f64.restype = ctypes.c_double
    lib.vqfx_sum.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_int64, ctypes.c_int32,
                             ctypes.c_int32, ctypes.c_uint64]
    lib.vqfx_sum.restype = ctypes.c_double
    return lib


def scale(bound: float, n: int):
    m = bound * (1 + 1e-9)
    E = math.frexp(m)[1] if m > 0 else 0
    L = max(0, math.ceil(math.log2(n + 1)))
    w = min(40, 62 - L)
    return E - 3 * w, w


def fsum(lib, xs, lsb, w, mult=1):
    a = np.ascontiguousarray(xs, dtype=np.float64)
    return lib.vqfx_sum(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(a), lsb, w, mult)


def exact(xs, mult=1):
    return float(sum((Fraction(float(x)) for x in xs), Fraction(0)) * mult)


def bits(x: float) -> int:
    return int(np.float64(x).view(np.int64))


def test_chunks_recompose(fx):
    rng = random.Random(7)
    c = (ctypes.c_int64 * 3)()
    for _ in range(2000):
        x = rng.uniform(-1e5, 1e5) * 10 ** rng.randint(-3, 0)
        lsb, w = scale(1e5, 1000)
        assert fx.vqfx_chunks(x, lsb, w, c) == 1
        back = (Fraction(c[0]) + Fraction(c[1]) * 2 ** w + Fraction(c[2]) * 2 ** (2 * w)) * \
            Fraction(2) ** lsb
        assert back == Fraction(x)
        for v in c:
            assert abs(v) < 2 ** w
    assert fx.vqfx_chunks(0.0, -100, 40, c) == 0
    assert fx.vqfx_chunks(-0.0, -100, 40, c) == 0


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sums_match_exact_rounding(fx, seed):
    """Join-shaped addends: price * (1 - discount) at cents resolution."""
    rng = np.random.default_rng(seed)
    n = 5000
    q = rng.integers(1, 51, n)
    price = np.round(q * rng.integers(90000, 10495001, n) / 100.0) / 100.0
    disc = rng.integers(0, 11, n) / 100.0
    x = price * (1.0 - disc)
    lsb, w = scale(float(np.abs(x).max()), n)
    for mult in (1, 3, 7):
        got = fsum(fx, x, lsb, w, mult)
        assert bits(got) == bits(exact(x, mult))
    # order independence: any permutation gives the same bits
    perm = rng.permutation(n)
    assert bits(fsum(fx, x[perm], lsb, w)) == bits(fsum(fx, x, lsb, w))


def test_mixed_signs_and_cancellation(fx):
    rng = np.random.default_rng(11)
    x = rng.normal(0, 1e3, 4000)
    x = np.concatenate([x, -x])  # cancels exactly
    rng.shuffle(x)
    lsb, w = scale(float(np.abs(x).max()), len(x))
    assert fsum(fx, x, lsb, w) == 0.0
    y = np.concatenate([x, [1e-3, -2.5e-4]])
    lsb, w = scale(float(np.abs(y).max()), len(y))
    assert bits(fsum(fx, y, lsb, w)) == bits(exact(y))


def test_ties_round_to_even(fx):
    # 2^53 + 1 is halfway between 2^53 and 2^53 + 2: rounds to even (2^53);
    # 2^53 + 3 is halfway between 2^53 + 2 and 2^53 + 4: rounds to 2^53 + 4.
    big = float(2 ** 53)
    for extra, want in ((1.0, 2.0 ** 53), (3.0, 2.0 ** 53 + 4)):
        xs = [big, extra]
        lsb, w = scale(big, 2)
        assert fsum(fx, xs, lsb, w) == want
        assert exact(xs) == want
    # just above the halfway point rounds up (the sticky bit)
    xs = [big, 1.0, 2.0 ** -20]
    lsb, w = scale(big, 3)
    assert fsum(fx, xs, lsb, w) == big + 2


def test_naive_sum_differs_but_exact_does_not(fx):
    """A case where left-to-right float64 addition loses the small terms."""
    xs = [1e16, 1.0, 1.0, 1.0, 1.0, -1e16]
    lsb, w = scale(1e16, len(xs))
    naive = 0.0
    for v in xs:
        naive += v
    assert naive != 4.0
    assert fsum(fx, xs, lsb, w) == 4.0


def test_multiplicity_scaling(fx):
    rng = np.random.default_rng(5)
    x = rng.uniform(0.5, 5e4, 777)
    lsb, w = scale(float(x.max()), len(x))
    for mult in (1, 2, 5, 1000, 2 ** 31 - 1):
        assert bits(fsum(fx, x, lsb, w, mult)) == bits(exact(x, mult))


def test_to_f64_limbs_directly(fx):
    # value = (l0 + l1 2^w + l2 2^2w) 2^lsb with negative limbs (carry-save form)
    w, lsb = 40, -80
    for l0, l1, l2 in ((5, 0, 0), (-(2 ** 50), 3, 0), (2 ** 61, -(2 ** 61), 7), (-1, -1, -1)):
        want = float((Fraction(l0) + Fraction(l1) * 2 ** w + Fraction(l2) * 2 ** (2 * w)) *
                     Fraction(2) ** lsb)
        assert fx.vqfx_to_f64(l0, l1, l2, lsb, w, 1) == want
