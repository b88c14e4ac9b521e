"""The correctly rounded acos / cos / pow(x, 1/3) that the cubic solver falls back to
when a root lands on a segment end (diffvg_amd/csrc/dvg_crmath.h), built for the host.

Pinned two ways: against high-precision values (mpmath at 200 bits, rounded once), which
must agree on every input; and against the host libm the reference is built with, which
must agree on all but a small fraction of inputs (the host libm itself is not correctly
rounded everywhere)."""
import ctypes
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, 'csrc', 'libcrmath_host.so')


def _lib():
    if not os.path.exists(SO):
        pytest.skip('tests/csrc/libcrmath_host.so not built (make -C tests/csrc)')
    L = ctypes.CDLL(SO)
    L.crm_eval.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    return L


def _eval(fn, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    _lib().crm_eval(fn, x.ctypes.data, y.ctypes.data, len(x))
    return y


def _inputs(fn, n, seed):
    rng = np.random.default_rng(seed)
    if fn == 0:  # cos: the cubic solver's arguments lie in [-2.1, 3.2]
        return rng.uniform(-3.3, 3.3, n)
    if fn == 1:
        return np.concatenate([rng.uniform(-1, 1, n - 4), [1 - 2 ** -53, -1 + 2 ** -53, 0.5, -0.5]])
    return np.exp2(rng.uniform(-60, 60, n))


REF = {0: math.cos, 1: math.acos, 2: lambda v: math.pow(v, 1.0 / 3.0)}


@pytest.mark.parametrize('fn', [0, 1, 2])
def test_correctly_rounded_against_high_precision(fn):
    mpmath = pytest.importorskip('mpmath')
    mpmath.mp.prec = 200
    x = _inputs(fn, 400, fn)
    y = _eval(fn, x)
    third = mpmath.mpf(1.0 / 3.0)
    f = {0: mpmath.cos, 1: mpmath.acos, 2: lambda v: mpmath.power(v, third)}[fn]
    for xi, yi in zip(x, y):
        assert yi == float(f(mpmath.mpf(float(xi)))), (fn, float(xi))


@pytest.mark.parametrize('fn', [0, 1, 2])
def test_agrees_with_host_libm(fn):
    x = _inputs(fn, 200000, 10 + fn)
    y = _eval(fn, x)
    ref = np.array([REF[fn](v) for v in x])
    assert np.mean(y != ref) < 3e-3
    assert np.all(np.abs(y - ref) <= np.spacing(np.abs(ref)))  # never more than one ulp apart


def test_special_values():
    assert _eval(1, [1.0, -1.0])[0] == 0.0 and _eval(1, [-1.0])[0] == math.pi
    assert _eval(2, [0.0])[0] == 0.0 and _eval(2, [8.0])[0] == math.pow(8.0, 1.0 / 3.0)
