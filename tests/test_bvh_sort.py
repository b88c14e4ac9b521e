"""The device BVH build's sort (diffvg_amd/csrc/dvg_sort.h) reproduces the reference's
std::sort leaf order exactly, ties included: compared with the C++ library's std::sort
on index arrays keyed by small alphabets, sorted/reversed/organ-pipe inputs, and
McIlroy's adversary for median-of-three partitioning (which drives std::sort into its
heap-sort fallback).  Built for the host from the same header the kernels compile."""
import ctypes
import os

import pytest

SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'csrc', 'libsort_check.so')


def _lib():
    if not os.path.exists(SO):
        pytest.skip('tests/csrc not built (make -C tests/csrc)')
    return ctypes.CDLL(SO)


def test_sort_matches_std_sort_on_ties():
    L = _lib()
    out = (ctypes.c_int64 * 3)()
    L.sort_fuzz(ctypes.c_int64(20000), ctypes.c_uint64(12345), out)
    assert out[1] == 0, ('first failing n', out[2])


@pytest.mark.parametrize('n', [17, 64, 1000, 20000])
def test_sort_matches_std_sort_on_adversary(n):
    L = _lib()
    out = (ctypes.c_int64 * 1)()
    L.sort_adversary(n, out)
    assert out[0] == 0
