"""The renderer's fast cubic winding and stroke-rejection paths (diffvg_amd/csrc/dvg_winding_fast.h) against the
oracle's restatement of the reference's double-precision solve: on random and adversarial
segments (rays through control-point y values, near-tangent rays, large coordinates,
nearly flat curves, straight segments written as cubics, points inside the control box,
both sides of boundary points as the
edge kernel shades them) every decision it takes must equal the reference's.  Built for the host from the same header the GPU kernels compile."""
import ctypes
import os

import pytest

SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'csrc', 'libwinding_fast_check.so')


@pytest.mark.parametrize('mode, min_decided', [(0, 0.99), (1, 0.5), (2, 0.5), (3, 0.99), (4, 0.9), (5, 0.99), (7, 0.99),
                                               (6, 0.99)])
def test_fast_cubic_winding_never_disagrees(mode, min_decided):
    if not os.path.exists(SO):
        pytest.skip('tests/csrc not built (make -C tests/csrc)')
    L = ctypes.CDLL(SO)
    L.wf_fuzz.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64),
                          ctypes.POINTER(ctypes.c_float)]
    out = (ctypes.c_int64 * 3)()
    bad = (ctypes.c_float * 10)()
    L.wf_fuzz(mode, 400000, 1000 + mode, out, bad)
    assert out[1] == 0, ('mismatch', list(bad))
    assert out[0] / out[2] >= min_decided


@pytest.mark.parametrize('mode, min_decided', [(0, 0.8), (1, 0.9), (2, 0.6), (3, 0.8)])
def test_fast_cubic_stroke_rejection_never_disagrees(mode, min_decided):
    """Whenever the certain-rejection test of a cubic stroke fires, the oracle's
    within-distance test (within_distance.h:119-272) rejects the point too (modes 0-2);
    and the oracle's closest point of the segment (compute_distance.h:127-280) is not
    closer than that radius, so the closest-point search may skip the segment (mode 3)."""
    if not os.path.exists(SO):
        pytest.skip('tests/csrc not built (make -C tests/csrc)')
    L = ctypes.CDLL(SO)
    L.wf_within_fuzz.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_float)]
    out = (ctypes.c_int64 * 3)()
    bad = (ctypes.c_float * 10)()
    L.wf_within_fuzz(mode, 200000, 77 + mode, out, bad)
    assert out[1] == 0, ('mismatch', list(bad))
    assert out[0] / out[2] >= min_decided
