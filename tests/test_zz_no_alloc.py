"""dvg_render_noalloc, the render entry point that allocates no device memory: through
the C ABI, two calls in a row on one GPU (the second smaller than the first, so it runs
on what the first left in the fixed workspace), each compared with dvg_render on a scene
built from the same description -- forward images and SDFs bitwise, gradients up to
atomic summation order -- for colour, prefiltered colour and SDF renders, and the first
forward against the oracle.  A render larger than the fixed workspace is refused with an
error before anything is enqueued."""
import ctypes

import numpy as np
import pytest
import torch

from diffvg_amd import _lib
import scenes
from refdrive import OracleRunner
from test_oracle_pins import GRAD_KEYS, _rel

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda:0')


def _render(scene, noalloc, w, h, nsx, nsy, seed, pre=False, sdf=False, d=None, bg=None):
    """One forward (d is None) or backward through dvg_render or dvg_render_noalloc."""
    out = None
    kw = dict(use_prefiltering=pre, background=bg)
    if d is None:
        out = torch.zeros(h, w, 1 if sdf else 4, device=DEV)
        kw['render_sdf' if sdf else 'render_image'] = out
    else:
        scene.zero_grads(None)
        kw['d_render_sdf' if sdf else 'd_render_image'] = d
        kw['d_translation'] = out = torch.zeros(h, w, 2, device=DEV)
    a = scene.make_args(w, h, nsx, nsy, seed, **kw)
    L = _lib.lib()
    fn = L.dvg_render_noalloc if noalloc else L.dvg_render
    _lib.check(fn(scene.handle, ctypes.byref(a)))
    torch.cuda.synchronize()
    scene.status()
    if d is None:
        return out.cpu().numpy(), None
    return out.cpu().numpy(), scene.grads()


# (case, width, height, nsx, nsy, prefilter, sdf): per case, the second call is smaller
CALLS = {
    'colour': [('strokes', 128, 128, 2, 2, False, False), ('overlap_eo', 40, 24, 2, 3, False, False)],
    'prefiltered': [('mixed_segments', 96, 80, 2, 2, True, False), ('curve_fill', 33, 17, 1, 1, True, False)],
    'sdf': [('transformed', 128, 96, 1, 1, False, True), ('primitives', 50, 30, 1, 1, False, True)],
}


@pytest.mark.parametrize('kind', sorted(CALLS))
def test_noalloc_matches_render_twice(kind):
    for k, (case, w, h, nsx, nsy, pre, sdf) in enumerate(CALLS[kind]):
        print('noalloc', kind, 'call', k, case, w, h, nsx, nsy, flush=True)
        f = scenes.flat_case(case)
        a, b = _lib.DeviceScene(f, 0), _lib.DeviceScene(f, 0)
        seed = 11 + k
        rng = np.random.default_rng(k)
        bg = None
        if not sdf and k == 1:
            bg = torch.from_numpy(rng.uniform(0, 1, (h, w, 4)).astype(np.float32)).to(DEV)
        got, _ = _render(a, True, w, h, nsx, nsy, seed, pre, sdf, bg=bg)
        want, _ = _render(b, False, w, h, nsx, nsy, seed, pre, sdf, bg=bg)
        assert np.array_equal(got, want), (kind, k)
        if k == 0:
            o = OracleRunner(f)
            ref = o.render(w, h, nsx, nsy, seed, prefilter=pre, output='sdf' if sdf else 'color')
            assert np.abs(got - ref['sdf' if sdf else 'image']).max() <= 1e-5, (kind, k)
        d = torch.from_numpy(rng.uniform(-1, 1, (h, w, 1 if sdf else 4)).astype(np.float32)).to(DEV)
        tr_a, ga = _render(a, True, w, h, nsx, nsy, seed, pre, sdf, d=d, bg=bg)
        tr_b, gb = _render(b, False, w, h, nsx, nsy, seed, pre, sdf, d=d, bg=bg)
        assert _rel(tr_a, tr_b) <= 1e-4, (kind, k, 'd_translation')
        for key in GRAD_KEYS:
            if gb[key].size:
                assert _rel(ga[key], gb[key]) <= 1e-4, (kind, k, key)


def test_noalloc_refuses_renders_beyond_its_workspace():
    f = scenes.flat_case('single_circle_64')
    sc = _lib.DeviceScene(f, 0)
    L = _lib.lib()
    for w, h, nsx, nsy in ((1024, 1024, 1, 1), (512, 512, 2, 2), (16, 16, 17, 16)):
        a = sc.make_args(w, h, nsx, nsy, 0, render_image=None)
        rc = L.dvg_render_noalloc(sc.handle, ctypes.byref(a))
        assert rc != 0
        msg = L.dvg_last_error().decode()
        assert 'dvg_render_noalloc' in msg, msg
