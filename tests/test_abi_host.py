"""CPU tests of the C-ABI library and the Python front end (no GPU needed).

* libdiffvg_amd.so loads and exports every function include/diffvg_amd.h declares;
* scene validation errors come back through dvg_last_error before any device is touched;
* the front end's scene marshalling (render_pytorch.flatten_args) produces the same
  dvg_scene_desc arrays as the independent flattening used by the checkers;
* without the library, rendering raises instead of falling back to anything else.
"""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import diffvg_amd as pd
from diffvg_amd import _lib
from diffvg_amd.render_pytorch import flatten_args
import scenes
from refdrive import flatten

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'diffvg_amd.h')


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    text = re.sub(r'//[^\n]*', '', text)
    return sorted(set(re.findall(r'\b(dvg_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_the_abi():
    names = declared_functions()
    for must in ('dvg_scene_create', 'dvg_scene_destroy', 'dvg_render', 'dvg_scene_get_grads',
                 'dvg_scene_zero_grads', 'dvg_last_error', 'dvg_abi_version'):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert L.dvg_abi_version() >= 1


def _desc(flat):
    d, keep = _lib.make_desc(flat)
    return d, keep


def _create(flat):
    L = _lib.lib()
    d, keep = _desc(flat)
    h = ctypes.c_void_p()
    rc = L.dvg_scene_create(ctypes.byref(d), 0, ctypes.byref(h))
    return rc, h, L.dvg_last_error().decode()


@pytest.mark.parametrize('mutate, message', [
    (lambda f: f.update(canvas_width=0), 'canvas'),
    (lambda f: f.update(filter_type=9), 'filter type'),
    (lambda f: f.update(filter_radius=0.0), 'filter radius'),
    (lambda f: f.update(filter_radius=float('nan')), 'filter radius'),
    (lambda f: f['shape_type'].__setitem__(0, 7), 'shape type'),
])
def test_scene_validation_errors(mutate, message):
    f = scenes.flat_case('primitives')
    f = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in f.items()}
    mutate(f)
    rc, h, err = _create(f)
    assert rc != 0
    assert message in err


def test_null_arguments():
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.dvg_scene_create(None, 0, ctypes.byref(h)) != 0
    assert 'null' in L.dvg_last_error().decode()
    assert L.dvg_render(None, None) != 0
    assert L.dvg_scene_destroy(None) == 0


@pytest.mark.skipif(torch.cuda.is_available(), reason='checks the no-device error path')
def test_valid_scene_without_device_reports_hip_error():
    rc, h, err = _create(scenes.flat_case('overlap_eo'))
    assert rc != 0 and 'hip' in err.lower()


@pytest.mark.parametrize('case', sorted(scenes.CASES))
def test_flatten_args_matches_independent_flattening(case):
    sc = scenes.CASES[case]()
    w, h = sc['canvas']
    ftype, frad = sc['filter']
    filt = pd.PixelFilter(type=getattr(pd.diffvg.FilterType, ftype), radius=torch.tensor(frad))
    args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'], filt)
    got = flatten_args(args).desc
    ref = flatten(w, h, sc['shapes'], sc['groups'], ftype, frad)
    assert set(got) == set(ref)
    for k, v in ref.items():
        g = got[k]
        if isinstance(v, np.ndarray):
            n = len(np.asarray(g))
            # the checker pads empty arrays to one element
            assert np.array_equal(np.asarray(g), v[:n]) and (n == len(v) or (n == 0 and len(v) == 1)), k
        else:
            assert g == v, k


def test_gradient_map_covers_every_differentiable_argument():
    sc = scenes.gradients()
    args = pd.RenderFunction.serialize_scene(64, 64, sc['shapes'], sc['groups'],
                                             pd.PixelFilter(type=pd.diffvg.FilterType.box,
                                                            radius=torch.tensor(0.5)))
    f = flatten_args(args)
    idx = {e[0] for e in f.grad_map}
    float_args = {i for i, a in enumerate(args) if isinstance(a, torch.Tensor) and a.is_floating_point()
                  and i >= 7}
    assert float_args <= idx


def test_serialize_scene_layout_matches_reference():
    """render_pytorch.py:35-195 layout: 7 header entries, per shape type + params + stroke
    width, per group ids, fill, stroke, even-odd flag, transform, then filter type and radius."""
    sc = scenes.primitives()
    args = pd.RenderFunction.serialize_scene(64, 64, sc['shapes'], sc['groups'],
                                             pd.PixelFilter(type=pd.diffvg.FilterType.box,
                                                            radius=torch.tensor(0.5)))
    assert args[:4] == [64, 64, 4, 4]
    assert args[7] == pd.diffvg.ShapeType.rect
    # rect: type, p_min, p_max, stroke_width
    assert torch.equal(args[8], sc['shapes'][0].p_min)
    n = 7 + (4) + (4) + (4) + (1 + 5 + 1)  # rect, ellipse, circle, polygon (path: 5 fields)
    assert args[n] is not None and torch.equal(args[n], torch.tensor([0]))
    assert args[-2] == pd.diffvg.FilterType.box


def test_non_finite_points_rejected():
    p = pd.Path(num_control_points=torch.tensor([0, 0]),
                points=torch.tensor([[0.0, 0.0], [float('nan'), 1.0], [3.0, 3.0]]), is_closed=True)
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=torch.tensor([1.0, 0, 0, 1]))
    args = pd.RenderFunction.serialize_scene(8, 8, [p], [g])
    with pytest.raises(AssertionError):
        flatten_args(args)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, '_lib', None)
    monkeypatch.setattr(_lib, 'LIB_PATH', os.path.join(ROOT, 'no_such_dir', 'libdiffvg_amd.so'))
    with pytest.raises(RuntimeError, match='not built'):
        _lib.lib()


def test_render_without_gpu_raises():
    if torch.cuda.is_available():
        pytest.skip('a GPU is present')
    sc = scenes.single_circle(32)
    args = pd.RenderFunction.serialize_scene(32, 32, sc['shapes'], sc['groups'])
    with pytest.raises(RuntimeError):
        pd.RenderFunction.apply(32, 32, 2, 2, 0, None, *args)


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    p = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2', '--no-cpu-baseline'],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0 and 'one GPU' in p.stderr


@pytest.mark.parametrize('case', sorted(scenes.CASES))
def test_drop_in_module_builds_the_same_scene(case, monkeypatch):
    """The `diffvg` module shim, given pybind-style objects the way render_pytorch.py
    builds them, produces the same dvg_scene_desc as the independent flattening."""
    from diffvg_amd import diffvg as dv
    from refdrive import RefRunner
    captured = {}

    class Capture:
        def __init__(self, flat, gpu_index=-1, mode=0):
            captured['flat'] = flat

    monkeypatch.setattr(dv._lib, 'DeviceScene', Capture)
    f = scenes.flat_case(case)
    RefRunner(f, dv=dv, use_gpu=True, gpu_index=0)
    got = captured['flat']
    for k, v in f.items():
        if isinstance(v, np.ndarray):
            n = len(np.asarray(got[k]))
            assert np.array_equal(np.asarray(got[k]), v[:n]), k
        else:
            assert got[k] == v, k
