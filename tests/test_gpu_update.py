"""Scene construction on the device (dvg_scene_update) against construction on the host
(dvg_scene_create, which follows scene.cpp with the reference's own std::sort).

A scene created from values V1 and rebuilt on the device from values V2 must hold the
same tables, byte for byte, as a scene created from V2: boxes, boundary lengths,
sampling CDFs, canvas_to_shape, SDF anchors and the BVH leaf visit orders, including
scenes whose sort keys tie (segments with equal box y-centres, group members in one
Morton cell, more than 16 of each, so that the sort's partitioning decides the order).
Renders of the two scenes then agree bitwise in the forward pass.  The autograd front
end reuses pooled scenes of an unchanged topology through this rebuild; its results
must equal those of a fresh scene.
"""
import numpy as np
import pytest
import torch

import diffvg_amd as pd
from diffvg_amd import _lib
from diffvg_amd import render_pytorch as rp
import scenes
from refdrive import flatten

pytestmark = pytest.mark.gpu

FLOAT_KEYS = ('shape_params', 'points', 'thickness', 'stroke_width', 'color_params', 'shape_to_canvas')


def ties(size=64):
    """A polyline of 20 horizontal segments at one height (equal y-centres) and a group of
    20 tiny circles in one Morton cell, next to ordinary shapes."""
    xs = np.linspace(4, 60, 21)
    pts = np.stack([xs, np.full_like(xs, 30.0)], 1)
    line = pd.Path(num_control_points=torch.zeros(20, dtype=torch.int32), points=scenes.T(pts), is_closed=False,
                   stroke_width=scenes.T(1.5))
    dots = [pd.Circle(radius=scenes.T(1.0 + 0.01 * k), center=scenes.T([20.0 + 0.001 * k, 20.0])) for k in range(20)]
    zig = np.array([[8 + 2.5 * k, 50 + (3 if k % 2 else -3)] for k in range(21)], np.float32)
    zz = pd.Path(num_control_points=torch.zeros(20, dtype=torch.int32), points=scenes.T(zig), is_closed=False,
                 stroke_width=scenes.T(1.0))
    shapes = [line] + dots + [zz]
    groups = [pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=None, stroke_color=scenes.T([1, 0, 0, 1])),
              pd.ShapeGroup(shape_ids=torch.arange(1, 21), fill_color=scenes.T([0, 0.5, 1, 0.8]),
                            stroke_color=scenes.T([0, 0, 0, 1])),
              pd.ShapeGroup(shape_ids=torch.tensor([21]), fill_color=None, stroke_color=scenes.T([0, 1, 0, 1]))]
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))


def long_paths(size=96):
    """Paths longer than one wave: a closed 200-cubic loop (distinct sort keys over
    several lanes per segment), a 150-segment staircase whose risers share y-centres
    (ties), a 600-segment zigzag (past the rebuild's per-wave segment limit: one
    workgroup, distinct keys), a 700-segment staircase (one workgroup, ties), a
    2,500-segment polyline whose segments all share one y-centre (one workgroup, every
    element a stopper of every partition) and a 4,200-segment one like it (past the
    workgroup's limit: one thread)."""
    rng = np.random.default_rng(7)
    t = np.linspace(0, 2 * np.pi, 601)[:-1]
    rad = 30 + 6 * np.sin(5 * t) + rng.uniform(-1, 1, t.shape)
    loop = np.stack([48 + rad * np.cos(t), 48 + rad * np.sin(t)], 1)
    p_loop = pd.Path(num_control_points=torch.full((200,), 2, dtype=torch.int32), points=scenes.T(loop),
                     is_closed=True, stroke_width=scenes.T(1.0))
    stair = [[4.0, 90.0]]
    for k in range(75):
        x, y = stair[-1]
        stair += [[x + 1.1, y], [x + 1.1, y - 1.0 * ((k % 5) == 0)]]
    p_stair = pd.Path(num_control_points=torch.zeros(150, dtype=torch.int32), points=scenes.T(np.array(stair)),
                      is_closed=False, stroke_width=scenes.T(0.8))
    xs = np.linspace(2, 94, 601)
    zig = np.stack([xs, 10 + 4 * (np.arange(601) % 2) + 0.01 * np.arange(601)], 1)
    p_zig = pd.Path(num_control_points=torch.zeros(600, dtype=torch.int32), points=scenes.T(zig), is_closed=False,
                    stroke_width=scenes.T(0.7))
    stair2 = [[2.0, 70.0]]
    for k in range(350):
        x, y = stair2[-1]
        stair2 += [[x + 0.25, y], [x + 0.25, y - 0.5 * ((k % 7) == 0)]]
    p_stair2 = pd.Path(num_control_points=torch.zeros(700, dtype=torch.int32), points=scenes.T(np.array(stair2)),
                       is_closed=False, stroke_width=scenes.T(0.6))
    xs = np.linspace(3, 93, 4201)
    flat = np.stack([xs, 50 + 3.0 * (np.arange(4201) % 2)], 1)
    p_flat = pd.Path(num_control_points=torch.zeros(4200, dtype=torch.int32), points=scenes.T(flat),
                     is_closed=False, stroke_width=scenes.T(0.5))
    xs = np.linspace(3, 93, 2501)
    flat2 = np.stack([xs, 20 + 3.0 * (np.arange(2501) % 2)], 1)
    p_flat2 = pd.Path(num_control_points=torch.zeros(2500, dtype=torch.int32), points=scenes.T(flat2),
                      is_closed=False, stroke_width=scenes.T(0.5))
    shapes = [p_loop, p_stair, p_zig, p_stair2, p_flat, p_flat2]
    groups = [pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=scenes.T([0.2, 0.4, 0.9, 0.7]),
                            stroke_color=scenes.T([0, 0, 0, 1])),
              pd.ShapeGroup(shape_ids=torch.tensor([1]), fill_color=None, stroke_color=scenes.T([1, 0, 0, 1])),
              pd.ShapeGroup(shape_ids=torch.tensor([2]), fill_color=None, stroke_color=scenes.T([0, 0.6, 0, 1])),
              pd.ShapeGroup(shape_ids=torch.tensor([3]), fill_color=None, stroke_color=scenes.T([0.5, 0, 0.5, 1])),
              pd.ShapeGroup(shape_ids=torch.tensor([4]), fill_color=None, stroke_color=scenes.T([0, 0, 0.8, 0.6])),
              pd.ShapeGroup(shape_ids=torch.tensor([5]), fill_color=None, stroke_color=scenes.T([0.8, 0.6, 0, 0.6]))]
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))


def big_groups(size=64):
    """Groups past the rebuild's per-lane member limit: 700 small circles in one group, a
    quarter of them stacked four to a Morton cell (ties; one workgroup), and 4,500 in
    another (past the workgroup's limit: one thread)."""
    rng = np.random.default_rng(11)
    shapes = []
    for k in range(700):
        c = [4 + (k % 25) * 2.2, 4 + (k // 25) * 2.0] if k % 4 else [20.0 + 0.0001 * k, 30.0]
        shapes.append(pd.Circle(radius=scenes.T(0.4 + 0.001 * k), center=scenes.T(c)))
    for k in range(4500):
        shapes.append(pd.Circle(radius=scenes.T(0.2), center=scenes.T(rng.uniform(2, size - 2, 2))))
    groups = [pd.ShapeGroup(shape_ids=torch.arange(0, 700), fill_color=scenes.T([0.2, 0.4, 0.9, 0.7])),
              pd.ShapeGroup(shape_ids=torch.arange(700, 5200), fill_color=scenes.T([0.9, 0.3, 0.1, 0.5]))]
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))


EXTRA = {'ties': ties, 'long_paths': long_paths, 'big_groups': big_groups}
CASES = sorted(scenes.CASES) + sorted(EXTRA)


def _flat(case):
    return scenes.flat_case(EXTRA[case]() if case in EXTRA else case)


def _perturb(f, seed, rigid=False):
    """The same topology with other float values (`rigid`: points moved by one common
    offset and shape parameters kept, which keeps ties between sort keys)."""
    rng = np.random.default_rng(seed)
    g = dict(f)
    if rigid:
        pts = f['points']  # a scene without paths holds one placeholder value
        g['points'] = pts if pts.size % 2 else (pts.reshape(-1, 2) + np.array([0.5, 0.25], np.float32)).reshape(
            pts.shape)
    else:
        g['points'] = (f['points'] + rng.normal(0, 0.7, f['points'].shape)).astype(np.float32)
    g['stroke_width'] = (f['stroke_width'] * rng.uniform(0.7, 1.3, f['stroke_width'].shape)).astype(np.float32)
    if f['thickness'] is not None and f['thickness'].size:
        g['thickness'] = (f['thickness'] * rng.uniform(0.7, 1.3, f['thickness'].shape)).astype(np.float32)
    scale = rng.uniform(0.9, 1.1, f['shape_params'].shape)
    g['shape_params'] = f['shape_params'].copy() if rigid else (f['shape_params'] * scale).astype(np.float32)
    g['color_params'] = np.clip(f['color_params'] + rng.normal(0, 0.05, f['color_params'].shape), 0, 1).astype(
        np.float32)
    x = f['shape_to_canvas'].reshape(-1, 3, 3).copy()
    x[:, :2, 2] += rng.normal(0, 0.5, (x.shape[0], 2))
    g['shape_to_canvas'] = x.reshape(-1).astype(np.float32)
    g['filter_radius'] = float(f['filter_radius'])
    return g


def _dev(a):
    if a is None or a.size == 0:
        return None
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _render(scene, w, h, d_img=None):
    img = None if d_img is not None else torch.zeros(h, w, 4, device='cuda')
    di = None if d_img is None else torch.from_numpy(d_img).cuda()
    scene.zero_grads(None)
    scene.render(w, h, 2, 2, 3, render_image=img, d_render_image=di)
    torch.cuda.synchronize()
    return img.cpu().numpy() if img is not None else scene.grads()


@pytest.mark.parametrize('case', CASES)
def test_device_rebuild_equals_host_build(case):
    print('rebuild', case, flush=True)
    f1 = _flat(case)
    f2 = _perturb(f1, 1, rigid=case in ('ties', 'long_paths', 'big_groups'))
    a = _lib.DeviceScene(f1, 0)
    tensors = {k: _dev(f2[k]) for k in FLOAT_KEYS}
    a.update(f2['filter_radius'], None, **tensors)
    a.status()
    b = _lib.DeviceScene(f2, 0)
    ba, bb = a.debug_block(), b.debug_block()
    assert ba.shape == bb.shape
    diff = np.flatnonzero(ba != bb)
    assert diff.size == 0, ('first differing byte', int(diff[0]))
    w, h = f1['canvas_width'], f1['canvas_height']
    assert np.array_equal(_render(a, w, h), _render(b, w, h))
    d = np.random.default_rng(2).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    ga, gb = _render(a, w, h, d), _render(b, w, h, d)
    for k in gb:
        if gb[k].size == 0:
            continue
        scale = max(float(np.abs(gb[k]).max()), 1e-12)
        # same tables: only the order of the float atomics differs (a lone circle's radius
        # adjoint sums 10^5 contributions of both signs), held to the parity bar
        assert np.abs(ga[k] - gb[k]).max() <= 1e-3 * scale, k


def test_rebuild_twice_and_keep_arrays():
    """A second rebuild from other values, and a rebuild that passes only the points."""
    f1 = _flat('overlap_eo')
    f2, f3 = _perturb(f1, 2), _perturb(f1, 3)
    a = _lib.DeviceScene(f1, 0)
    a.update(f2['filter_radius'], None, **{k: _dev(f2[k]) for k in FLOAT_KEYS})
    a.update(f3['filter_radius'], None, **{k: _dev(f3[k]) for k in FLOAT_KEYS})
    a.status()
    assert np.array_equal(a.debug_block(), _lib.DeviceScene(f3, 0).debug_block())
    f4 = dict(f3)
    f4['points'] = _perturb(f1, 4)['points']
    a.update(f4['filter_radius'], None, points=_dev(f4['points']))
    a.status()
    assert np.array_equal(a.debug_block(), _lib.DeviceScene(f4, 0).debug_block())


def test_rebuild_reports_what_construction_rejects():
    f1 = _flat('curve_fill')
    a = _lib.DeviceScene(f1, 0)
    bad = f1['points'].copy()
    bad[3] = np.nan
    a.update(f1['filter_radius'], None, points=_dev(bad))
    with pytest.raises(RuntimeError, match='non-finite path points'):
        a.status()
    a.update(f1['filter_radius'], None, points=_dev(f1['points']))
    a.status()
    # a triangle collapsed to a point: every boundary length is exactly 0
    tri = dict(canvas=(64, 64), shapes=[pd.Polygon(points=scenes.T([[10, 10], [30, 12], [20, 30]]), is_closed=True)],
               groups=[pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=scenes.T([1, 0, 0, 1]))],
               filter=('box', 0.5))
    ft = scenes.flat_case(tri)
    t = _lib.DeviceScene(ft, 0)
    collapsed = np.tile(ft['points'].reshape(-1, 2)[:1], (3, 1)).reshape(ft['points'].shape).astype(np.float32)
    t.update(ft['filter_radius'], None, points=_dev(collapsed))
    with pytest.raises(RuntimeError, match='total length'):
        t.status()
    fc = dict(ft)
    fc['points'] = collapsed
    with pytest.raises(RuntimeError, match='total length'):
        _lib.DeviceScene(fc, 0)


@pytest.mark.parametrize('case', ['overlap_eo', 'strokes', 'transformed', 'ties'])
def test_plain_schedule_renders_updated_scenes(case):
    """The PLAIN schedule walks the reference's top-level BVH (scene.cpp:648-683), which
    the device rebuild does not maintain: its next PLAIN render rebuilds the tree from the
    updated group boxes.  Renders of an updated scene then equal those of a scene created
    from the same values, and the oracle's, in the PLAIN schedule -- also after a second
    update and a TILED render in between."""
    from refdrive import OracleRunner
    print('plain after update', case, flush=True)
    f1 = _flat(case)
    w, h = f1['canvas_width'], f1['canvas_height']
    d = np.random.default_rng(4).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    a = _lib.DeviceScene(f1, 0)
    a.set_mode(_lib.MODE_PLAIN)
    _render(a, w, h)
    for seed in (5, 6):
        f2 = _perturb(f1, seed, rigid=case == 'ties')
        a.update(f2['filter_radius'], None, **{k: _dev(f2[k]) for k in FLOAT_KEYS})
        a.status()
        if seed == 6:
            a.set_mode(_lib.MODE_TILED)
            _render(a, w, h)
            a.set_mode(_lib.MODE_PLAIN)
        b = _lib.DeviceScene(f2, 0, _lib.MODE_PLAIN)
        img_a, img_b = _render(a, w, h), _render(b, w, h)
        want = OracleRunner(f2).render(w, h, 2, 2, 3)['image']
        # PLAIN splats with float atomics in arrival order (5e-5, as for every PLAIN render)
        assert np.abs(img_a - img_b).max() <= 5e-5
        assert np.abs(img_a - want).max() <= 5e-5
        ga, gb = _render(a, w, h, d), _render(b, w, h, d)
        for k in gb:
            if gb[k].size == 0:
                continue
            scale = max(float(np.abs(gb[k]).max()), 1e-12)
            assert np.abs(ga[k] - gb[k]).max() <= 1e-3 * scale, k


def _leaf_scene():
    sc = scenes.overlap(size=64, n=10)
    for s in sc['shapes']:
        s.points.requires_grad_(True)
        s.stroke_width.requires_grad_(True)
    for g in sc['groups']:
        g.fill_color.requires_grad_(True)
    return sc


def _step(sc, target):
    args = pd.RenderFunction.serialize_scene(64, 64, sc['shapes'], sc['groups'])
    img = pd.RenderFunction.apply(64, 64, 2, 2, 5, None, *args)
    loss = ((img - target) ** 2).sum()
    for s in sc['shapes']:
        s.points.grad = None
    loss.backward()
    return img.detach().cpu().numpy(), [s.points.grad.cpu().numpy().copy() for s in sc['shapes']]


def test_render_function_pools_scenes_and_matches_fresh():
    pd.set_use_gpu(True)
    pd.set_device(torch.device('cuda:0'))
    rp.clear_scene_cache()
    sc = _leaf_scene()
    target = torch.rand(64, 64, 4, device='cuda', generator=torch.Generator('cuda').manual_seed(0))
    _step(sc, target)  # host build; starts a pool for this topology
    assert len(rp._templates()) == 1
    with torch.no_grad():
        for s in sc['shapes']:
            s.points += 0.8
    img_pooled, g_pooled = _step(sc, target)  # device rebuild of the pooled scene
    pd.synchronize()  # the last backward's scene returns to the pool once its findings are read
    assert len(rp._templates()) == 1 and len(rp._templates()[0].idle) >= 1
    rp.clear_scene_cache()
    img_fresh, g_fresh = _step(sc, target)  # host build from the same values
    assert np.array_equal(img_pooled, img_fresh)
    for a, b in zip(g_pooled, g_fresh):
        assert np.abs(a - b).max() <= 1e-5 * max(float(np.abs(b).max()), 1e-12)


def test_render_function_rebuild_rejects_non_finite_points():
    pd.set_use_gpu(True)
    pd.set_device(torch.device('cuda:0'))
    rp.clear_scene_cache()
    sc = _leaf_scene()
    target = torch.zeros(64, 64, 4, device='cuda')
    _step(sc, target)
    with torch.no_grad():
        sc['shapes'][2].points[1, 0] = float('nan')
    with pytest.raises(AssertionError):
        _step(sc, target)
    rp.clear_scene_cache()


def test_render_function_more_than_256_spp_on_a_pooled_scene():
    """17 x 16 samples per pixel exceed the tiled kernels' 256 lanes, so such renders run
    the PLAIN schedule -- also on the pooled scene the second call of one layout
    rebuilds on the device.  Both calls (forward and every parameter gradient) agree with
    the oracle."""
    from refdrive import OracleRunner
    pd.set_use_gpu(True)
    pd.set_device(torch.device('cuda:0'))
    rp.clear_scene_cache()
    size = 24
    sc = scenes.overlap(size=size, n=4)
    for s in sc['shapes']:
        s.points.requires_grad_(True)
    rng = np.random.default_rng(9)
    for step in range(2):
        if step:
            with torch.no_grad():
                for s in sc['shapes']:
                    s.points += torch.from_numpy(rng.normal(0, 0.4, tuple(s.points.shape)).astype(np.float32))
        for s in sc['shapes']:
            s.points.grad = None
        args = pd.RenderFunction.serialize_scene(size, size, sc['shapes'], sc['groups'])
        img = pd.RenderFunction.apply(size, size, 17, 16, 1, None, *args)
        d = torch.from_numpy(rng.uniform(-1, 1, (size, size, 4)).astype(np.float32)).cuda()
        (img * d).sum().backward()
        pd.synchronize()
        flat = flatten(size, size, sc['shapes'], sc['groups'], 0, 0.5)
        orc = OracleRunner(flat)
        want = orc.render(size, size, 17, 16, 1)['image']
        assert np.abs(img.detach().cpu().numpy() - want).max() <= 5e-5, step
        gw = orc.render(size, size, 17, 16, 1, d_image=d.cpu().numpy())['grads']['points']
        got = np.concatenate([s.points.grad.cpu().numpy().reshape(-1) for s in sc['shapes']])
        assert np.abs(got - gw).max() <= 1e-3 * max(float(np.abs(gw).max()), 1e-12), step
    rp.clear_scene_cache()


def test_backward_findings_are_reported_at_the_next_sync():
    """The backward returns while its kernels run; what the reference's backward would
    stop on (here a non-finite grad_img, render_pytorch.py:678) is raised at
    synchronize() or at the next entry into the render path, and then cleared."""
    pd.set_use_gpu(True)
    pd.set_device(torch.device('cuda:0'))
    rp.synchronize()
    sc = _leaf_scene()
    args = pd.RenderFunction.serialize_scene(64, 64, sc['shapes'], sc['groups'])
    img = pd.RenderFunction.apply(64, 64, 2, 2, 5, None, *args)
    img.backward(torch.full_like(img, float('nan')))
    with pytest.raises(AssertionError):
        pd.synchronize()
    pd.synchronize()
    target = torch.zeros(64, 64, 4, device='cuda')
    img2, g2 = _step(sc, target)
    pd.synchronize()
    assert np.isfinite(img2).all() and all(np.isfinite(g).all() for g in g2)


def test_prefiltered_circle_stroke_is_reported_at_the_next_sync():
    """The reference's prefiltered forward asserts on a stroked circle (diffvg.cpp:868).
    Here the forward returns while its kernels run; the condition is raised at
    synchronize() (or at the next render call) -- also after the backward has rendered
    on the same scene -- and is then cleared."""
    pd.set_use_gpu(True)
    pd.set_device(torch.device('cuda:0'))
    rp.synchronize()
    c = pd.Circle(radius=scenes.T(10.0), center=scenes.T([32.0, 32.0]))
    c.radius.requires_grad_(True)
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=scenes.T([0.2, 0.4, 0.6, 1.0]),
                      stroke_color=scenes.T([0.0, 0.0, 0.0, 1.0]))
    args = pd.RenderFunction.serialize_scene(64, 64, [c], [g], use_prefiltering=True)
    img = pd.RenderFunction.apply(64, 64, 2, 2, 0, None, *args)
    img.sum().backward()
    with pytest.raises(RuntimeError):
        pd.synchronize()
    pd.synchronize()
