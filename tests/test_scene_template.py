"""Scene templates (the topology a pooled device scene is rebuilt for) on the CPU: a later
argument list matches when its topology is the template's; the same tensor objects skip
the shape comparison, and a tensor resized in place since is caught by the gathers."""
import torch

import diffvg_amd as pd
from diffvg_amd import render_pytorch as rp
import scenes


def _args(sc):
    w, h = sc['canvas']
    return pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'])


def test_template_matches_same_topology_and_rejects_other_shapes():
    sc = scenes.CASES['strokes']()
    a = _args(sc)
    tpl = rp._SceneTemplate(a, rp.flatten_args(a))
    assert tpl.matches(_args(sc)) and tpl.matches(_args(sc))
    # new tensors with the same shapes: still the topology
    for s in sc['shapes']:
        s.points = s.points.clone() + 1.0
    assert tpl.matches(_args(sc))
    # a path with one more point: another topology
    p = sc['shapes'][3]
    p.points = torch.cat([p.points, p.points[-1:]])
    p.num_control_points = torch.cat([p.num_control_points, torch.zeros(1, dtype=torch.int32)])
    assert not tpl.matches(_args(sc))


def test_in_place_resize_is_caught_by_the_gather():
    sc = scenes.CASES['strokes']()
    a = _args(sc)
    tpl = rp._SceneTemplate(a, rp.flatten_args(a))
    assert tpl.matches(_args(sc))
    vals = tpl.device_params(_args(sc), torch.device('cpu'))
    assert vals['points'].numel() == 2 * tpl.flat.desc['num_points_total']
    p = sc['shapes'][0].points
    p.data = torch.cat([p.data, p.data[:1]])  # same tensor object, one more point
    b = _args(sc)
    assert tpl.matches(b)  # the identity path does not compare shapes ...
    try:
        tpl.device_params(b, torch.device('cpu'))
        raise AssertionError('resize not caught')
    except rp._Resized:
        pass  # ... the gather's element count does
    tpl._float_seen = None
    assert not tpl.matches(b)


def test_batch_gather_equals_per_scene_gathers():
    """render_batch gathers the float arrays of all pooled scenes with one op per kind and
    cuts them per scene; each scene's arrays must equal its own gather, values and
    gradients (scenes with circles, rects, gradients, per-point thickness, transforms,
    the same topology twice)."""
    names = ['primitives', 'gradients', 'thick_stroke', 'transformed', 'strokes', 'mixed_segments', 'strokes']
    cases = [scenes.CASES[n]() for n in names]
    for sc in cases:
        for s in sc['shapes']:
            for a in ('points', 'thickness', 'radius', 'center', 'p_min', 'p_max', 'stroke_width'):
                t = getattr(s, a, None)
                if isinstance(t, torch.Tensor) and t.is_floating_point():
                    t.requires_grad_(True)
    arg_lists = [_args(sc) for sc in cases]
    tpls = []
    for a in arg_lists:
        tp = next((t for t in tpls if t.matches(a)), None) or rp._SceneTemplate(a, rp.flatten_args(a))
        tpls.append(tp)
    cpu = torch.device('cpu')
    batch = rp._batch_params(tpls, arg_lists, cpu, True)
    weights = {}
    loss_b, loss_s = 0, 0
    for tp, a, vb in zip(tpls, arg_lists, batch):
        vs = tp.device_params(a, cpu, True)
        for k in rp._KINDS:
            assert (vs[k] is None) == (vb[k] is None), k
            if vs[k] is None:
                continue
            assert torch.equal(vs[k], vb[k]), k
            w = weights.setdefault((id(tp), k, len(weights)), torch.randn(vs[k].shape))
            loss_s = loss_s + (vs[k] * w).sum()
            loss_b = loss_b + (vb[k] * w).sum()
    leaves = [t for sc in cases for s in sc['shapes'] for t in (getattr(s, a, None) for a in
              ('points', 'thickness', 'radius', 'center', 'p_min', 'p_max', 'stroke_width'))
              if isinstance(t, torch.Tensor) and t.requires_grad]
    gb = torch.autograd.grad(loss_b, leaves, allow_unused=True)
    gs = torch.autograd.grad(loss_s, leaves, allow_unused=True)
    for a, b in zip(gb, gs):
        assert (a is None) == (b is None)
        if a is not None:
            assert torch.allclose(a, b, rtol=1e-6, atol=1e-6)


def test_template_keeps_no_argument_tensor_alive():
    import gc
    import weakref
    sc = scenes.CASES['strokes']()
    for s in sc['shapes']:
        s.points.requires_grad_(True)
    a = _args(sc)
    tpl = rp._SceneTemplate(a, rp.flatten_args(a))
    b = _args(sc)
    assert tpl.matches(b) and tpl.direct_leaves(b, torch.device('cpu')) is not None
    ref = weakref.ref(sc['shapes'][0].points)
    del sc, a, b
    gc.collect()
    assert ref() is None


def test_clearing_the_scene_cache_drops_every_template_reference():
    import gc
    import weakref
    from diffvg_amd import batch_render
    sc = scenes.CASES['strokes']()
    a = _args(sc)
    tpl = rp._SceneTemplate(a, rp.flatten_args(a))
    rp._TEMPLATES.setdefault(len(a), []).insert(0, tpl)
    rp._batch_tables([tpl, tpl], torch.device('cpu'))
    batch_render._HINTS[:] = [tpl]
    ref = weakref.ref(tpl)
    del tpl
    rp.clear_scene_cache()
    gc.collect()
    assert ref() is None


def test_int64_shape_ids_serialize_to_the_same_int32_tensor_until_modified():
    import gc
    ids = torch.tensor([0, 2])
    ncp = torch.tensor([2, 2, 1])
    sc = dict(canvas=(8, 8),
              shapes=[pd.Path(num_control_points=ncp, points=torch.rand(8, 2) * 8, is_closed=True),
                      pd.Circle(radius=torch.tensor(1.0), center=torch.tensor([4.0, 4.0])),
                      pd.Circle(radius=torch.tensor(2.0), center=torch.tensor([3.0, 3.0]))],
              groups=[pd.ShapeGroup(shape_ids=ids, fill_color=torch.tensor([1.0, 0, 0, 1]))])
    a, b = _args(sc), _args(sc)
    pos_ids = next(i for i, x in enumerate(a) if isinstance(x, torch.Tensor) and x.tolist() == [0, 2])
    assert a[pos_ids] is b[pos_ids] and a[pos_ids].dtype == torch.int32
    pos_ncp = next(i for i, x in enumerate(a) if isinstance(x, torch.Tensor) and x.tolist() == [2, 2, 1])
    assert a[pos_ncp] is b[pos_ncp] and a[pos_ncp].dtype == torch.int32
    # an in-place change is seen
    ids[1] = 1
    c = _args(sc)
    assert c[pos_ids].tolist() == [0, 1] and c[pos_ids] is not a[pos_ids]
    # the conversion of a tensor that died is dropped
    n = len(rp._I32)
    sc['groups'][0].shape_ids = torch.tensor([1, 2])
    _args(sc)
    del ids
    gc.collect()
    assert len(rp._I32) == n
