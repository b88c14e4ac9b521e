"""Direct delivery of scene gradients into the argument tensors' .grad
(set_direct_leaf_gradients) against the autograd engine's delivery of the same gradient
block, on the CPU: both routes take one gradient block laid out by kind and must give
every argument tensor the same .grad, bit for bit (both only route values).  Delivery
runs through the C++ helper (_dvg_deliver.so) when it is built and through the Python
cut of the views otherwise; both are checked."""
import numpy as np
import pytest
import torch

import diffvg_amd as pd
from diffvg_amd import render_pytorch as rp
import scenes

FLOAT_ATTRS = ('points', 'thickness', 'radius', 'center', 'p_min', 'p_max', 'stroke_width')
PAINT_ATTRS = ('fill_color', 'stroke_color')


def _leaves(sc):
    out = []
    for s in sc['shapes']:
        for a in FLOAT_ATTRS:
            t = getattr(s, a, None)
            if isinstance(t, torch.Tensor) and t.is_floating_point():
                out.append(t)
    for g in sc['groups']:
        for a in PAINT_ATTRS:
            c = getattr(g, a, None)
            if isinstance(c, torch.Tensor):
                out.append(c)
            elif c is not None:
                out += [t for t in (getattr(c, 'begin', None), getattr(c, 'end', None), getattr(c, 'center', None),
                                    getattr(c, 'radius', None),
                                    c.offsets, c.stop_colors) if isinstance(t, torch.Tensor)]
        out.append(g.shape_to_canvas)
    uniq, seen = [], set()
    for t in out:
        if id(t) not in seen:
            seen.add(id(t))
            uniq.append(t)
    return uniq


def _setup(case, skip_every=0):
    sc = scenes.CASES[case]()
    leaves = _leaves(sc)
    for i, t in enumerate(leaves):
        t.requires_grad_(not (skip_every and i % skip_every == 0))
    w, h = sc['canvas']
    args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'])
    flat = rp.flatten_args(args)
    tpl = rp._SceneTemplate(args, flat)
    d = flat.desc
    sizes = dict(shape_params=d['num_shape_params'], points=2 * d['num_points_total'],
                 thickness=d['num_points_total'], stroke_width=d['num_shapes'],
                 color_params=d['num_color_params'], shape_to_canvas=9 * d['num_groups'])
    layout, o = {}, 0
    for k in rp._KINDS:
        layout[k] = o
        o += sizes[k] + 3  # gaps between the kinds, as a block may have
    layout['filter_radius'] = o
    block = torch.from_numpy(np.random.default_rng(1).standard_normal(o + 1).astype(np.float32))
    return sc, args, tpl, leaves, layout, block, sizes


def _engine_grads(args, tpl, leaves, layout, block, sizes):
    for t in leaves:
        t.grad = None
    vals = tpl.device_params(args, torch.device('cpu'), grad=True)
    loss = 0
    for k in rp._KINDS:
        v = vals[k]
        if v is not None and v.requires_grad:
            loss = loss + (v * block[layout[k]:layout[k] + sizes[k]]).sum()
    loss.backward()
    return [None if t.grad is None else t.grad.clone() for t in leaves]


@pytest.fixture(params=['native', 'python'])
def route(request, monkeypatch):
    if request.param == 'native':
        if rp._native() is None:
            pytest.skip('_dvg_deliver.so not built (diffvg_amd.build.build_deliver)')
    else:
        monkeypatch.setattr(rp, '_NATIVE', [None])
    return request.param


@pytest.mark.parametrize('case', ['primitives', 'gradients', 'transformed', 'thick_stroke', 'mixed_segments',
                                  'strokes', 'overlap_eo'])
@pytest.mark.parametrize('skip_every', [0, 3])
def test_direct_delivery_equals_engine(case, skip_every, route):
    sc, args, tpl, leaves, layout, block, sizes = _setup(case, skip_every)
    want = _engine_grads(args, tpl, leaves, layout, block, sizes)
    for t in leaves:
        t.grad = None
    plan = tpl.direct_leaves(args, torch.device('cpu'))
    assert plan is not None
    # each backward delivers from a fresh block (a tensor used several times sums its
    # slots into the first one)
    tpl.deliver(block.clone(), layout, plan)
    uses = {}
    for a in args:
        uses[id(a)] = uses.get(id(a), 0) + 1

    def same(a, b, t):
        # a tensor used several times (the shared default transform) sums its slots, in
        # an order the engine does not promise
        return torch.equal(a, b) if uses[id(t)] == 1 else torch.allclose(a, b, rtol=1e-6, atol=1e-5)

    for t, w in zip(leaves, want):
        if w is None:
            assert t.grad is None
        else:
            assert t.grad is not None and same(t.grad, w, t), t.shape
    # a second backward adds, as the engine's accumulation does
    tpl.deliver(block.clone(), layout, plan)
    for t, w in zip(leaves, want):
        if w is not None:
            assert same(t.grad, 2 * w, t)


def test_direct_delivery_refuses_what_the_engine_must_route():
    sc, args, tpl, leaves, layout, block, sizes = _setup('strokes')
    cpu = torch.device('cpu')
    assert tpl.direct_leaves(args, cpu) is not None
    # an argument computed from another tensor (not a leaf)
    p = sc['shapes'][0].points
    args2 = list(args)
    i = next(k for k, a in enumerate(args) if a is p)
    args2[i] = p * 1.0
    assert tpl.direct_leaves(args2, cpu) is None
    # a tensor with a gradient hook
    h = p.register_hook(lambda g: g)
    assert tpl.direct_leaves(args, cpu) is None
    h.remove()
    # float64 arguments
    args4 = list(args)
    args4[i] = p.detach().double().requires_grad_(True)
    assert tpl.direct_leaves(args4, cpu) is None
    # nothing requires a gradient
    for t in leaves:
        t.requires_grad_(False)
    assert tpl.direct_leaves(args, cpu) is None


def test_direct_delivery_not_for_multi_element_stroke_widths():
    sc = scenes.CASES['strokes']()
    sc['shapes'][0].stroke_width = torch.tensor([1.5, 2.0])
    w, h = sc['canvas']
    args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'])
    tpl = rp._SceneTemplate(args, rp.flatten_args(args))
    assert tpl.direct_leaves(args, torch.device('cpu')) is None


def test_direct_delivery_rechecks_after_an_accepted_call():
    """The second call with the same tensors reuses the first call's checks except for
    hooks; other tensors or changed gradient flags are checked in full."""
    sc, args, tpl, leaves, layout, block, sizes = _setup('strokes')
    cpu = torch.device('cpu')
    assert tpl.direct_leaves(args, cpu) is not None
    assert tpl.direct_leaves(args, cpu) is not None
    p = sc['shapes'][0].points
    h = p.register_hook(lambda g: g)
    assert tpl.direct_leaves(args, cpu) is None
    h.remove()
    i = next(k for k, a in enumerate(args) if a is p)
    args2 = list(args)
    args2[i] = p * 1.0
    assert tpl.direct_leaves(args2, cpu) is None
    q = sc['shapes'][1].points
    q.requires_grad_(False)
    plan = tpl.direct_leaves(args, cpu)
    assert plan is not None
    tpl.deliver(block.clone(), layout, plan)
    assert q.grad is None and p.grad is not None
