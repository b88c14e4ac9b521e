"""RenderFunction with the gradients written straight into the argument tensors' .grad
(set_direct_leaf_gradients) against the same renders with the autograd engine delivering
them: images equal bitwise, gradients within float summation order (the backward's
atomics may sum in another order from run to run)."""
import pytest
import torch

import diffvg_amd as pd
import scenes
from test_direct_grads import _leaves

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    pd.set_use_gpu(True)
    pd.set_device(torch.device('cuda:0'))
    yield
    pd.set_direct_leaf_gradients(False)


def _to_cuda(sc):
    for s in sc['shapes']:
        for a in ('points', 'thickness', 'radius', 'center', 'p_min', 'p_max', 'stroke_width'):
            t = getattr(s, a, None)
            if isinstance(t, torch.Tensor) and t.is_floating_point():
                setattr(s, a, t.cuda())
    for g in sc['groups']:
        for a in ('fill_color', 'stroke_color'):
            c = getattr(g, a, None)
            if isinstance(c, torch.Tensor):
                setattr(g, a, c.cuda())
            elif c is not None:
                for k in ('begin', 'end', 'center', 'radius', 'offsets', 'stop_colors'):
                    t = getattr(c, k, None)
                    if isinstance(t, torch.Tensor):
                        setattr(c, k, t.cuda())
        g.shape_to_canvas = g.shape_to_canvas.cuda()
    return sc


def _step(sc, target, direct, steps=2):
    pd.set_direct_leaf_gradients(direct)
    w, h = sc['canvas']
    leaves = _leaves(sc)
    for t in leaves:
        t.requires_grad_(True)
        t.grad = None
    img = None
    for k in range(steps):
        args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'])
        img = pd.RenderFunction.apply(w, h, 2, 2, 5, None, *args)
        kinds = {type(f).__name__ for f, _ in img.grad_fn.next_functions if f is not None}
        ((img - target) ** 2).sum().backward()
    pd.synchronize()
    return img.detach(), [t.grad.clone() for t in leaves], kinds


@pytest.mark.parametrize('case', ['strokes', 'gradients', 'transformed', 'thick_stroke', 'primitives'])
def test_direct_leaf_gradients_match_engine(case):
    print('direct leaf gradients', case, flush=True)
    sc = _to_cuda(scenes.CASES[case]())
    w, h = sc['canvas']
    target = torch.rand(h, w, 4, device='cuda', generator=torch.Generator(device='cuda').manual_seed(4))
    # two steps each: the first render of a topology builds it on the host and goes
    # through the engine; the second rebuilds a pooled scene and may deliver directly.
    # Gradients of the two steps accumulate, as in a loop without zero_grad.
    img_e, g_e, kinds_e = _step(sc, target, False)
    img_d, g_d, kinds_d = _step(sc, target, True)
    assert kinds_d == {'AccumulateGrad'}, kinds_d  # only the anchor: no per-kind gather nodes
    assert kinds_e != kinds_d
    assert torch.equal(img_e, img_d)
    for a, b in zip(g_d, g_e):
        scale = max(float(b.abs().max()), 1e-6)
        assert float((a - b).abs().max()) <= 1e-4 * scale


def test_direct_leaf_gradients_fall_back_for_non_leaves():
    sc = _to_cuda(scenes.CASES['strokes']())
    w, h = sc['canvas']
    pd.set_direct_leaf_gradients(True)
    base = sc['shapes'][0].points.clone().requires_grad_(True)
    for k in range(2):
        sc['shapes'][0].points = base * 1.0  # not a leaf: the engine routes this render
        args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'])
        img = pd.RenderFunction.apply(w, h, 2, 2, 5, None, *args)
        img.sum().backward()
    pd.synchronize()
    assert base.grad is not None and torch.isfinite(base.grad).all()
    assert float(base.grad.abs().max()) > 0


def test_render_batch_direct_leaf_gradients_match_engine():
    """render_batch over scenes of several topologies: direct delivery gives every
    argument tensor the gradient the engine gives it."""
    from diffvg_amd.batch_render import render_batch
    print('render_batch direct leaf gradients', flush=True)
    cases = ['strokes', 'gradients', 'transformed', 'overlap_eo', 'primitives']
    items = [_to_cuda(scenes.CASES[c]()) for c in cases]
    targets = [torch.rand(sc['canvas'][1], sc['canvas'][0], 4, device='cuda',
                          generator=torch.Generator(device='cuda').manual_seed(i)) for i, sc in enumerate(items)]
    leaves = [t for sc in items for t in _leaves(sc)]

    def run(direct):
        pd.set_direct_leaf_gradients(direct)
        for t in leaves:
            t.requires_grad_(True)
            t.grad = None
        lst = [(sc['canvas'][0], sc['canvas'][1], sc['shapes'], sc['groups']) for sc in items]
        for _ in range(2):  # the first call builds the topologies on the host
            imgs = render_batch(lst, 2, 2, list(range(len(items))))
            sum(((im - t) ** 2).sum() for im, t in zip(imgs, targets)).backward()
        pd.synchronize()
        return [im.detach() for im in imgs], [t.grad.clone() for t in leaves]

    ie, ge = run(False)
    idr, gd = run(True)
    for a, b in zip(ie, idr):
        assert torch.equal(a, b)
    for a, b in zip(gd, ge):
        scale = max(float(b.abs().max()), 1e-6)
        assert float((a - b).abs().max()) <= 1e-4 * scale
