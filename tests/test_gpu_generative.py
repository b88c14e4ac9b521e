"""Batched differentiable rendering (BatchRenderFunction, render_batch) and the batched
generative-model front end (diffvg_amd.generative) against one RenderFunction call per
scene, the way apps/generative_models/rendering.py renders a batch.

With equal seeds the batched images equal the per-scene images bitwise (the same
sample kernels run; only the launch is shared) and gradients agree within float
summation order.  The per-item loops below restate rendering.py's loops (169-305) with
RenderFunction, drawing random numbers in the same order."""
import random

import numpy as np
import pytest
import torch

import diffvg_amd as pd
from diffvg_amd import generative
from diffvg_amd.batch_render import render_batch
import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _gpu():
    pd.set_use_gpu(True)
    pd.set_device(torch.device('cuda:0'))


def _rel(a, b):
    return float((a - b).abs().max()) / max(float(b.abs().max()), 1e-12)


def test_render_batch_equals_per_scene_calls():
    print('render_batch', flush=True)
    items = []
    for k, case in enumerate(['overlap_eo', 'strokes', 'gradients', 'mixed_segments', 'transformed']):
        sc = scenes.CASES[case]()
        for s in sc['shapes']:
            for attr in ('points', 'center', 'radius', 'p_min', 'p_max'):
                t = getattr(s, attr, None)
                if isinstance(t, torch.Tensor) and t.dtype == torch.float32:
                    t.requires_grad_(True)
        items.append(sc)
    seeds = [3, 9, 27, 81, 243]
    targets = [torch.rand(sc['canvas'][1], sc['canvas'][0], 4, device='cuda') for sc in items]

    def leaves():
        out = []
        for sc in items:
            for s in sc['shapes']:
                for attr in ('points', 'center', 'radius', 'p_min', 'p_max'):
                    t = getattr(s, attr, None)
                    if isinstance(t, torch.Tensor) and t.requires_grad:
                        out.append(t)
        return out

    imgs = render_batch([(sc['canvas'][0], sc['canvas'][1], sc['shapes'], sc['groups']) for sc in items], 2, 2, seeds)
    loss = sum(((im - t) ** 2).sum() for im, t in zip(imgs, targets))
    loss.backward()
    g_batch = [t.grad.clone() for t in leaves()]
    for t in leaves():
        t.grad = None
    singles = []
    for sc, seed in zip(items, seeds):
        w, h = sc['canvas']
        args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'])
        singles.append(pd.RenderFunction.apply(w, h, 2, 2, seed, None, *args))
    loss = sum(((im - t) ** 2).sum() for im, t in zip(singles, targets))
    loss.backward()
    for a, b in zip(imgs, singles):
        assert torch.equal(a.detach(), b.detach())
    # gradients: the same contributions summed by float atomics in another order (a
    # circle's radius adjoint sums ~10^4 terms of both signs), held to the parity bar
    for a, t in zip(g_batch, leaves()):
        assert _rel(a, t.grad) <= 1e-3


def _loop_segments(all_points, all_widths, all_alphas, canvas_size, colors, ncp):
    """rendering.py:169-305 with one RenderFunction call per item."""
    all_points = 0.5 * (all_points + 1.0) * canvas_size
    all_points = all_points + 1e-4 * torch.randn_like(all_points)
    bs, n = all_points.shape[0], all_points.shape[1]
    n_out = 3 if colors is not None else 1
    output = torch.zeros(bs, n_out, canvas_size, canvas_size, device=all_points.device)
    for k in range(bs):
        shapes, groups = [], []
        for p in range(n):
            color = colors[k, p] if colors is not None else torch.ones(3, device=all_points.device)
            color = torch.cat([color, all_alphas[k, p].view(1,)])
            shapes.append(pd.Path(num_control_points=ncp(), points=all_points[k, p].contiguous(),
                                  stroke_width=all_widths[k, p], is_closed=False))
            groups.append(pd.ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]), fill_color=None,
                                        stroke_color=color))
        raster = generative.render(canvas_size, canvas_size, shapes, groups, samples=2)
        raster = raster.permute(2, 0, 1).view(4, canvas_size, canvas_size)
        alpha = raster[3:4]
        if colors is not None:
            image, alpha = raster[:3], alpha.repeat(3, 1, 1)
        else:
            image = raster[:1]
        output[k] = image * alpha
    return output


@pytest.mark.parametrize('kind', ['line', 'bezier', 'bezier_color'])
def test_generative_front_end_matches_per_item_loop(kind):
    print('generative', kind, flush=True)
    bs, n = 6, 5
    g = torch.Generator(device='cuda').manual_seed(1)
    npts = 2 if kind == 'line' else 7
    pts0 = torch.rand(bs, n, npts, 2, device='cuda', generator=g) * 1.6 - 0.8
    widths0 = torch.rand(bs, n, device='cuda', generator=g) * 1.5 + 0.5
    alphas0 = torch.rand(bs, n, device='cuda', generator=g) * 0.8 + 0.2
    colors0 = torch.rand(bs, n, 3, device='cuda', generator=g) if kind == 'bezier_color' else None
    nseg = (npts - 1) // 3

    def run(batched):
        pts = pts0.clone().requires_grad_(True)
        widths = widths0.clone().requires_grad_(True)
        alphas = alphas0.clone().requires_grad_(True)
        colors = None if colors0 is None else colors0.clone().requires_grad_(True)
        torch.manual_seed(5)
        random.seed(7)
        if batched:
            fn = generative.line_render if kind == 'line' else generative.bezier_render
            out, _ = fn(pts, widths, alphas, canvas_size=32, colors=colors)
        else:
            ncp = (lambda: torch.zeros(1, dtype=torch.int32)) if kind == 'line' else \
                (lambda: torch.zeros(nseg, dtype=torch.int32) + 2)
            out = _loop_segments(pts, widths, alphas, 32, colors, ncp)
        (out * torch.linspace(0, 1, out.numel(), device='cuda').view(out.shape)).sum().backward()
        grads = [pts.grad, widths.grad, alphas.grad] + ([] if colors is None else [colors.grad])
        return out.detach(), grads

    ob, gb = run(True)
    ol, gl = run(False)
    assert torch.equal(ob, ol)
    for a, b in zip(gb, gl):
        assert _rel(a, b) <= 1e-3


def test_opacity_strokes_match_per_item_loop():
    print('opacity strokes', flush=True)
    g = torch.Generator(device='cuda').manual_seed(2)
    strokes = torch.rand(4, 9, 3, device='cuda', generator=g) * 0.3 - 0.15
    strokes[..., 2] = torch.rand(4, 9, device='cuda', generator=g)
    torch.manual_seed(11)
    random.seed(3)
    batched = generative.opacityStroke2diffvg(strokes, canvas_size=48)
    torch.manual_seed(11)
    random.seed(3)
    out = []
    for stroke in strokes:
        all_points = 0.5 * (stroke[..., :2].cumsum(0) + 1.0) * 48
        all_points = all_points + 1e-4 * torch.randn_like(all_points)
        shapes, groups = [], []
        for i in range(stroke.shape[0] - 1):
            shapes.append(pd.Path(num_control_points=torch.zeros(1, dtype=torch.int32),
                                  points=all_points[i:i + 2].contiguous().float(), stroke_width=torch.ones(1),
                                  is_closed=False))
            groups.append(pd.ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]), fill_color=None,
                                        stroke_color=torch.cat([torch.ones(3, device='cuda'),
                                                                stroke[i, 2].unsqueeze(0)])))
        out.append(generative.render(48, 48, shapes, groups, samples=4))
    loop = torch.stack(out, 0).permute(0, 3, 1, 2).contiguous()
    assert torch.equal(batched, loop)


def test_render_batch_scenes_return_to_the_pool():
    """A batched step releases its pooled scenes when its backward's findings are read,
    so the next step of the same topologies rebuilds them instead of building new ones."""
    from diffvg_amd import render_pytorch as rp
    rp.synchronize()
    rp.clear_scene_cache()
    items = [scenes.CASES[c]() for c in ('overlap_eo', 'strokes', 'curve_fill')]
    for sc in items:
        for s in sc['shapes']:
            s.points.requires_grad_(True)
    lst = [(sc['canvas'][0], sc['canvas'][1], sc['shapes'], sc['groups']) for sc in items]
    for step in range(3):
        imgs = render_batch(lst, 2, 2, [1, 2, 3])
        sum(im.sum() for im in imgs).backward()
        del imgs
        pd.synchronize()
        tpls = rp._templates()
        assert len(tpls) == 3
        assert all(len(t.idle) == 1 for t in tpls), [len(t.idle) for t in tpls]


def test_render_batch_reports_non_finite_grad_at_sync():
    sc = scenes.CASES['overlap_eo']()
    for s in sc['shapes']:
        s.points.requires_grad_(True)
    (img,) = render_batch([(sc['canvas'][0], sc['canvas'][1], sc['shapes'], sc['groups'])], 2, 2, [0])
    img.backward(torch.full_like(img, float('nan')))
    with pytest.raises(AssertionError):
        pd.synchronize()
    pd.synchronize()


@pytest.mark.parametrize('kind', ['line', 'bezier_color'])
def test_generative_front_end_matches_oracle(kind):
    """line_render / bezier_render against the oracle: each item's scene (as the front end
    returns it) rendered by the oracle at the seed the front end drew for it, put through
    the same alpha product gives the batched output (within 2e-5: products of values within
    1e-5); the oracle's backward for that item's image gradient, chained through
    points = 0.5 (p + 1) canvas_size + jitter and colour = (rgb, alpha), gives the
    gradients of the generative inputs (within 1e-3 relative)."""
    from refdrive import OracleRunner, flatten
    print('generative vs oracle', kind, flush=True)
    bs, n, cs = 4, 5, 32
    g = torch.Generator(device='cuda').manual_seed(3)
    npts = 2 if kind == 'line' else 7
    pts = (torch.rand(bs, n, npts, 2, device='cuda', generator=g) * 1.6 - 0.8).requires_grad_(True)
    widths = (torch.rand(bs, n, device='cuda', generator=g) * 1.5 + 0.5).requires_grad_(True)
    alphas = (torch.rand(bs, n, device='cuda', generator=g) * 0.8 + 0.2).requires_grad_(True)
    colors = (torch.rand(bs, n, 3, device='cuda', generator=g).requires_grad_(True)
              if kind == 'bezier_color' else None)
    torch.manual_seed(5)
    random.seed(7)
    fn = generative.line_render if kind == 'line' else generative.bezier_render
    out, items = fn(pts, widths, alphas, canvas_size=cs, colors=colors)
    weight = torch.linspace(0, 1, out.numel(), device='cuda').view(out.shape)
    (out * weight).sum().backward()
    random.seed(7)
    seeds = [random.randint(0, 1000000) for _ in range(bs)]  # one per item, as the front end draws them
    want_pts = np.zeros((bs, n, npts, 2), np.float32)
    want_w, want_a = np.zeros((bs, n), np.float32), np.zeros((bs, n), np.float32)
    want_c = np.zeros((bs, n, 3), np.float32)
    for k in range(bs):
        w, h, shapes, groups = items[k]
        f = flatten(w, h, shapes, groups, 0, 0.5)
        o = OracleRunner(f)
        raster = torch.from_numpy(o.render(w, h, 2, 2, seeds[k])['image']).requires_grad_(True)
        r = raster.permute(2, 0, 1).reshape(4, cs, cs)
        alpha = r[3:4]
        img = r[:3] * alpha.repeat(3, 1, 1) if colors is not None else r[:1] * alpha
        assert float((img.detach() - out[k].detach().cpu()).abs().max()) <= 2e-5, k
        (img * weight[k].cpu()).sum().backward()
        gr = o.render(w, h, 2, 2, seeds[k], d_image=raster.grad.numpy())['grads']
        want_pts[k] = gr['points'].reshape(n, npts, 2) * (0.5 * cs)
        want_w[k] = gr['stroke_width']
        col = np.stack([gr['color_params'][f['stroke_offset'][i]:f['stroke_offset'][i] + 4] for i in range(n)])
        want_a[k] = col[:, 3]
        want_c[k] = col[:, :3]
    got = [pts.grad, widths.grad, alphas.grad] + ([] if colors is None else [colors.grad])
    want = [want_pts, want_w, want_a] + ([] if colors is None else [want_c])
    for name, a, b in zip(('points', 'widths', 'alphas', 'colors'), got, want):
        assert _rel(a.cpu(), torch.from_numpy(b)) <= 1e-3, name


def test_generative_forwards_in_flight_keep_their_scenes():
    """Two bezier_render passes of one topology before either backward: each holds its own
    set of pooled scenes, so each backward returns its own pass's gradients (equal to
    those of the pass run alone); after both, the sets go back to the pool."""
    print('generative in flight', flush=True)
    bs, n = 3, 4
    g = torch.Generator(device='cuda').manual_seed(9)
    inputs = [(torch.rand(bs, n, 7, 2, device='cuda', generator=g) * 1.6 - 0.8,
               torch.rand(bs, n, device='cuda', generator=g) + 0.5,
               torch.rand(bs, n, device='cuda', generator=g) * 0.8 + 0.2) for _ in range(2)]

    def forward(i):
        pts, w, a = (t.clone().requires_grad_(True) for t in inputs[i])
        torch.manual_seed(i)
        random.seed(i)
        out, _ = generative.bezier_render(pts, w, a, canvas_size=24)
        return out, (pts, w, a)

    alone = []
    for i in range(2):
        out, leaves = forward(i)
        out.sum().backward()
        alone.append((out.detach(), [t.grad.clone() for t in leaves]))
    outs = [forward(i) for i in range(2)]
    for i in (1, 0):
        out, leaves = outs[i]
        out.sum().backward()
        assert torch.equal(out.detach(), alone[i][0])
        for a, b in zip((t.grad for t in leaves), alone[i][1]):
            assert _rel(a, b) <= 1e-5
    del outs, out
    pd.synchronize()
    pools = [p for k, p in generative._POOLS.items() if k[:4] == (bs, n, 7, 24)]
    assert len(pools) == 1 and len(pools[0].idle) == 2


def test_generative_non_finite_points_raise():
    pts = torch.zeros(2, 3, 2, 2, device='cuda')
    pts[1, 2, 0, 0] = float('nan')
    with pytest.raises(AssertionError):
        generative.line_render(pts, torch.ones(2, 3, device='cuda'), torch.ones(2, 3, device='cuda'))
    pd.synchronize()
