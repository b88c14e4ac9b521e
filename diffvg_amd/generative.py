"""Rasterisation front end of the generative-model examples, batched.

The functions of apps/generative_models/rendering.py with their signatures and results:
`render`, `opacityStroke2diffvg`, `stroke2diffvg`, `line_render` and `bezier_render`.
The reference builds one scene per batch item and calls RenderFunction once per item
(rendering.py:36-98, 185-234, 256-305), moving everything to the CPU first
(`force_cpu`).  Here every batch is one BatchRenderFunction call -- one kernel per stage
over all items, on the GPU -- and each item's scene is rebuilt on the device from the
previous step's pooled scene when its topology repeats.  Random draws happen in the
reference's order (torch.randn_like for the point jitter, then random.randint for the
seed, per item), so with equal RNG states the images are the reference's images.
`force_cpu` is accepted and ignored: rendering stays on the GPU.
"""
import ctypes
import random

import numpy as np
import torch

from . import _lib
from .batch import Batch
from .batch_render import _BatchCheck, render_batch
from .render_pytorch import _PENDING, RenderFunction, _check_gpu, _stream, check_rebuild, flatten_args, synchronize
from .shape import Path, ShapeGroup

__all__ = ['render', 'opacityStroke2diffvg', 'stroke2diffvg', 'line_render', 'bezier_render']


def render(canvas_width, canvas_height, shapes, shape_groups, samples=2, seed=None):
    """One scene through RenderFunction (rendering.py:14-25)."""
    if seed is None:
        seed = random.randint(0, 1000000)
    scene_args = RenderFunction.serialize_scene(canvas_width, canvas_height, shapes, shape_groups)
    return RenderFunction.apply(canvas_width, canvas_height, samples, samples, seed, None, *scene_args)


def _render_items(items, canvas_w, canvas_h, samples, device):
    """items: per batch item a (shapes, groups) pair or None (nothing to draw); one
    seed is drawn per drawn item, in item order, as the per-item loop draws them."""
    drawn, seeds = [], []
    for it in items:
        if it is not None:
            drawn.append((canvas_w, canvas_h, it[0], it[1]))
            seeds.append(random.randint(0, 1000000))
    imgs = render_batch(drawn, samples, samples, seeds) if drawn else []
    out, k = [], 0
    for it in items:
        if it is None:
            out.append(torch.zeros(canvas_h, canvas_w, 4, device=device))
        else:
            out.append(imgs[k])
            k += 1
    return out


def _seed_order(builders):
    """Runs per-item scene builders and seed draws interleaved, item by item, as the
    reference's loop does (jitter of item k, seed of item k, jitter of item k + 1, ...)."""
    items, seeds = [], []
    for b in builders:
        it = b()
        items.append(it)
        seeds.append(random.randint(0, 1000000) if it is not None else None)
    return items, seeds


def _render_seeded(items, seeds, canvas_w, canvas_h, samples, device):
    drawn = [(canvas_w, canvas_h, it[0], it[1]) for it in items if it is not None]
    ds = [s for it, s in zip(items, seeds) if it is not None]
    imgs = render_batch(drawn, samples, samples, ds) if drawn else []
    out, k = [], 0
    for it in items:
        if it is None:
            out.append(torch.zeros(canvas_h, canvas_w, 4, device=device))
        else:
            out.append(imgs[k])
            k += 1
    return out


def opacityStroke2diffvg(strokes, canvas_size=128, debug=False, relative=True, force_cpu=True):
    """Strokes (dx, dy, opacity) [bs, nsegs, 3] -> images [bs, 4, H, W]
    (rendering.py:28-98): one white line segment per step, 4x4 samples."""
    dev = strokes.device
    bs, nsegs, dims = strokes.shape

    def builder(stroke):
        def build():
            all_points = stroke[..., :2].cumsum(0) if relative else stroke[..., :2]
            all_opacities = stroke[..., 2]
            all_points = 0.5 * (all_points + 1.0) * canvas_size
            all_points = all_points + 1e-4 * torch.randn_like(all_points)
            shapes, groups = [], []
            for start in range(0, nsegs - 1):
                points = all_points[start:start + 2].contiguous().float()
                opacity = all_opacities[start]
                shapes.append(Path(num_control_points=torch.zeros(points.shape[0] - 1, dtype=torch.int32),
                                   points=points, stroke_width=torch.ones(1, device=points.device), is_closed=False))
                color = torch.cat([torch.ones(3, device=opacity.device), opacity.unsqueeze(0)], 0)
                groups.append(ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]), fill_color=None,
                                         stroke_color=color))
            return (shapes, groups) if shapes else None
        return build

    items, seeds = _seed_order([builder(s) for s in strokes])
    out = _render_seeded(items, seeds, canvas_size, canvas_size, 4, dev)
    return torch.stack(out, 0).permute(0, 3, 1, 2).contiguous().to(dev)


def stroke2diffvg(strokes, canvas_size=128):
    """Pen-plotter sequences [bs, n, 5] (dx, dy, ., pen-up, end-of-stroke) -> RGB
    images [bs, 3, H, W] (rendering.py:101-166): one white polyline per pen-down run
    of more than two points, 2x2 samples."""
    bs, nsegs, dims = strokes.shape

    def builder(stroke):
        def build():
            end_of_stroke = stroke[:, 4] == 1
            last = end_of_stroke.cpu().numpy().argmax()
            s = stroke[:last + 1, :]
            split_idx = s[:, 3].nonzero().squeeze(1)
            all_points = s[..., :2].cumsum(0)
            all_points[..., 0] += 0.5
            all_points[..., 0] *= canvas_size
            all_points[..., 1] += 0.5
            all_points[..., 1] *= canvas_size
            all_points[..., :2] = torch.clamp(all_points[..., :2], 0, canvas_size)
            shapes, groups = [], []
            start_idx = 0
            for end_idx in split_idx:
                points = all_points[start_idx:end_idx + 1].contiguous().float()
                if points.shape[0] <= 2:
                    continue
                shapes.append(Path(num_control_points=torch.zeros(points.shape[0] - 1, dtype=torch.int32),
                                   points=points, stroke_width=torch.ones(1, device=points.device), is_closed=False))
                start_idx = end_idx + 1
                groups.append(ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]), fill_color=None,
                                         stroke_color=torch.ones(4, 1, device=points.device)))
            return (shapes, groups) if shapes else None
        return build

    items, seeds = _seed_order([builder(s) for s in strokes])
    out = _render_seeded(items, seeds, canvas_size, canvas_size, 2, strokes.device)
    return torch.stack(out, 0).permute(0, 3, 1, 2)[:, :3].contiguous()


class _StrokePool:
    """Device scenes of one stroke-batch topology (n open paths of the same segment kinds,
    one constant stroke colour each, no fill, box filter of radius 0.5), created from the
    first batch's values and rebuilt on the device from every later batch's values.  A
    forward pass leases one set of scenes (one per batch item) until its backward has run
    or its graph is freed, so two forward passes in flight never share a scene."""

    MAX_IDLE = 4

    def __init__(self, desc):
        self.desc = desc
        self.color_idx = torch.from_numpy(
            np.asarray(desc['stroke_offset'], np.int64)[:, None] + np.arange(4, dtype=np.int64))
        self.idle = []
        self._dev_idx = {}

    def color_index(self, dev):
        t = self._dev_idx.get(dev)
        if t is None:
            t = self._dev_idx[dev] = self.color_idx.to(dev)
        return t

    def acquire(self, bs, dev_index):
        if self.idle:
            return self.idle.pop()
        return [_lib.DeviceScene(self.desc, dev_index) for _ in range(bs)]

    def release(self, scenes):
        if len(self.idle) < self.MAX_IDLE:
            self.idle.append(scenes)


class _SetLease:
    def __init__(self, pool, scenes):
        self.pool, self.scenes = pool, scenes

    def __del__(self):
        try:
            self.pool.release(self.scenes)
        except Exception:
            pass


_POOLS = {}


class _StrokeBatch(torch.autograd.Function):
    """A batch of stroke scenes of one topology rendered from batched parameter arrays:
    points [bs, 2 * num_points], stroke widths [bs, n], stroke colours [bs, n, 4].  Each
    item's pooled scene is rebuilt on the device from its rows (dvg_scene_update), all
    items render in one forward batch and one backward batch, and the backward returns
    the gradients of the batched arrays as slices of the scenes' gradient blocks: the
    autograd graph holds one node for the whole batch."""

    @staticmethod
    def forward(ctx, meta, points, widths, colors):
        pool, seeds, size, samples = meta
        dev = points.device
        st = _stream()
        bs = widths.shape[0]
        lease = _SetLease(pool, pool.acquire(bs, dev.index if dev.index is not None else -1))
        col = torch.zeros(bs, pool.desc['num_color_params'], device=dev)
        col[:, pool.color_index(dev)] = colors.detach()
        pts = points.detach().contiguous()
        sw = widths.detach().contiguous()
        for k, sc in enumerate(lease.scenes):
            sc.update(pool.desc['filter_radius'], st, points=pts[k], stroke_width=sw[k], color_params=col[k])
        for sc in lease.scenes:
            # what Scene construction rejects (non-finite points) is reported now
            check_rebuild(sc)
        images = torch.zeros(bs, size, size, 4, device=dev)
        fwd = Batch([(sc, dict(width=size, height=size, nsx=samples, nsy=samples, seed=seeds[k],
                               render_image=images[k])) for k, sc in enumerate(lease.scenes)], zero_grads=False)
        fwd.render(st)
        ctx.meta, ctx.lease, ctx.fwd = meta, lease, fwd
        return images

    @staticmethod
    def backward(ctx, grad_images):
        pool, seeds, size, samples = ctx.meta
        scenes = ctx.lease.scenes
        ctx.fwd = None
        dev = grad_images.device
        st = _stream()
        g = grad_images.to(torch.float32).contiguous()
        bwd = Batch([(sc, dict(width=size, height=size, nsx=samples, nsy=samples, seed=seeds[k],
                               d_render_image=g[k])) for k, sc in enumerate(scenes)], zero_grads=True)
        bwd.render(st)
        layout, total = scenes[0].grad_layout()
        block = torch.empty(len(scenes), total, device=dev)
        for k, sc in enumerate(scenes):
            _lib.check(sc._lib.dvg_scene_copy_grads(sc.handle, ctypes.c_void_p(block[k].data_ptr()), st))
        o = layout['points']
        d_points = block[:, o:o + 2 * pool.desc['num_points_total']]
        o = layout['stroke_width']
        d_widths = block[:, o:o + pool.desc['num_shapes']]
        o = layout['color_params']
        d_colors = block[:, o:o + pool.desc['num_color_params']][:, pool.color_index(dev)]
        # the batch's findings (and a non-finite grad_img) are read at the next render call
        # or synchronize(); the lease stays with them until then
        _PENDING.append((_BatchCheck(bwd, ctx.lease), g.isfinite().all()))
        ctx.lease = None
        return None, d_points, d_widths, d_colors


def _segments_render(all_points, all_widths, all_alphas, canvas_size, colors, ncp_of):
    dev = all_points.device
    all_points = 0.5 * (all_points + 1.0) * canvas_size
    all_points = all_points + 1e-4 * torch.randn_like(all_points)
    bs, n_strokes = all_points.shape[0], all_points.shape[1]
    n_out = 3 if colors is not None else 1
    scenes, seeds = [], []
    for k in range(bs):
        shapes, groups = [], []
        for p in range(n_strokes):
            points = all_points[k, p].contiguous()
            width = all_widths[k, p]
            alpha = all_alphas[k, p]
            color = colors[k, p] if colors is not None else torch.ones(3, device=alpha.device)
            color = torch.cat([color, alpha.view(1,)])
            shapes.append(Path(num_control_points=ncp_of(points), points=points, stroke_width=width,
                               is_closed=False))
            groups.append(ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]), fill_color=None,
                                     stroke_color=color))
        scenes.append((canvas_size, canvas_size, shapes, groups))
        seeds.append(random.randint(0, 1000000))
    if dev.type == 'cuda':
        rasters = _render_strokes_batched(all_points, all_widths, all_alphas, colors, scenes, seeds, canvas_size)
    else:
        rasters = _render_seeded([(sh, gr) for _, _, sh, gr in scenes], seeds, canvas_size, canvas_size, 2, dev)
        rasters = torch.stack(rasters, 0)
    raster = rasters.permute(0, 3, 1, 2)
    alpha = raster[:, 3:4]
    if colors is not None:
        output = raster[:, :3] * alpha.repeat(1, 3, 1, 1)
    else:
        output = raster[:, :1] * alpha
    return output.contiguous(), scenes


def _render_strokes_batched(all_points, all_widths, all_alphas, colors, scenes, seeds, size):
    """The batch's images [bs, size, size, 4] through _StrokeBatch: the scenes share one
    topology, so their parameters go to the device as three batched arrays."""
    _check_gpu(read_findings=False)
    synchronize()
    bs, n = all_widths.shape
    npts = all_points.shape[2]
    w, h, shapes, groups = scenes[0]
    key = (bs, n, npts, size, all_points.device.index, tuple(shapes[0].num_control_points.tolist()))
    pool = _POOLS.get(key)
    if pool is None:
        flat = flatten_args(RenderFunction.serialize_scene(w, h, shapes, groups))
        pool = _POOLS[key] = _StrokePool(flat.desc)
    pts = all_points.reshape(bs, n * npts * 2).to(torch.float32)
    rgb = colors if colors is not None else torch.ones(bs, n, 3, device=all_points.device)
    col = torch.cat([rgb, all_alphas.unsqueeze(-1)], -1).to(torch.float32)
    return _StrokeBatch.apply((pool, list(seeds), size, 2), pts, all_widths.to(torch.float32), col)


def line_render(all_points, all_widths, all_alphas, force_cpu=True, canvas_size=32, colors=None):
    """Line segments [bs, n, 2, 2] in [-1, 1] with widths and alphas [bs, n] (and colours
    [bs, n, 3]) -> (images [bs, 1 or 3, H, W] premultiplied by alpha, scenes)
    (rendering.py:169-234)."""
    return _segments_render(all_points, all_widths, all_alphas, canvas_size, colors,
                            lambda pts: torch.zeros(1, dtype=torch.int32))


def bezier_render(all_points, all_widths, all_alphas, force_cpu=True, canvas_size=32, colors=None):
    """Cubic strokes [bs, n, 3 m + 1, 2] in [-1, 1] -> (images, scenes)
    (rendering.py:237-305)."""
    num_segments = (all_points.shape[2] - 1) // 3
    return _segments_render(all_points, all_widths, all_alphas, canvas_size, colors,
                            lambda pts: torch.zeros(num_segments, dtype=torch.int32) + 2)
