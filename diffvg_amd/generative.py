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
import random

import torch

from .batch_render import render_batch
from .render_pytorch import RenderFunction
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


def _segments_render(all_points, all_widths, all_alphas, canvas_size, colors, ncp_of):
    dev = all_points.device
    all_points = 0.5 * (all_points + 1.0) * canvas_size
    all_points = all_points + 1e-4 * torch.randn_like(all_points)
    bs, n_strokes = all_points.shape[0], all_points.shape[1]
    n_out = 3 if colors is not None else 1
    scenes, items, seeds = [], [], []
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
        items.append((shapes, groups))
        seeds.append(random.randint(0, 1000000))
    rasters = _render_seeded(items, seeds, canvas_size, canvas_size, 2, dev)
    output = torch.zeros(bs, n_out, canvas_size, canvas_size, device=dev)
    for k, raster in enumerate(rasters):
        raster = raster.permute(2, 0, 1).view(4, canvas_size, canvas_size)
        alpha = raster[3:4]
        if colors is not None:
            image = raster[:3]
            alpha = alpha.repeat(3, 1, 1)
        else:
            image = raster[:1]
        output[k] = image * alpha
    return output, scenes


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
