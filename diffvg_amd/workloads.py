"""Seeded synthetic scenes for the benchmark configurations.

Each generator returns a dict with canvas size, shapes, shape groups, filter, sample
counts and, where the configuration defines one, a target image.  Generators draw from
numpy.random.default_rng(workload number); parameters are float32 torch tensors on
`device`.

  1 single_circle      circle (128,128) r=40, fill (0.3,0.6,0.3,1), 256x256, 2x2, box
           0.5; target circle (100,120) r=60 (apps/single_circle.py:9-27)
  2 closed_paths_300   512x512, 2x2, Gaussian filter (sigma 0.5, radius 1.5): 300 closed cubic
           paths with 4-20 segments, points = centre + N(0, 15^2); fill alpha in
           [0.3, 1]; 30% also stroked with width in [0.5, 3]; even-odd fill
  3 strokes_2048       512x512, 2x2, box 0.5: 2048 open 3-cubic strokes, random walk N(0, 20^2),
           width in [1, 4], uniform-noise target
  4 uneven_scenes      one scene of the uneven batch: path count log-uniform in [1, 512], 2-8
           cubics per path, resolution in {28, 64, 128, 256}, 10% linear/radial fills
  5 serving_request    one request of the serving stream: path count log-uniform in [1, 4096],
           resolution in {128, 256, 512, 1024}, colour or SDF, prefilter on or off

Scenes of unusual scale (SCALE): many_groups (12,000 small groups, 256x256),
crowded_tile (3,138 groups over a 16x16 canvas), long_paths (closed and open paths of
1,500 and 600 cubic segments, 128x128).
"""
import math

import numpy as np
import torch

from . import diffvg
from .color import LinearGradient, RadialGradient
from .pixel_filter import PixelFilter
from .shape import Circle, Path, Rect, ShapeGroup


def _t(x, device):
    return torch.tensor(np.asarray(x, dtype=np.float32), device=device)


def single_circle(device='cpu'):
    c = Circle(radius=_t(40.0, device), center=_t([128.0, 128.0], device))
    g = ShapeGroup(shape_ids=torch.tensor([0], dtype=torch.int32), fill_color=_t([0.3, 0.6, 0.3, 1.0], device))
    tc = Circle(radius=_t(60.0, device), center=_t([100.0, 120.0], device))
    tg = ShapeGroup(shape_ids=torch.tensor([0], dtype=torch.int32), fill_color=_t([0.3, 0.6, 0.3, 1.0], device))
    return dict(name='single_circle_256', canvas=(256, 256), shapes=[c], groups=[g],
                filter=PixelFilter(type=diffvg.FilterType.box, radius=torch.tensor(0.5)), nsx=2, nsy=2,
                target_scene=([tc], [tg]))


def closed_paths(rng, n_paths, size, seg_lo, seg_hi, spread, device, stroke_frac=0.3, grad_frac=0.0,
                 even_odd=True):
    shapes, groups = [], []
    for i in range(n_paths):
        nseg = int(rng.integers(seg_lo, seg_hi + 1))
        centre = rng.uniform(0, size, 2)
        pts = centre + rng.normal(0, spread, (3 * nseg, 2))
        stroked = rng.uniform() < stroke_frac
        width = rng.uniform(0.5, 3.0) if stroked else 1.0
        shapes.append(Path(num_control_points=torch.full((nseg,), 2, dtype=torch.int32), points=_t(pts, device),
                           is_closed=True, stroke_width=_t(width, device)))
        fill = _t(np.r_[rng.uniform(0, 1, 3), rng.uniform(0.3, 1.0)], device)
        if grad_frac > 0 and rng.uniform() < grad_frac:
            nstops = int(rng.integers(2, 5))
            offs = np.sort(rng.uniform(0, 1, nstops))
            offs[0] = 0.0
            cols = rng.uniform(0, 1, (nstops, 4))
            if rng.uniform() < 0.5:
                fill = LinearGradient(begin=_t(centre - spread, device), end=_t(centre + spread, device),
                                      offsets=_t(offs, device), stop_colors=_t(cols, device))
            else:
                fill = RadialGradient(center=_t(centre, device), radius=_t([2 * spread, 2 * spread], device),
                                      offsets=_t(offs, device), stop_colors=_t(cols, device))
        stroke = _t(rng.uniform(0, 1, 4), device) if stroked else None
        groups.append(ShapeGroup(shape_ids=torch.tensor([i], dtype=torch.int32), fill_color=fill,
                                 stroke_color=stroke, use_even_odd_rule=even_odd))
    return shapes, groups


def closed_paths_300(device='cpu', filter_type='gaussian'):
    rng = np.random.default_rng(2)
    shapes, groups = closed_paths(rng, 300, 512.0, 4, 20, 15.0, device)
    if filter_type == 'gaussian':
        filt = PixelFilter(type=diffvg.FilterType.gaussian, radius=torch.tensor(1.5))
    else:  # parity twin (the reference has no Gaussian filter)
        filt = PixelFilter(type=diffvg.FilterType.tent, radius=torch.tensor(1.0))
    return dict(name='closed_paths_300_512', canvas=(512, 512), shapes=shapes, groups=groups,
                filter=filt, nsx=2, nsy=2)


def strokes(rng, n_paths, size, step, device, wlo=1.0, whi=4.0):
    shapes, groups = [], []
    for i in range(n_paths):
        p0 = rng.uniform(0, size, 2)
        pts = np.vstack([p0, p0 + np.cumsum(rng.normal(0, step, (9, 2)), axis=0)])
        shapes.append(Path(num_control_points=torch.full((3,), 2, dtype=torch.int32), points=_t(pts, device),
                           is_closed=False, stroke_width=_t(rng.uniform(wlo, whi), device)))
        groups.append(ShapeGroup(shape_ids=torch.tensor([i], dtype=torch.int32), fill_color=None,
                                 stroke_color=_t(rng.uniform(0, 1, 4), device)))
    return shapes, groups


def strokes_2048(device='cpu'):
    rng = np.random.default_rng(3)
    shapes, groups = strokes(rng, 2048, 512.0, 20.0, device)
    target = torch.from_numpy(rng.uniform(0, 1, (512, 512, 4)).astype(np.float32)).to(device)
    return dict(name='strokes_2048_512', canvas=(512, 512), shapes=shapes, groups=groups,
                filter=PixelFilter(type=diffvg.FilterType.box, radius=torch.tensor(0.5)), nsx=2, nsy=2,
                target=target)


def uneven_scenes(n_scenes=256, device='cpu', seed=4):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_scenes):
        res = int(rng.choice([28, 64, 128, 256]))
        n = max(1, int(math.floor(math.exp(rng.uniform(0, math.log(512))))))
        shapes, groups = closed_paths(rng, n, float(res), 2, 8, res * 15.0 / 512.0 * 4, device, stroke_frac=0.0,
                                      grad_frac=0.1)
        out.append(dict(name='uneven_scene%d' % k, canvas=(res, res), shapes=shapes, groups=groups,
                        filter=PixelFilter(type=diffvg.FilterType.box, radius=torch.tensor(0.5)), nsx=2, nsy=2))
    return out


def serving_request(rng, device='cpu', res=None, n_paths=None, sdf=None, prefilter=None):
    """One request; `res`, `n_paths`, `sdf` and `prefilter` pin what is otherwise drawn."""
    r_res = int(rng.choice([128, 256, 512, 1024]))
    r_n = max(1, int(math.floor(math.exp(rng.uniform(0, math.log(4096))))))
    res = r_res if res is None else int(res)
    n = r_n if n_paths is None else int(n_paths)
    # the path size shrinks with the path count so that the expected overlap depth stays
    # bounded (the reference renders at most 64 prefiltered / 256 plain fragments per sample)
    spread = max(1.0, 0.6 * res / math.sqrt(n))
    shapes, groups = closed_paths(rng, n, float(res), 2, 8, spread, device, stroke_frac=0.3)
    r_sdf, r_pre = rng.uniform() < 0.5, rng.uniform() < 0.5
    sdf = r_sdf if sdf is None else bool(sdf)
    prefilter = r_pre if prefilter is None else bool(prefilter)
    return dict(name='serving_request', canvas=(res, res), shapes=shapes, groups=groups,
                filter=PixelFilter(type=diffvg.FilterType.box, radius=torch.tensor(0.5)),
                nsx=1 if sdf else 2, nsy=1 if sdf else 2, sdf=sdf, prefilter=prefilter)


def many_groups(device='cpu', size=256, n=12000, seed=5):
    """10^4 small groups: closed quadratic triangles, half of them stroked, spread over
    the canvas (a tile's candidate list holds tens of groups, the scene ten thousand)."""
    rng = np.random.default_rng(seed)
    shapes, groups = [], []
    for i in range(n):
        c = rng.uniform(2, size - 2, 2)
        r = rng.uniform(0.8, 3.0)
        ang = rng.uniform(0, 2 * math.pi) + np.array([0, 2.1, 4.2])
        pts = []
        for k in range(3):
            a0, a1 = ang[k], ang[(k + 1) % 3]
            p0 = c + r * np.array([math.cos(a0), math.sin(a0)])
            pm = c + 0.7 * r * np.array([math.cos(0.5 * (a0 + a1)), math.sin(0.5 * (a0 + a1))])
            pts += [p0, pm]
        stroked = i % 2 == 0
        shapes.append(Path(num_control_points=torch.ones(3, dtype=torch.int32), points=_t(pts, device),
                           is_closed=True, stroke_width=_t(rng.uniform(0.2, 0.8), device)))
        groups.append(ShapeGroup(shape_ids=torch.tensor([i]), fill_color=_t(np.r_[rng.uniform(0, 1, 3), 0.7], device),
                                 stroke_color=_t(rng.uniform(0, 1, 4), device) if stroked else None))
    return dict(name='many_groups_12000_256', canvas=(size, size), shapes=shapes, groups=groups,
                filter=PixelFilter(type=diffvg.FilterType.box, radius=torch.tensor(0.5)), nsx=2, nsy=2)


def crowded_tile(device='cpu', size=16, grid=56, seed=6):
    """More groups overlapping one tile than one candidate pass of the tiled kernels holds
    (grid^2 = 3,136 tiny triangles in a 16x16 canvas) under two larger translucent paths."""
    rng = np.random.default_rng(seed)
    shapes, groups = [], []
    step = size / grid
    for iy in range(grid):
        for ix in range(grid):
            c = np.array([(ix + 0.5) * step, (iy + 0.5) * step]) + rng.uniform(-0.1, 0.1, 2) * step
            r = 0.45 * step
            pts = [c + [r, 0], c + [-0.5 * r, 0.87 * r], c + [-0.5 * r, -0.87 * r]]
            shapes.append(Path(num_control_points=torch.zeros(3, dtype=torch.int32), points=_t(pts, device),
                               is_closed=True, stroke_width=_t(0.1, device)))
            groups.append(ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]),
                                     fill_color=_t(np.r_[rng.uniform(0, 1, 3), 0.6], device)))
    for k in range(2):
        c = np.array([size * (0.35 + 0.3 * k), size * 0.5])
        pts = c + rng.normal(0, size * 0.25, (9, 2))
        shapes.append(Path(num_control_points=torch.full((3,), 2, dtype=torch.int32), points=_t(pts, device),
                           is_closed=True, stroke_width=_t(0.6, device)))
        groups.append(ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]),
                                 fill_color=_t(np.r_[rng.uniform(0, 1, 3), 0.5], device),
                                 stroke_color=_t([0.1, 0.2, 0.9, 0.8], device)))
    return dict(name='crowded_tile_3138_16', canvas=(size, size), shapes=shapes, groups=groups,
                filter=PixelFilter(type=diffvg.FilterType.box, radius=torch.tensor(0.5)), nsx=2, nsy=2)


def long_paths(device='cpu', size=128, seed=7, n_closed=1500, n_open=600):
    """A closed path of 1,500 cubic segments (a wobbling ring) and an open stroke of 600
    cubic segments (a spiral), over a translucent rectangle."""
    rng = np.random.default_rng(seed)
    c = np.array([size / 2, size / 2])
    t = np.linspace(0, 2 * math.pi, 3 * n_closed, endpoint=False)
    rad = size * 0.36 + size * 0.05 * np.sin(37 * t) + rng.normal(0, 0.4, t.shape)
    ring = c + np.stack([rad * np.cos(t), rad * np.sin(t)], 1)
    u = np.linspace(0, 6 * math.pi, 3 * n_open + 1)
    sr = size * 0.04 + size * 0.018 * u
    spiral = c + np.stack([sr * np.cos(u), sr * np.sin(u)], 1)
    shapes = [Rect(p_min=_t([size * 0.1, size * 0.2], device), p_max=_t([size * 0.8, size * 0.7], device)),
              Path(num_control_points=torch.full((n_closed,), 2, dtype=torch.int32), points=_t(ring, device),
                   is_closed=True, stroke_width=_t(1.2, device)),
              Path(num_control_points=torch.full((n_open,), 2, dtype=torch.int32), points=_t(spiral, device),
                   is_closed=False, stroke_width=_t(1.5, device))]
    groups = [ShapeGroup(shape_ids=torch.tensor([0]), fill_color=_t([0.2, 0.5, 0.9, 0.6], device)),
              ShapeGroup(shape_ids=torch.tensor([1]), fill_color=_t([0.9, 0.4, 0.1, 0.7], device),
                         stroke_color=_t([0.1, 0.1, 0.1, 0.9], device), use_even_odd_rule=False),
              ShapeGroup(shape_ids=torch.tensor([2]), fill_color=None, stroke_color=_t([0.2, 0.8, 0.3, 0.8], device))]
    return dict(name='long_paths_1500_600_128', canvas=(size, size), shapes=shapes, groups=groups,
                filter=PixelFilter(type=diffvg.FilterType.box, radius=torch.tensor(0.5)), nsx=2, nsy=2)


# scenes of unusual scale (the GPU scale tests and the bench's scale leg)
SCALE = (many_groups, crowded_tile, long_paths)

WORKLOADS = {1: single_circle, 2: closed_paths_300, 3: strokes_2048}
