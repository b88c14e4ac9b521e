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
"""
import math

import numpy as np
import torch

from . import diffvg
from .color import LinearGradient, RadialGradient
from .pixel_filter import PixelFilter
from .shape import Circle, Path, ShapeGroup


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


WORKLOADS = {1: single_circle, 2: closed_paths_300, 3: strokes_2048}
