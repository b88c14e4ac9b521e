"""Scene zoo for parity tests (test infrastructure).

Every case returns (canvas_w, canvas_h, shapes, shape_groups, filter_type, filter_radius,
render options).  Scenes are built with the pydiffvg object model of diffvg_amd and are
deterministic (numpy Generator seeded per case).
"""
import math

import numpy as np
import torch

import diffvg_amd as pd
from diffvg_amd import workloads


def T(x):
    return torch.tensor(np.asarray(x, dtype=np.float32))


def _xform(theta, sx, sy, tx, ty):
    c, s = math.cos(theta), math.sin(theta)
    return T([[sx * c, -sy * s, tx], [sx * s, sy * c, ty], [0, 0, 1]])


def single_circle(size=256):
    # apps/single_circle.py:9-27
    s = size / 256.0
    c = pd.Circle(radius=T(40.0 * s), center=T([128.0 * s, 128.0 * s]))
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=T([0.3, 0.6, 0.3, 1.0]))
    return dict(canvas=(size, size), shapes=[c], groups=[g], filter=('box', 0.5))


def curve_fill(size=64):
    # apps/single_curve_sdf.py:9-27 scaled to the canvas
    s = size / 256.0
    pts = T(np.array([[120.0, 30.0], [150.0, 60.0], [90.0, 198.0], [60.0, 218.0], [90.0, 180.0],
                      [200.0, 65.0], [210.0, 98.0], [220.0, 70.0], [130.0, 55.0]]) * s)
    p = pd.Path(num_control_points=torch.tensor([2, 2, 2]), points=pts, is_closed=True)
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=T([0.3, 0.6, 0.3, 1.0]))
    return dict(canvas=(size, size), shapes=[p], groups=[g], filter=('box', 0.5))


def open_stroke(size=64, width=3.0):
    s = size / 64.0
    pts = T(np.array([[8, 50], [20, 5], [40, 60], [56, 14], [60, 30], [40, 40], [30, 20]]) * s)
    p = pd.Path(num_control_points=torch.tensor([2, 2]), points=pts, is_closed=False,
                stroke_width=T(width))
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=None, stroke_color=T([0.8, 0.2, 0.1, 0.9]))
    return dict(canvas=(size, size), shapes=[p], groups=[g], filter=('box', 0.5))


def thick_stroke(size=64):
    pts = T([[8, 40], [20, 8], [36, 60], [56, 20], [40, 30]])
    p = pd.Path(num_control_points=torch.tensor([2, 0]), points=pts, is_closed=False,
                stroke_width=T([1.0, 2.0, 3.0, 4.0, 2.5]))
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=None, stroke_color=T([0.1, 0.3, 0.9, 1.0]))
    return dict(canvas=(size, size), shapes=[p], groups=[g], filter=('box', 0.5))


def mixed_segments(size=64):
    # line + quadratic + cubic closed path, stroked and filled
    pts = T([[10, 10], [50, 8], [58, 30], [40, 50], [30, 62], [12, 40]])
    p = pd.Path(num_control_points=torch.tensor([0, 1, 2]), points=pts, is_closed=True, stroke_width=T(1.5))
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=T([0.2, 0.7, 0.4, 0.8]),
                      stroke_color=T([0.9, 0.1, 0.1, 1.0]))
    return dict(canvas=(size, size), shapes=[p], groups=[g], filter=('box', 0.5))


def quad_approx(size=64):
    pts = T([[6, 50], [30, -10], [58, 50], [40, 60], [20, 20]])
    p = pd.Path(num_control_points=torch.tensor([1, 1]), points=pts, is_closed=False, stroke_width=T(2.0),
                use_distance_approx=True)
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=None, stroke_color=T([0.5, 0.5, 0.1, 1.0]))
    return dict(canvas=(size, size), shapes=[p], groups=[g], filter=('box', 0.5))


def primitives(size=64):
    shapes = [pd.Rect(p_min=T([6.0, 8.0]), p_max=T([30.0, 40.0]), stroke_width=T(1.5)),
              pd.Ellipse(radius=T([14.0, 8.0]), center=T([42.0, 42.0])),
              pd.Circle(radius=T(10.0), center=T([44.0, 18.0]), stroke_width=T(2.0)),
              pd.Polygon(points=T([[8, 56], [24, 44], [30, 60]]), is_closed=True)]
    groups = [pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=T([0.9, 0.5, 0.1, 0.7]),
                            stroke_color=T([0.1, 0.1, 0.8, 1.0])),
              pd.ShapeGroup(shape_ids=torch.tensor([1]), fill_color=T([0.2, 0.6, 0.9, 0.6])),
              pd.ShapeGroup(shape_ids=torch.tensor([2]), fill_color=T([0.5, 0.9, 0.3, 0.5]),
                            stroke_color=T([0.0, 0.0, 0.0, 1.0])),
              pd.ShapeGroup(shape_ids=torch.tensor([3]), fill_color=T([0.7, 0.2, 0.7, 1.0]), use_even_odd_rule=False)]
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))


def gradients(size=64):
    pts = T([[8, 8], [56, 8], [56, 56], [8, 56]])
    p0 = pd.Polygon(points=pts, is_closed=True)
    p1 = pd.Circle(radius=T(18.0), center=T([32.0, 32.0]))
    lin = pd.LinearGradient(begin=T([8.0, 8.0]), end=T([56.0, 40.0]), offsets=T([0.0, 0.4, 1.0]),
                            stop_colors=T([[1, 0, 0, 1], [0, 1, 0, 0.8], [0, 0, 1, 1]]))
    rad = pd.RadialGradient(center=T([30.0, 34.0]), radius=T([16.0, 12.0]), offsets=T([0.1, 0.6, 0.9]),
                            stop_colors=T([[1, 1, 0, 1], [0, 1, 1, 0.5], [1, 0, 1, 0.9]]))
    groups = [pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=lin),
              pd.ShapeGroup(shape_ids=torch.tensor([1]), fill_color=rad)]
    return dict(canvas=(size, size), shapes=[p0, p1], groups=groups, filter=('box', 0.5))


def transformed(size=64):
    pts = T([[0, 0], [20, -10], [30, 20], [10, 25], [-5, 18], [-10, 5]])
    p = pd.Path(num_control_points=torch.tensor([2, 1]), points=pts, is_closed=True, stroke_width=T(1.0))
    c = pd.Circle(radius=T(6.0), center=T([8.0, 6.0]))
    groups = [pd.ShapeGroup(shape_ids=torch.tensor([0, 1]), fill_color=T([0.3, 0.4, 0.9, 0.9]),
                            stroke_color=T([0.9, 0.9, 0.1, 1.0]), shape_to_canvas=_xform(0.4, 1.3, 0.9, 24, 18))]
    return dict(canvas=(size, size), shapes=[p, c], groups=groups, filter=('box', 0.5))


def overlap(size=64, seed=7, n=12, even_odd=True):
    rng = np.random.default_rng(seed)
    shapes, groups = [], []
    for i in range(n):
        nseg = int(rng.integers(3, 6))
        c = rng.uniform(10, size - 10, 2)
        pts = c + rng.normal(0, 9, (3 * nseg, 2))
        shapes.append(pd.Path(num_control_points=torch.full((nseg,), 2, dtype=torch.int32), points=T(pts),
                              is_closed=True, stroke_width=T(rng.uniform(0.5, 2.0))))
        fill = T(np.r_[rng.uniform(0, 1, 3), rng.uniform(0.3, 1.0)])
        stroke = T(rng.uniform(0, 1, 4)) if i % 3 == 0 else None
        groups.append(pd.ShapeGroup(shape_ids=torch.tensor([i]), fill_color=fill, stroke_color=stroke,
                                    use_even_odd_rule=even_odd))
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))


def strokes(size=64, seed=11, n=24):
    # painterly-style open strokes (config 3 in miniature)
    rng = np.random.default_rng(seed)
    shapes, groups = [], []
    for i in range(n):
        p0 = rng.uniform(0, size, 2)
        steps = rng.normal(0, size * 20.0 / 512.0 * 2, (9, 2))
        pts = np.vstack([p0, p0 + np.cumsum(steps, axis=0)])
        shapes.append(pd.Path(num_control_points=torch.full((3,), 2, dtype=torch.int32), points=T(pts),
                              is_closed=False, stroke_width=T(rng.uniform(1.0, 4.0) * 0.5)))
        groups.append(pd.ShapeGroup(shape_ids=torch.tensor([i]), fill_color=None,
                                    stroke_color=T(rng.uniform(0, 1, 4))))
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))


CASES = {
    'single_circle_256': lambda: single_circle(256),
    'single_circle_64': lambda: single_circle(64),
    'curve_fill': curve_fill,
    'open_stroke': open_stroke,
    'thick_stroke': thick_stroke,
    'mixed_segments': mixed_segments,
    'quad_approx': quad_approx,
    'primitives': primitives,
    'gradients': gradients,
    'transformed': transformed,
    'overlap_eo': lambda: overlap(even_odd=True),
    'overlap_nz': lambda: overlap(seed=8, even_odd=False),
    'strokes': strokes,
}


def flat_case(case, filt=None):
    from refdrive import flatten
    sc = CASES[case]() if isinstance(case, str) else case
    ftype, frad = filt if filt is not None else sc['filter']
    w, h = sc['canvas']
    return flatten(w, h, sc['shapes'], sc['groups'], ftype, frad)



def _from_workload(c):
    return dict(canvas=c['canvas'], shapes=c['shapes'], groups=c['groups'], filter=('box', 0.5))


def many_groups():
    """10^4 small groups (diffvg_amd.workloads.many_groups)."""
    return _from_workload(workloads.many_groups())


def crowded_tile():
    """More groups over one tile than one candidate pass holds (workloads.crowded_tile)."""
    return _from_workload(workloads.crowded_tile())


def long_paths():
    """Paths of 1,500 and 600 cubic segments (workloads.long_paths)."""
    return _from_workload(workloads.long_paths())
