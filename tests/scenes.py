"""Scene zoo for parity tests (test infrastructure).

Every case returns (canvas_w, canvas_h, shapes, shape_groups, filter_type, filter_radius,
render options).  Scenes are built with the pydiffvg object model of diffvg_amd and are
deterministic (numpy Generator seeded per case).
"""
import math

import numpy as np
import torch

import diffvg_amd as pd


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


def many_groups(size=256, n=12000, seed=5):
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
        shapes.append(pd.Path(num_control_points=torch.ones(3, dtype=torch.int32), points=T(pts), is_closed=True,
                              stroke_width=T(rng.uniform(0.2, 0.8))))
        groups.append(pd.ShapeGroup(shape_ids=torch.tensor([i]), fill_color=T(np.r_[rng.uniform(0, 1, 3), 0.7]),
                                    stroke_color=T(rng.uniform(0, 1, 4)) if stroked else None))
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))


def crowded_tile(size=16, grid=56, seed=6):
    """More groups overlapping one tile than one candidate pass holds (grid^2 = 3,136 tiny
    triangles in a 16x16 canvas) under two larger translucent paths."""
    rng = np.random.default_rng(seed)
    shapes, groups = [], []
    step = size / grid
    for iy in range(grid):
        for ix in range(grid):
            c = np.array([(ix + 0.5) * step, (iy + 0.5) * step]) + rng.uniform(-0.1, 0.1, 2) * step
            r = 0.45 * step
            pts = [c + [r, 0], c + [-0.5 * r, 0.87 * r], c + [-0.5 * r, -0.87 * r]]
            shapes.append(pd.Path(num_control_points=torch.zeros(3, dtype=torch.int32), points=T(pts),
                                  is_closed=True, stroke_width=T(0.1)))
            groups.append(pd.ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]),
                                        fill_color=T(np.r_[rng.uniform(0, 1, 3), 0.6])))
    for k in range(2):
        c = np.array([size * (0.35 + 0.3 * k), size * 0.5])
        pts = c + rng.normal(0, size * 0.25, (9, 2))
        shapes.append(pd.Path(num_control_points=torch.full((3,), 2, dtype=torch.int32), points=T(pts),
                              is_closed=True, stroke_width=T(0.6)))
        groups.append(pd.ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]),
                                    fill_color=T(np.r_[rng.uniform(0, 1, 3), 0.5]),
                                    stroke_color=T([0.1, 0.2, 0.9, 0.8])))
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))


def long_paths(size=128, seed=7, n_closed=1500, n_open=600):
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
    shapes = [pd.Rect(p_min=T([size * 0.1, size * 0.2]), p_max=T([size * 0.8, size * 0.7])),
              pd.Path(num_control_points=torch.full((n_closed,), 2, dtype=torch.int32), points=T(ring),
                      is_closed=True, stroke_width=T(1.2)),
              pd.Path(num_control_points=torch.full((n_open,), 2, dtype=torch.int32), points=T(spiral),
                      is_closed=False, stroke_width=T(1.5))]
    groups = [pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=T([0.2, 0.5, 0.9, 0.6])),
              pd.ShapeGroup(shape_ids=torch.tensor([1]), fill_color=T([0.9, 0.4, 0.1, 0.7]),
                            stroke_color=T([0.1, 0.1, 0.1, 0.9]), use_even_odd_rule=False),
              pd.ShapeGroup(shape_ids=torch.tensor([2]), fill_color=None, stroke_color=T([0.2, 0.8, 0.3, 0.8]))]
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))
