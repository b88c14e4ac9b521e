"""Shape containers with the constructor signatures of pydiffvg/shape.py:5-61, and
``from_svg_path`` (pydiffvg/shape.py:63-172) over diffvg_amd.svg_path.
"""
import torch

from . import svg_path as _svg_path


class Circle:
    def __init__(self, radius, center, stroke_width=torch.tensor(1.0), id=''):
        self.radius = radius
        self.center = center
        self.stroke_width = stroke_width
        self.id = id


class Ellipse:
    def __init__(self, radius, center, stroke_width=torch.tensor(1.0), id=''):
        self.radius = radius
        self.center = center
        self.stroke_width = stroke_width
        self.id = id


class Path:
    def __init__(self,
                 num_control_points,
                 points,
                 is_closed,
                 stroke_width=torch.tensor(1.0),
                 id='',
                 use_distance_approx=False):
        self.num_control_points = num_control_points
        self.points = points
        self.is_closed = is_closed
        self.stroke_width = stroke_width
        self.id = id
        self.use_distance_approx = use_distance_approx


class Polygon:
    def __init__(self, points, is_closed, stroke_width=torch.tensor(1.0), id=''):
        self.points = points
        self.is_closed = is_closed
        self.stroke_width = stroke_width
        self.id = id


class Rect:
    def __init__(self, p_min, p_max, stroke_width=torch.tensor(1.0), id=''):
        self.p_min = p_min
        self.p_max = p_max
        self.stroke_width = stroke_width
        self.id = id


class ShapeGroup:
    def __init__(self,
                 shape_ids,
                 fill_color,
                 use_even_odd_rule=True,
                 stroke_color=None,
                 shape_to_canvas=torch.eye(3),
                 id=''):
        self.shape_ids = shape_ids
        self.fill_color = fill_color
        self.use_even_odd_rule = use_even_odd_rule
        self.stroke_color = stroke_color
        self.shape_to_canvas = shape_to_canvas
        self.id = id


def from_svg_path(path_str, shape_to_canvas=torch.eye(3), force_close=False):
    """SVG path data -> list of Path, one per continuous subpath, with points mapped
    through `shape_to_canvas` (pydiffvg/shape.py:63-172).  Near-closed subpaths (ends
    within 1e-5) are snapped shut; with `force_close` open subpaths get a closing line;
    arcs become cubics."""
    segs = _svg_path.parse_path(path_str)
    if len(segs) == 0:
        return []
    return _svg_path.segments_to_paths(segs, shape_to_canvas, force_close, Path)
