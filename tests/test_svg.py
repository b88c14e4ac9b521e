"""SVG loading and saving (diffvg_amd.svg_path, parse_svg, save_svg) on the CPU.

The reference's loader needs svgpathtools and cssutils, which are not installed, so its
results are pinned here by the SVG 1.1 specification and closed forms: hand-derived
segment lists for each path command, arc centres and sweep directions from the circle
through the end points, arc radii scaled up when too small (F.6.6), transform matrices
written out by hand, and the save -> load round trip being exact for float32 points.
"""
import math
import os

import numpy as np
import pytest
import torch

import diffvg_amd as pydiffvg
from diffvg_amd import svg_path
from diffvg_amd.parse_svg import parse_color, parse_transform

pytestmark = pytest.mark.filterwarnings('ignore')


def _pts(path):
    return path.points.numpy().astype(np.float64)


def _ncp(path):
    return path.num_control_points.tolist()


# ---- path data -----------------------------------------------------------------------

def test_lines_h_v_close():
    (p,) = pydiffvg.from_svg_path('M 10 20 L 30 40 H 50 V 60 Z')
    assert p.is_closed
    assert _ncp(p) == [0, 0, 0, 0]
    np.testing.assert_array_equal(_pts(p), [[10, 20], [30, 40], [50, 40], [50, 60]])


def test_relative_equals_absolute():
    a = pydiffvg.from_svg_path('M 10 20 L 30 40 C 35 45 40 45 50 40 Q 55 30 60 40 L 70 70')
    r = pydiffvg.from_svg_path('m 10 20 l 20 20 c 5 5 10 5 20 0 q 5 -10 10 0 l 10 30')
    assert len(a) == len(r) == 1
    assert _ncp(a[0]) == _ncp(r[0]) == [0, 2, 1, 0]
    np.testing.assert_array_equal(_pts(a[0]), _pts(r[0]))
    assert not a[0].is_closed


def test_implicit_lineto_after_moveto():
    (p,) = pydiffvg.from_svg_path('M 0 0 10 0 10 10')
    assert _ncp(p) == [0, 0]
    np.testing.assert_array_equal(_pts(p), [[0, 0], [10, 0], [10, 10]])
    (q,) = pydiffvg.from_svg_path('m 1 1 2 0 0 2')
    np.testing.assert_array_equal(_pts(q), [[1, 1], [3, 1], [3, 3]])


def test_smooth_cubic_and_quadratic_reflect_controls():
    (p,) = pydiffvg.from_svg_path('M 0 0 C 0 10 10 10 10 0 S 20 -10 20 0')
    # S reflects the previous second control (10,10) about (10,0) -> (10,-10)
    np.testing.assert_array_equal(_pts(p), [[0, 0], [0, 10], [10, 10], [10, 0], [10, -10], [20, -10], [20, 0]])
    (q,) = pydiffvg.from_svg_path('M 0 0 Q 5 10 10 0 T 20 0')
    np.testing.assert_array_equal(_pts(q), [[0, 0], [5, 10], [10, 0], [15, -10], [20, 0]])
    # without a preceding C/S (Q/T) the first control is the current point
    (s,) = pydiffvg.from_svg_path('M 0 0 L 5 5 S 10 10 15 5')
    np.testing.assert_array_equal(_pts(s)[2], [5, 5])
    (t,) = pydiffvg.from_svg_path('M 0 0 L 5 5 T 15 5')
    np.testing.assert_array_equal(_pts(t)[2], [5, 5])


def test_subpaths_split_and_relative_move_after_close():
    ps = pydiffvg.from_svg_path('M 0 0 L 10 0 L 10 10 Z m 20 0 l 5 0 l 0 5 z')
    assert len(ps) == 2
    assert ps[0].is_closed and ps[1].is_closed
    np.testing.assert_array_equal(_pts(ps[1]), [[20, 0], [25, 0], [25, 5]])


def test_close_snapping_and_force_close():
    # an explicit return to the start makes the Z line zero-length: it is dropped
    (p,) = pydiffvg.from_svg_path('M 0 0 L 10 0 L 10 10 L 0 0 Z')
    assert p.is_closed and _ncp(p) == [0, 0, 0]
    # an end within 1e-5 of the start snaps shut
    (q,) = pydiffvg.from_svg_path('M 0 0 L 10 0 L 10 10 L 0.000001 0')
    assert q.is_closed and _ncp(q) == [0, 0, 0]
    # open paths stay open unless forced (filled shapes are forced)
    (r,) = pydiffvg.from_svg_path('M 0 0 L 10 0 L 10 10')
    assert not r.is_closed and _ncp(r) == [0, 0]
    (f,) = pydiffvg.from_svg_path('M 0 0 L 10 0 L 10 10', force_close=True)
    assert f.is_closed and _ncp(f) == [0, 0, 0]
    np.testing.assert_array_equal(_pts(f), [[0, 0], [10, 0], [10, 10]])


def test_shape_to_canvas_applied_to_points():
    m = torch.tensor([[2.0, 0.0, 5.0], [0.0, 3.0, -1.0], [0.0, 0.0, 1.0]])
    (p,) = pydiffvg.from_svg_path('M 1 1 L 2 3', m)
    np.testing.assert_array_equal(_pts(p), [[7, 2], [9, 8]])


def test_number_forms():
    (p,) = pydiffvg.from_svg_path('M-1.5.5L1e1-2E0l.5.5')
    np.testing.assert_array_equal(_pts(p), [[-1.5, 0.5], [10, -2], [10.5, -1.5]])


def test_bad_path_data_raises():
    with pytest.raises(ValueError):
        pydiffvg.from_svg_path('10 10 L 20 20')
    with pytest.raises(ValueError):
        pydiffvg.from_svg_path('M 10 10 L 20')


# ---- arcs ------------------------------------------------------------------------------

def _cubic(p0, p1, p2, p3, t):
    s = 1 - t
    return s ** 3 * p0 + 3 * s * s * t * p1 + 3 * s * t * t * p2 + t ** 3 * p3


def _arc_cubics(path):
    pts = _pts(path)
    assert set(_ncp(path)) == {2}
    return [pts[3 * k:3 * k + 4] for k in range(len(_ncp(path)))]


@pytest.mark.parametrize('sweep, mid_y', [(1, -10.0), (0, 10.0)])
def test_semicircle_arc_direction_and_accuracy(sweep, mid_y):
    (p,) = pydiffvg.from_svg_path('M 0 0 A 10 10 0 0 %d 20 0' % sweep)
    cubics = _arc_cubics(p)
    assert len(cubics) == 2  # 180 degrees in quarter turns
    c = np.array([10.0, 0.0])
    for q in cubics:
        for t in np.linspace(0, 1, 33):
            r = np.linalg.norm(_cubic(*q, t) - c)
            assert abs(r - 10.0) < 3e-3  # quarter-turn cubics deviate up to ~2.5e-4 r
        assert abs(np.linalg.norm(q[0] - c) - 10) < 1e-4 and abs(np.linalg.norm(q[3] - c) - 10) < 1e-4
    # the joint between the two pieces is the arc's midpoint (10, mid_y)
    np.testing.assert_allclose(cubics[0][3], [10.0, mid_y], atol=1e-4)
    np.testing.assert_array_equal(cubics[-1][3], [20, 0])


@pytest.mark.parametrize('large, sweep, centre', [(0, 1, (0, 10)), (1, 0, (0, 10)), (1, 1, (10, 0)),
                                                  (0, 0, (10, 0))])
def test_quarter_arc_centre_choice(large, sweep, centre):
    # circle of radius 10 through (0,0) and (10,10): centres (10,0) and (0,10)
    seg = svg_path.parse_path('M 0 0 A 10 10 0 %d %d 10 10' % (large, sweep))[0]
    assert abs(seg.center - complex(*centre)) < 1e-9
    assert abs(abs(seg.delta) - (270.0 if large else 90.0)) < 1e-9
    assert (seg.delta > 0) == bool(sweep)
    (p,) = pydiffvg.from_svg_path('M 0 0 A 10 10 0 %d %d 10 10' % (large, sweep))
    assert len(_ncp(p)) == (3 if large else 1)
    for q in _arc_cubics(p):
        for t in np.linspace(0, 1, 17):
            assert abs(np.linalg.norm(_cubic(*q, t) - np.array(centre)) - 10.0) < 3e-3


def test_arc_radius_scaled_up_when_too_small():
    seg = svg_path.parse_path('M 0 0 A 1 1 0 0 1 20 0')[0]
    assert abs(seg.radius - complex(10, 10)) < 1e-12
    assert abs(seg.center - complex(10, 0)) < 1e-9
    assert abs(seg.delta - 180.0) < 1e-9


def test_ellipse_arc_points_on_ellipse():
    (p,) = pydiffvg.from_svg_path('M 0 0 A 20 10 0 1 1 20 10')
    seg = svg_path.parse_path('M 0 0 A 20 10 0 1 1 20 10')[0]
    cx, cy = seg.center.real, seg.center.imag
    for q in _arc_cubics(p):
        for v in (q[0], q[3]):
            assert abs(((v[0] - cx) / 20) ** 2 + ((v[1] - cy) / 10) ** 2 - 1) < 1e-5


def test_zero_radius_arc_is_a_line_and_compact_flags():
    (p,) = pydiffvg.from_svg_path('M 0 0 A 0 5 0 0 1 10 10')
    assert _ncp(p) == [0]
    a = pydiffvg.from_svg_path('M0 0a10 10 0 0010 10')[0]
    b = pydiffvg.from_svg_path('M 0 0 a 10 10 0 0 0 10 10')[0]
    np.testing.assert_array_equal(_pts(a), _pts(b))


def test_full_circle_from_two_arcs_closes():
    (p,) = pydiffvg.from_svg_path('M 0 10 A 10 10 0 0 1 20 10 A 10 10 0 0 1 0 10 Z')
    assert p.is_closed
    assert len(_ncp(p)) == 4
    area = 0.0  # shoelace over a dense sampling: pi r^2
    poly = np.concatenate([np.array([_cubic(*q, t) for t in np.linspace(0, 1, 200)[:-1]])
                           for q in _arc_cubics_closed(p)])
    x, y = poly[:, 0], poly[:, 1]
    area = 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))
    assert abs(area - math.pi * 100) < 0.3  # 0.1%: the cubics bulge by at most 2.5e-4 r


def _arc_cubics_closed(path):
    pts = _pts(path)
    pts = np.vstack([pts, pts[:1]])
    return [pts[3 * k:3 * k + 4] for k in range(len(_ncp(path)))]


# ---- transforms and colours ---------------------------------------------------------------

def test_transforms_closed_form():
    m = parse_transform('translate(10,20) scale(2) rotate(90)').numpy()
    np.testing.assert_allclose(m, [[0, -2, 10], [2, 0, 20], [0, 0, 1]], atol=1e-6)
    m = parse_transform('rotate(90 5 5)').numpy()
    np.testing.assert_allclose(m @ [5, 5, 1], [5, 5, 1], atol=1e-6)
    np.testing.assert_allclose(m @ [6, 5, 1], [5, 6, 1], atol=1e-6)
    m = parse_transform('matrix(1 2 3 4 5 6)').numpy()
    np.testing.assert_array_equal(m, [[1, 3, 5], [2, 4, 6], [0, 0, 1]])
    m = parse_transform('skewX(45) skewY(45)').numpy()
    np.testing.assert_allclose(m, [[2, 1, 0], [1, 1, 0], [0, 0, 1]], atol=1e-6)
    np.testing.assert_array_equal(parse_transform('scale(2, 3)').numpy(), np.diag([2, 3, 1]))


def test_colours():
    np.testing.assert_allclose(parse_color('#f80', {}).numpy(), [1, 0x88 / 255, 0, 1], rtol=1e-6)
    np.testing.assert_allclose(parse_color('#102030', {}).numpy(), [16 / 255, 32 / 255, 48 / 255, 1], rtol=1e-6)
    np.testing.assert_allclose(parse_color('rgb(255, 0, 51)', {}).numpy(), [1, 0, 0.2, 1], rtol=1e-6)
    np.testing.assert_allclose(parse_color('red', {}).numpy(), [1, 0, 0, 1])
    np.testing.assert_allclose(parse_color('steelblue', {}).numpy(), [70 / 255, 130 / 255, 180 / 255, 1], rtol=1e-6)
    assert parse_color('none', {}) is None
    marker = object()
    assert parse_color('url(#g1)', {'g1': marker}) is marker


# ---- documents --------------------------------------------------------------------------

DOC = '''<?xml version="1.0"?>
<svg xmlns="http://www.w3.org/2000/svg" xmlns:xlink="http://www.w3.org/1999/xlink"
     width="200px" height="100px">
  <defs>
    <linearGradient id="lg" x1="0" y1="0" x2="100" y2="0">
      <stop offset="0" stop-color="#ff0000"/>
      <stop offset="1" style="stop-color:#0000ff;stop-opacity:0.5"/>
    </linearGradient>
    <radialGradient id="rg" cx="50" cy="50" r="20">
      <stop offset="0" stop-color="#00ff00"/>
      <stop offset="1" stop-color="#000000"/>
    </radialGradient>
    <style>
      /* classes */
      .thick { stroke: #00ff00; stroke-width: 6 }
    </style>
  </defs>
  <path id="a" d="M 10 10 L 50 10 L 50 50" fill="#ff0000" fill-opacity="0.5" fill-rule="evenodd"/>
  <g transform="translate(100,0)" fill="#0000ff">
    <path d="M 0 0 L 10 0 L 10 10 M 20 20 L 30 20 L 30 30"/>
    <circle cx="5" cy="6" r="7" opacity="0.25"/>
  </g>
  <path d="M 0 60 L 40 60" fill="none" class="thick"/>
  <path d="M 0 80 L 40 80 L 40 90" style="fill:url(#lg);stroke:#000;stroke-width:2px"/>
  <rect x="30" y="40" width="20" height="10" fill="url(#rg)"/>
  <polygon points="1,2 3,4 5,6" stroke="#123456" stroke-width="3" fill="none"/>
  <line x1="0" y1="0" x2="5" y2="5" stroke="#000000"/>
  <ellipse cx="1" cy="1" rx="2" ry="3"/>
</svg>
'''


@pytest.fixture()
def doc(tmp_path):
    f = tmp_path / 'doc.svg'
    f.write_text(DOC)
    return str(f)


def test_document_structure(doc):
    w, h, shapes, groups = pydiffvg.svg_to_scene(doc)
    assert (w, h) == (200, 100)
    # path a; two subpaths in <g>; circle; class path; style path; rect; polygon; line
    # (<ellipse> is skipped)
    assert len(groups) == 8
    assert [len(g.shape_ids) for g in groups] == [1, 2, 1, 1, 1, 1, 1, 1]
    kinds = [type(s).__name__ for s in shapes]
    assert kinds == ['Path', 'Path', 'Path', 'Circle', 'Path', 'Path', 'Rect', 'Polygon', 'Polygon']


def test_document_attributes(doc):
    w, h, shapes, groups = pydiffvg.svg_to_scene(doc)
    ga = groups[0]
    np.testing.assert_allclose(ga.fill_color.numpy(), [1, 0, 0, 0.5])
    assert ga.use_even_odd_rule is True and ga.stroke_color is None
    assert float(shapes[0].stroke_width) == 0.5  # default half width
    assert shapes[0].is_closed  # filled paths are closed
    # group transform baked into path points; group fill inherited
    np.testing.assert_array_equal(_pts(shapes[1]), [[100, 0], [110, 0], [110, 10]])
    np.testing.assert_array_equal(_pts(shapes[2]), [[120, 20], [130, 20], [130, 30]])
    # the inherited fill tensor is shared; the circle's opacity is written into it
    assert groups[1].fill_color is groups[2].fill_color
    np.testing.assert_allclose(groups[2].fill_color.numpy(), [0, 0, 1, 0.25])
    np.testing.assert_allclose(groups[2].shape_to_canvas.numpy(), [[1, 0, 100], [0, 1, 0], [0, 0, 1]])
    # CSS class: stroke colour and half width; fill none leaves the path open
    gc = groups[3]
    assert gc.fill_color is None
    np.testing.assert_allclose(gc.stroke_color.numpy(), [0, 1, 0, 1])
    assert float(shapes[4].stroke_width) == 3.0
    assert not shapes[4].is_closed
    # style attribute with a linear gradient fill
    gs = groups[4]
    assert isinstance(gs.fill_color, pydiffvg.LinearGradient)
    np.testing.assert_allclose(gs.fill_color.begin.numpy(), [0, 0])
    np.testing.assert_allclose(gs.fill_color.end.numpy(), [100, 0])
    np.testing.assert_allclose(gs.fill_color.offsets.numpy(), [0, 1])
    np.testing.assert_allclose(gs.fill_color.stop_colors.numpy(), [[1, 0, 0, 1], [0, 0, 1, 0.5]])
    np.testing.assert_allclose(gs.stroke_color.numpy(), [0, 0, 0, 1])
    assert float(shapes[5].stroke_width) == 1.0
    # rect spans (0,0)-(width,height); radial gradient keeps centre/radius at zero
    r = shapes[6]
    np.testing.assert_array_equal(r.p_min.numpy(), [0, 0])
    np.testing.assert_array_equal(r.p_max.numpy(), [20, 10])
    rg = groups[5].fill_color
    assert isinstance(rg, pydiffvg.RadialGradient)
    np.testing.assert_array_equal(rg.center.numpy(), [0, 0])
    np.testing.assert_array_equal(rg.radius.numpy(), [0, 0])
    np.testing.assert_allclose(rg.stop_colors.numpy(), [[0, 1, 0, 1], [0, 0, 0, 1]])
    # polygon and line
    np.testing.assert_array_equal(_pts(shapes[7]), [[1, 2], [3, 4], [5, 6]])
    assert not shapes[7].is_closed
    assert float(shapes[7].stroke_width) == 1.5
    assert not shapes[8].is_closed


def test_view_box_sets_canvas(tmp_path):
    f = tmp_path / 'vb.svg'
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" width="10" height="10" viewBox="0 0 64 48">'
                 '<path d="M 0 0 L 1 0 L 1 1 Z"/></svg>')
    w, h, shapes, groups = pydiffvg.svg_to_scene(str(f))
    assert (w, h) == (64, 48)
    np.testing.assert_allclose(groups[0].fill_color.numpy(), [0, 0, 0, 1])


# ---- saving ------------------------------------------------------------------------------

def _scene():
    rng = np.random.default_rng(7)
    shapes, groups = [], []
    pts = torch.tensor(rng.uniform(0, 100, (9, 2)), dtype=torch.float32)
    shapes.append(pydiffvg.Path(num_control_points=torch.tensor([2, 2, 2]), points=pts, is_closed=True,
                                stroke_width=torch.tensor(1.25)))
    groups.append(pydiffvg.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=torch.tensor([0.2, 0.4, 0.6, 0.8]),
                                      stroke_color=torch.tensor([1.0, 0.0, 0.0, 0.5])))
    pts = torch.tensor(rng.uniform(0, 100, (7, 2)), dtype=torch.float32)
    shapes.append(pydiffvg.Path(num_control_points=torch.tensor([1, 0, 2]), points=pts, is_closed=False,
                                stroke_width=torch.tensor(2.0)))
    groups.append(pydiffvg.ShapeGroup(shape_ids=torch.tensor([1]), fill_color=None,
                                      stroke_color=torch.tensor([0.0, 0.0, 1.0, 1.0])))
    shapes.append(pydiffvg.Circle(radius=torch.tensor(12.5), center=torch.tensor([40.0, 30.0])))
    lg = pydiffvg.LinearGradient(begin=torch.tensor([0.0, 0.0]), end=torch.tensor([50.0, 10.0]),
                                 offsets=torch.tensor([0.0, 1.0]),
                                 stop_colors=torch.tensor([[1.0, 1.0, 0.0, 1.0], [0.0, 1.0, 1.0, 0.5]]))
    groups.append(pydiffvg.ShapeGroup(shape_ids=torch.tensor([2]), fill_color=lg))
    return shapes, groups


def test_save_svg_text(tmp_path):
    shapes, groups = _scene()
    f = tmp_path / 'out.svg'
    pydiffvg.save_svg(str(f), 100, 80, shapes, groups)
    text = f.read_text()
    assert text.startswith('<?xml version="1.0" ?>\n<svg xmlns="http://www.w3.org/2000/svg" version="1.1" '
                           'width="100" height="80">')
    assert '<linearGradient id="shape_2_fill" x1="0.0" y1="0.0" x2="50.0" y2="10.0">' in text
    assert 'stop-color="rgb(0, 255, 255)" stop-opacity="0.5"' in text
    assert 'fill="rgb(51, 102, 153)" opacity="0.8"' in text
    assert 'stroke="rgb(255, 0, 0)" stroke-opacity="0.5" stroke-linecap="round" stroke-linejoin="round"' in text
    assert '<circle r="12.5" cx="40.0" cy="30.0" stroke-width="2.0" fill="url(#shape_2_fill)"/>' in text


def test_save_load_round_trip(tmp_path):
    shapes, groups = _scene()
    f = tmp_path / 'rt.svg'
    pydiffvg.save_svg(str(f), 100, 80, shapes, groups)
    w, h, s2, g2 = pydiffvg.svg_to_scene(str(f))
    assert (w, h) == (100, 80)
    assert len(s2) == 3 and len(g2) == 3
    # closed path: exact float32 points, same segment kinds
    assert _ncp(s2[0]) == [2, 2, 2] and s2[0].is_closed
    np.testing.assert_array_equal(s2[0].points.numpy(), shapes[0].points.numpy())
    assert float(s2[0].stroke_width) == 1.25
    np.testing.assert_allclose(g2[0].fill_color.numpy(), [51 / 255, 102 / 255, 153 / 255, 0.8], rtol=1e-6)
    # open stroke path survives open (no fill)
    assert _ncp(s2[1]) == [1, 0, 2] and not s2[1].is_closed
    np.testing.assert_array_equal(s2[1].points.numpy(), shapes[1].points.numpy())
    assert g2[1].fill_color is None
    assert float(s2[2].radius) == 12.5
    np.testing.assert_array_equal(s2[2].center.numpy(), [40, 30])
