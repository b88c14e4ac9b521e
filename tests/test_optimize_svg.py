"""OptimizableSvg and SvgOptimizationSettings (diffvg_amd.optimize_svg) on the CPU.

Pins:
- the transform decomposition against apps/svg_parse_test.py:12-22's own check
  (recompose(decompose(M)) == M for random matrices) and against an independent
  computation through numpy's QR factorisation with signs normalised;
- SVG transform functions against their matrices written out by hand;
- the scene the document tree builds (transform composition through nested groups,
  opacity products, paint inheritance, gradient end points in canvas space) against
  values derived by hand for a small document;
- write_xml -> parse again giving the same scene, before and after parameters move.
"""
import io
import json
import math

import numpy as np
import pytest
import torch

import diffvg_amd as pydiffvg
from diffvg_amd import optimize_svg as osv
from diffvg_amd.optimize_svg import (OptimizableSvg, ParameterBank, SvgOptimizationSettings, TransformTools,
                                     colour_hex, parse_colour, parse_length, parse_opacity,
                                     parse_transform_list, transform_features)

pytestmark = pytest.mark.filterwarnings('ignore')


def _rot(deg):
    r = math.radians(deg)
    return np.array([[math.cos(r), -math.sin(r), 0], [math.sin(r), math.cos(r), 0], [0, 0, 1.]])


def _tr(x, y):
    return np.array([[1, 0, x], [0, 1, y], [0, 0, 1.]])


def _sc(x, y):
    return np.diag([x, y, 1.])


# ---- transforms ---------------------------------------------------------------------------

def test_decompose_recompose_round_trip():
    """apps/svg_parse_test.py:12-22 over random 2x3 matrices in [-1, 1]."""
    rng = np.random.default_rng(0)
    for _ in range(2000):
        M = np.eye(3)
        M[0:2, :] = (rng.random((2, 3)) - 0.5) * 2
        d = TransformTools.decompose(M)
        assert d.scale[0] > 0
        np.testing.assert_allclose(TransformTools.recompose_numpy(d.theta, d.scale, d.shear, d.translate), M,
                                   atol=1e-9)
        back = TransformTools.recompose(torch.tensor(d.theta, dtype=torch.float32),
                                        torch.tensor(d.scale, dtype=torch.float32),
                                        torch.tensor(d.shear, dtype=torch.float32),
                                        torch.tensor(d.translate, dtype=torch.float32)).numpy()
        assert np.linalg.norm(M - back) < 1e-3


def test_decompose_matches_sign_normalised_qr():
    """The same factorisation through numpy's QR: A = Q R, with the signs of Q's columns
    chosen so that R's diagonal is positive, then Q's reflection (if any) moved into the
    second scale."""
    rng = np.random.default_rng(1)
    for _ in range(500):
        A = (rng.random((2, 2)) - 0.5) * 4
        q, r = np.linalg.qr(A)
        s = np.diag(np.sign(np.diag(r)))
        q, r = q @ s, s @ r                       # now diag(r) > 0
        if np.linalg.det(q) < 0:                  # q = rot @ diag(1, -1)
            q, r = q @ np.diag([1., -1.]), np.diag([1., -1.]) @ r
        theta = math.atan2(q[1, 0], q[0, 0])
        M = np.eye(3)
        M[0:2, 0:2] = A
        d = TransformTools.decompose(M)
        assert math.isclose(d.theta, theta, abs_tol=1e-9)
        np.testing.assert_allclose(d.scale, [r[0, 0], r[1, 1]], atol=1e-9)
        assert math.isclose(d.shear, r[0, 1] / r[0, 0], abs_tol=1e-9)


@pytest.mark.parametrize('M, theta, scale, shear, tr', [
    (_rot(30), math.radians(30), [1, 1], 0, [0, 0]),
    (_sc(2, 3), 0, [2, 3], 0, [0, 0]),
    (_sc(1, -1), 0, [1, -1], 0, [0, 0]),
    (_sc(-2, -2), math.pi, [2, 2], 0, [0, 0]),
    (np.array([[1, 0.5, 0], [0, 1, 0], [0, 0, 1.]]), 0, [1, 1], 0.5, [0, 0]),
    (_rot(90) @ _tr(3, 4), math.pi / 2, [1, 1], 0, [3, 4]),
])
def test_decompose_closed_forms(M, theta, scale, shear, tr):
    d = TransformTools.decompose(M)
    assert math.isclose(d.theta, theta, abs_tol=1e-12)
    np.testing.assert_allclose(d.scale, scale, atol=1e-12)
    assert math.isclose(d.shear, shear, abs_tol=1e-12)
    np.testing.assert_allclose(d.translate, tr, atol=1e-12)


def test_transform_features():
    p = transform_features(TransformTools.decompose(_rot(10) @ _sc(2, 2) @ _tr(1, 0)))
    assert p['rotation'] and p['scale'] and not p['mirror'] and p['uniform']
    assert not p['shear'] and p['translation']
    p = transform_features(TransformTools.decompose(_sc(1, -2)))
    assert p['mirror'] and not p['uniform'] and not p['rotation']
    p = transform_features(TransformTools.decompose(np.eye(3)))
    assert p['uniform'] and not any(v for k, v in p.items() if k != 'uniform')


@pytest.mark.parametrize('text, M', [
    ('translate(10 20)', _tr(10, 20)),
    ('translate(10)', _tr(10, 0)),
    ('scale(2)', _sc(2, 2)),
    ('scale(2,3)', _sc(2, 3)),
    ('rotate(30)', _rot(30)),
    ('rotate(90 10 0)', _tr(10, 0) @ _rot(90) @ _tr(-10, 0)),
    ('matrix(1 2 3 4 5 6)', np.array([[1, 3, 5], [2, 4, 6], [0, 0, 1.]])),
    ('skewX(0.5)', np.array([[1, 0.5, 0], [0, 1, 0], [0, 0, 1.]])),
    ('skewY(0.25)', np.array([[1, 0, 0], [0.25, 1, 0], [0, 0, 1.]])),
    ('translate(1,2) scale(3)', _tr(1, 2) @ _sc(3, 3)),
    ('translate(1,2)rotate(45)', _tr(1, 2) @ _rot(45)),
])
def test_parse_transform(text, M):
    np.testing.assert_allclose(parse_transform_list(text), M, atol=1e-12)


def test_parse_transform_errors():
    for bad in ('spin(3)', 'translate(1 2 3)', 'rotate(1 2)', 'scale 2'):
        with pytest.raises(ValueError):
            parse_transform_list(bad)


def test_rotate_about_centre_keeps_it_fixed():
    M = parse_transform_list('rotate(37 12 -5)')
    np.testing.assert_allclose(M @ [12, -5, 1], [12, -5, 1], atol=1e-12)


def test_lengths_opacities_colours():
    assert parse_length('3') == 3.0
    assert parse_length('2.5px') == 2.5
    assert parse_length('2mm') == 8.0
    assert parse_length('1in') == pytest.approx(101.6)
    assert parse_length('72pt') == pytest.approx(101.6)
    with pytest.raises(ValueError):
        parse_length('3em')
    assert parse_opacity('50%') == 0.5
    assert parse_opacity('1.5') == 1.0
    assert parse_opacity('-1') == 0.0
    np.testing.assert_allclose(parse_colour('#ff8000').numpy(), [1, 128 / 255, 0], rtol=1e-6)
    np.testing.assert_allclose(parse_colour('#f80').numpy(), [1, 8 / 15, 0], rtol=1e-6)
    for bad in ('red', '#12345', '#ggg'):
        with pytest.raises(ValueError):
            parse_colour(bad)
    assert colour_hex(torch.tensor([1.0, 0.5, 1.2])) == '#ff7fff'


# ---- documents ---------------------------------------------------------------------------

DOC = """<svg xmlns="http://www.w3.org/2000/svg" xmlns:xlink="http://www.w3.org/1999/xlink"
     version="2.0" width="64" height="48">
  <style type="text/css">
    .warm { fill: #ff0000; fill-opacity: 0.5 }
  </style>
  <defs>
    <linearGradient id="base">
      <stop offset="1" stop-color="#0000ff"/>
      <stop offset="0" stop-color="#00ff00" stop-opacity="0.5"/>
    </linearGradient>
    <linearGradient id="lin" xlink:href="#base" x1="0" y1="0" x2="10" y2="0"
                    gradientTransform="translate(1 2)"/>
  </defs>
  <g id="outer" transform="translate(10 5)" opacity="0.5">
    <g id="inner" transform="scale(2)" style="fill:#00ff00;stroke:#0000ff;stroke-width:3">
      <path id="tri" d="M 0 0 L 8 0 L 4 6 Z"/>
      <circle id="dot" cx="5" cy="6" r="2" class="warm" fill-rule="evenodd"/>
    </g>
    <rect id="box" x="1" y="2" width="3" height="4" fill="url(#lin)"/>
  </g>
  <ellipse id="ell" cx="30" cy="20" rx="4" ry="2" fill="#000080"/>
  <polygon id="poly" points="1,1 5,1 3,4" fill="#808080" opacity="25%"/>
</svg>
"""


@pytest.fixture
def doc(tmp_path):
    f = tmp_path / 'doc.svg'
    f.write_text(DOC)
    return str(f)


def _mat(t):
    return t.detach().numpy().astype(np.float64)


def test_document_scene(doc):
    o = OptimizableSvg(doc)
    w, h, shapes, groups = o.build_scene()
    assert (w, h) == (64, 48)
    assert [type(s).__name__ for s in shapes] == ['Path', 'Circle', 'Rect', 'Ellipse', 'Polygon']
    assert [g.id for g in groups] == ['tri', 'dot', 'box', 'ell', 'poly']
    assert [g.shape_ids.tolist() for g in groups] == [[0], [1], [2], [3], [4]]
    inner = _tr(10, 5) @ _sc(2, 2)
    np.testing.assert_allclose(_mat(groups[0].shape_to_canvas), inner, atol=1e-6)
    np.testing.assert_allclose(_mat(groups[1].shape_to_canvas), inner, atol=1e-6)
    np.testing.assert_allclose(_mat(groups[2].shape_to_canvas), _tr(10, 5), atol=1e-6)
    np.testing.assert_allclose(_mat(groups[3].shape_to_canvas), np.eye(3), atol=1e-6)
    # path: inherited green fill at the outer group's opacity 0.5, blue stroke, width 3
    np.testing.assert_allclose(groups[0].fill_color.detach().numpy(), [0, 1, 0, 0.5])
    np.testing.assert_allclose(groups[0].stroke_color.detach().numpy(), [0, 0, 1, 0.5])
    assert float(shapes[0].stroke_width) == 3.0
    assert shapes[0].is_closed and shapes[0].num_control_points.tolist() == [0, 0, 0]
    np.testing.assert_array_equal(shapes[0].points.detach().numpy(), [[0, 0], [8, 0], [4, 6]])
    # circle: class red fill with fill-opacity 0.5 times opacity 0.5, even-odd from its attribute
    np.testing.assert_allclose(groups[1].fill_color.detach().numpy(), [1, 0, 0, 0.25])
    assert groups[1].use_even_odd_rule and not groups[0].use_even_odd_rule
    np.testing.assert_allclose(shapes[1].center.numpy(), [5, 6])
    assert float(shapes[1].radius) == 2.0
    # rect: p_max = p_min + size; its gradient uses base's stops sorted by offset, alphas
    # times the group opacity, end points through the rect's transform then its own
    np.testing.assert_allclose(shapes[2].p_min.numpy(), [1, 2])
    np.testing.assert_allclose(shapes[2].p_max.numpy(), [4, 6])
    g = groups[2].fill_color
    assert isinstance(g, pydiffvg.LinearGradient)
    np.testing.assert_allclose(g.offsets.detach().numpy(), [0, 1])
    np.testing.assert_allclose(g.stop_colors.detach().numpy(), [[0, 1, 0, 0.25], [0, 0, 1, 0.5]])
    np.testing.assert_allclose(g.begin.detach().numpy(), [11, 7], atol=1e-6)
    np.testing.assert_allclose(g.end.detach().numpy(), [21, 7], atol=1e-6)
    assert groups[2].stroke_color is None
    # ellipse at the root: default stroke none, width 0
    np.testing.assert_allclose(shapes[3].radius.numpy(), [4, 2])
    np.testing.assert_allclose(groups[3].fill_color.detach().numpy(), [0, 0, 128 / 255, 1], rtol=1e-6)
    assert float(shapes[3].stroke_width) == 0.0
    np.testing.assert_allclose(groups[4].fill_color.detach().numpy(), [128 / 255, 128 / 255, 128 / 255, 0.25],
                               rtol=1e-6)
    np.testing.assert_array_equal(shapes[4].points.numpy(), [[1, 1], [5, 1], [3, 4]])


def test_serializes_for_render_function(doc):
    o = OptimizableSvg(doc)
    args = pydiffvg.RenderFunction.serialize_scene(*o.build_scene())
    assert args[0] == 64 and args[1] == 48 and args[2] == 5 and args[3] == 5


def test_local_attribute_beats_style_and_class(tmp_path):
    f = tmp_path / 'a.svg'
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="8" height="8">'
                 '<style type="text/css">.c { fill: #0000ff }</style>'
                 '<circle class="c" style="fill:#00ff00" fill="#ff0000" cx="4" cy="4" r="2"/>'
                 '<circle class="c" style="fill:#00ff00" cx="4" cy="4" r="2"/>'
                 '<circle class="c" cx="4" cy="4" r="2"/></svg>')
    _, _, _, groups = OptimizableSvg(str(f)).build_scene()
    cols = [g.fill_color.detach().numpy()[:3].tolist() for g in groups]
    assert cols == [[1, 0, 0], [0, 1, 0], [0, 0, 1]]


def test_style_keys_are_not_stripped(tmp_path):
    """A space after ';' makes the next style key unrecognised, as in the reference's
    parser, so the second declaration is dropped."""
    f = tmp_path / 'a.svg'
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="8" height="8">'
                 '<circle style="fill:#00ff00; stroke:#0000ff" cx="4" cy="4" r="2"/></svg>')
    _, _, _, (g,) = OptimizableSvg(str(f)).build_scene()
    assert g.stroke_color is None


def test_root_appearance_is_not_inherited(tmp_path):
    f = tmp_path / 'a.svg'
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="8" height="8" fill="#ff0000">'
                 '<circle cx="4" cy="4" r="2"/></svg>')
    _, _, _, (g,) = OptimizableSvg(str(f)).build_scene()
    np.testing.assert_array_equal(g.fill_color.detach().numpy(), [0, 0, 0, 1])


def test_view_box_and_errors(tmp_path):
    f = tmp_path / 'a.svg'
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" viewBox="0 0 20.5 10">'
                 '<circle cx="4" cy="4" r="2"/></svg>')
    assert OptimizableSvg(str(f)).canvas.tolist() == [21, 10]
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="8" height="8">'
                 '<circle cx="4" cy="4" r="2" fill="url(#nope)"/></svg>')
    with pytest.raises(ValueError):
        OptimizableSvg(str(f))
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="8" height="8">'
                 '<defs><g/></defs></svg>')
    with pytest.raises(NotImplementedError):
        OptimizableSvg(str(f))


def test_multi_subpath_ids(tmp_path):
    f = tmp_path / 'a.svg'
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="8" height="8">'
                 '<path id="p" d="M 0 0 L 1 0 L 1 1 Z M 3 3 L 4 3 L 4 4 Z"/></svg>')
    _, _, shapes, (g,) = OptimizableSvg(str(f)).build_scene()
    assert [s.id for s in shapes] == ['p-0', 'p-1'] and g.shape_ids.tolist() == [0, 1]


# ---- optimisation ---------------------------------------------------------------------------

def test_trainable_parameters(doc):
    o = OptimizableSvg(doc)
    assert o.node('tri').geometry['paths'][0].points.requires_grad
    # shapes other than paths keep their geometry (not trained)
    assert not o.node('box').geometry['values'].requires_grad
    # transforms (rigid mode by default) train from their decomposition
    outer = o.node('outer').transform
    assert outer.trained and outer.residual is None
    np.testing.assert_allclose(_mat(outer.matrix()), _tr(10, 5), atol=1e-6)
    # the translation's learning rate is scaled by the larger canvas side
    assert o.bank.learning_rates(outer.theta) == pytest.approx([2e-3])
    assert o.bank.learning_rates(outer.scale) == pytest.approx([2e-3])
    assert o.bank.learning_rates(outer.translate) == pytest.approx([2e-3 * 64])
    # colours
    assert o.node('ell').style['fill'].rgb.requires_grad
    # every registered tensor is a leaf that requires grad, registered once
    ts = o.bank.tensors()
    assert all(t.is_leaf and t.requires_grad for t in ts) and len(set(map(id, ts))) == len(ts)


@pytest.mark.parametrize('mode, text, residual', [
    ('move', 'rotate(20) translate(3 4)', False),
    ('move', 'scale(2)', True),
    ('rigid', 'rotate(-30) scale(1.5) translate(1 2)', False),
    ('rigid', 'skewX(0.3)', True),
    ('similarity', 'scale(2 -2) rotate(10)', False),
    ('similarity', 'scale(1 3)', True),
    ('affine', 'matrix(1 0.2 -0.4 1.5 3 -7)', False),
    ('affine', 'scale(1 -1)', False),
])
def test_transform_modes_start_at_the_parsed_transform(tmp_path, mode, text, residual):
    f = tmp_path / 'a.svg'
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="8" height="8">'
                 '<g id="g" transform="{}"><circle cx="4" cy="4" r="2"/></g></svg>'.format(text))
    s = SvgOptimizationSettings()
    s.global_override(["transforms", "transform_mode"], mode)
    o = OptimizableSvg(str(f), s)
    to = o.node('g').transform
    assert (to.residual is not None) == residual
    T = to.matrix()
    np.testing.assert_allclose(_mat(T), parse_transform_list(text), atol=1e-5)
    # gradients reach the parameters
    T.sum().backward()
    assert to.theta.grad is not None and to.translate.grad is not None
    # the written transform parses back to the current one after the parameters move
    with torch.no_grad():
        to.theta += 0.1
        to.translate += torch.tensor([0.5, -0.25])
    np.testing.assert_allclose(parse_transform_list(to.svg()), _mat(to.matrix()), atol=1e-5)


def test_projections_keep_ranges():
    """Colours stay in [1e-4, 1]; stop offsets are clamped to [0, 1], sorted, and their
    ends pinned to 0 and 1 after every step."""
    c = torch.tensor([0.5, 0.0, 1.0])
    s = torch.tensor([0.0, 0.3, 0.6, 1.0])
    bank = ParameterBank()
    bank.add(c, 1.0, 'SGD', osv._project_colour)
    bank.add(s, 1.0, 'SGD', osv._project_stops)
    c.grad = torch.tensor([1.0, 1.0, -1.0])
    s.grad = torch.tensor([-0.5, 0.5, -0.2, 0.0])
    bank.step()
    np.testing.assert_allclose(c.detach().numpy(), [1e-4, 1e-4, 1.0])
    # after the step [0.5, -0.2, 0.8, 1.0]; clamped [0.5, 0, 0.8, 1]; sorted
    # [0, 0.5, 0.8, 1]; ends pinned to 0 and 1
    np.testing.assert_allclose(s.detach().numpy(), [0.0, 0.5, 0.8, 1.0], atol=1e-7)
    # [0, 0.5, 0.8, 0.5] sorts to [0, 0.5, 0.5, 0.8], then the last is pinned to 1
    bank.zero_grad()
    c.grad = torch.zeros(3)
    s.grad = torch.tensor([0.0, 0.0, 0.0, 0.5])
    bank.step()
    np.testing.assert_allclose(s.detach().numpy(), [0.0, 0.5, 0.5, 1.0], atol=1e-7)


def test_one_bank_matches_separate_optimizers():
    """Adam over parameter groups in one optimizer steps each tensor exactly as a
    separate Adam per tensor does."""
    torch.manual_seed(0)
    a, b = torch.randn(5), torch.randn(3)
    a2, b2 = a.clone().requires_grad_(True), b.clone().requires_grad_(True)
    bank = ParameterBank()
    bank.add(a, 0.1, 'Adam')
    bank.add(b, 0.01, 'Adam')
    oa, ob = torch.optim.Adam([a2], lr=0.1), torch.optim.Adam([b2], lr=0.01)
    for _ in range(4):
        ga, gb = torch.randn(5), torch.randn(3)
        a.grad, b.grad, a2.grad, b2.grad = ga.clone(), gb.clone(), ga.clone(), gb.clone()
        bank.step()
        oa.step()
        ob.step()
    assert torch.equal(a, a2) and torch.equal(b, b2)


def test_step_moves_points_and_rebuilds(doc):
    o = OptimizableSvg(doc)
    w, h, shapes, groups = o.build_scene()
    assert o.build_scene()[2] is shapes           # cached until a step
    loss = shapes[0].points.sum() + groups[0].fill_color.sum()
    o.zero_grad()
    loss.backward()
    before = shapes[0].points.detach().clone()
    o.step()
    assert not torch.equal(before, shapes[0].points.detach())
    assert o.build_scene()[2] is not shapes


def test_optimize_alpha_trains_opacity(tmp_path):
    f = tmp_path / 'a.svg'
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="8" height="8">'
                 '<circle cx="4" cy="4" r="2" fill="#ff0000" fill-opacity="0.5"/></svg>')
    s = SvgOptimizationSettings()
    s.global_override(["optimize_alpha"], True)
    s.global_override(["alpha_lr"], 0.1)
    s.global_override(["optimizer"], "SGD")
    o = OptimizableSvg(str(f), s)
    _, _, _, (g,) = o.build_scene()
    g.fill_color[3].backward()
    o.step()
    assert float(o.nodes('circle')[0].style['fill-opacity']) == pytest.approx(0.4)


# ---- writing -------------------------------------------------------------------------------

def _scene_arrays(scene):
    w, h, shapes, groups = scene
    out = [w, h]
    for s in shapes:
        for a in ('points', 'center', 'radius', 'p_min', 'p_max', 'stroke_width'):
            v = getattr(s, a, None)
            if torch.is_tensor(v):
                out.append(v.detach().double().numpy().ravel())
    for g in groups:
        out.append(g.shape_to_canvas.detach().double().numpy().ravel())
        for col in (g.fill_color, g.stroke_color):
            if col is None:
                out.append(None)
            elif torch.is_tensor(col):
                out.append(col.detach().double().numpy())
            else:
                out += [col.begin.detach().double().numpy(), col.end.detach().double().numpy(),
                        col.offsets.detach().double().numpy(), col.stop_colors.detach().double().numpy()]
        out.append(g.use_even_odd_rule)
    return out


def _same(a, b, atol):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            np.testing.assert_allclose(x, y, atol=atol)
        else:
            assert x == y


def test_write_parse_round_trip(doc, tmp_path):
    o = OptimizableSvg(doc)
    text = o.write_xml()
    f = tmp_path / 'out.svg'
    f.write_text(text)
    o2 = OptimizableSvg(str(f))
    # colours are written as bytes (truncated), so compare colours at 1/255
    _same(_scene_arrays(o.build_scene()), _scene_arrays(o2.build_scene()), atol=1 / 255)


def test_write_parse_round_trip_after_training(doc, tmp_path):
    s = SvgOptimizationSettings()
    s.global_override(["optimizer"], "SGD")
    s.global_override(["optimize_color"], False)
    s.global_override(["gradients", "optimize_color"], False)
    s.global_override(["gradients", "optimize_stops"], False)
    o = OptimizableSvg(doc, s)
    torch.manual_seed(0)
    for _ in range(2):
        o.zero_grad()
        w, h, shapes, groups = o.build_scene()
        loss = sum((s_.points * torch.randn_like(s_.points)).sum() for s_ in shapes if hasattr(s_, 'points')
                   and s_.points.requires_grad)
        loss = loss + sum((g.shape_to_canvas * torch.randn(3, 3)).sum() for g in groups)
        loss.backward()
        o.step()
    f = tmp_path / 'out.svg'
    f.write_text(o.write_xml())
    o2 = OptimizableSvg(str(f), SvgOptimizationSettings())
    _same(_scene_arrays(o.build_scene()), _scene_arrays(o2.build_scene()), atol=2e-4)


def test_written_document_text(doc):
    text = OptimizableSvg(doc).write_xml()
    assert '<defs>' in text and '<style type="text/css">.warm {fill: #ff0000;fill-opacity: 0.5}' in text
    assert 'xlink:href="#base"' in text and 'gradientUnits="userSpaceOnUse"' in text
    assert 'd="M 0 0 L 8 0 L 4 6 L 0 0 Z"' in text
    assert 'opacity:0.5' in text and 'tensor' not in text


# ---- settings ---------------------------------------------------------------------------------

def test_settings_store():
    s = SvgOptimizationSettings()
    assert s.retrieve("x") == (s.store["default"], False)
    mine = s.undefault("x")
    mine["paths"]["shape_lr"] = 5.0
    assert s.retrieve("x")[1] and s.store["default"]["paths"]["shape_lr"] == 0.2
    s.global_override(["transforms", "transform_lr"], 0.5)
    assert s.store["x"]["transforms"]["transform_lr"] == 0.5 == s.store["default"]["transforms"]["transform_lr"]
    s.override_optimizer("SGD")
    assert all(v["optimizer"] == "SGD" for v in s.store.values())
    buf = io.StringIO()
    s.save(buf)
    s2 = SvgOptimizationSettings(io.StringIO(buf.getvalue()))
    assert s2.store == json.loads(buf.getvalue()) == s.store
    assert s.reset_to_defaults("x") is s.store["default"] and "x" not in s.store
