"""SVG documents loaded with diffvg_amd.svg_to_scene render on the GPU like the oracle
renders them: arcs (converted to cubics), quadratic and smooth segments, group
transforms, multi-subpath fills, polygons, circles, strokes and a linear gradient, in
the forward pass and the backward pass, under both schedules."""
import numpy as np
import pytest

import diffvg_amd as pd
from diffvg_amd import _lib
from gpudrive import GpuRunner
from refdrive import OracleRunner, flatten
from test_gpu_parity import IMAGE_TOL, MODES, _compare

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings('ignore')]

DOC = '''<svg xmlns="http://www.w3.org/2000/svg" width="96" height="80">
  <defs>
    <linearGradient id="g" x1="10" y1="0" x2="80" y2="0">
      <stop offset="0" stop-color="#ff8000"/><stop offset="0.6" stop-color="#0080ff" stop-opacity="0.7"/>
      <stop offset="1" stop-color="#20c040"/>
    </linearGradient>
  </defs>
  <path d="M 10 40 A 25 18 0 1 1 60 40 A 25 18 0 0 1 10 40 Z" fill="url(#g)" fill-rule="evenodd"/>
  <g transform="translate(40 10) rotate(20) scale(0.8)" fill="#3366cc" opacity="0.8">
    <path d="M 0 0 q 20 -15 40 0 t 40 0 v 30 h -80 z m 20 8 c 5 -6 15 -6 20 0 s 15 6 20 0 l 0 10 l -40 0 z"/>
  </g>
  <path d="M 5 70 C 25 50 45 90 65 70 S 90 50 92 75" fill="none" stroke="#aa2211" stroke-width="3"/>
  <polygon points="70,5 90,15 85,35 65,25" fill="#55aa33" stroke="#000000" stroke-width="1.5"/>
  <circle cx="20" cy="15" r="9" fill="#9933aa"/>
  <path d="M 50 55 a 8 8 0 1 0 16 0 a 8 8 0 1 0 -16 0" fill="#118877" fill-opacity="0.6"/>
</svg>
'''


@pytest.fixture(scope='module')
def flat(tmp_path_factory):
    f = tmp_path_factory.mktemp('svg') / 'doc.svg'
    f.write_text(DOC)
    w, h, shapes, groups = pd.svg_to_scene(str(f))
    assert len(groups) == 6
    return flatten(w, h, shapes, groups, 1, 1.0)


@pytest.mark.parametrize('mode', sorted(MODES))
def test_svg_scene_matches_oracle(flat, mode):
    print('svg scene', mode, flush=True)
    w, h = flat['canvas_width'], flat['canvas_height']
    o, g = OracleRunner(flat), GpuRunner(flat, MODES[mode])
    for pre in (False, True):
        kw = dict(prefilter=pre)
        _compare(g.render(w, h, 2, 2, 3, **kw), o.render(w, h, 2, 2, 3, **kw), ('svg', pre, 'fwd'), IMAGE_TOL[mode])
    d = np.random.default_rng(5).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    kw = dict(d_image=d, translation=True)
    _compare(g.render(w, h, 2, 2, 3, **kw), o.render(w, h, 2, 2, 3, **kw), ('svg', 'bwd'), IMAGE_TOL[mode])


def test_svg_scene_sdf_matches_oracle(flat):
    print('svg sdf', flush=True)
    w, h = flat['canvas_width'], flat['canvas_height']
    o, g = OracleRunner(flat), GpuRunner(flat, _lib.MODE_TILED)
    _compare(g.render(w, h, 1, 1, 0, output='sdf'), o.render(w, h, 1, 1, 0, output='sdf'), ('svg', 'sdf'))


def _saved_scene():
    """A scene whose every value survives save_svg's text exactly: colours are multiples
    of 1/255, opacities and coordinates are float32 values printed in full, the fill rule
    is SVG's default (non-zero), the rect sits at the origin.  Paths (cubic; mixed line/quadratic/cubic open stroke), a
    circle with a three-stop linear gradient, a filled and stroked rect, a polygon."""
    import torch
    T = lambda x: torch.tensor(x, dtype=torch.float32)
    c255 = lambda r, g, b, a: T([r / 255, g / 255, b / 255, a])
    shapes = [
        pd.Path(num_control_points=torch.tensor([2, 2, 2, 2], dtype=torch.int32),
                points=T([[20, 10], [45, 5], [70, 20], [75, 40], [80, 60], [50, 75], [30, 70], [10, 65], [5, 40],
                          [10, 30], [12, 20], [15, 15]]), is_closed=True, stroke_width=T(1.25)),
        pd.Path(num_control_points=torch.tensor([1, 0, 2], dtype=torch.int32),
                points=T([[10, 85], [30, 60], [50, 88], [60, 70], [70, 95], [85, 60], [92, 80]]), is_closed=False,
                stroke_width=T(2.0)),
        pd.Circle(radius=T(12.5), center=T([40.0, 30.0])),
        # at the origin: the reference's loader drops a rect's x and y (parse_svg.py:496-505)
        pd.Rect(p_min=T([0.0, 0.0]), p_max=T([30.25, 25.25]), stroke_width=T(1.5)),
        pd.Polygon(points=T([[62, 40], [95, 45], [88, 75], [66, 68]]), is_closed=True, stroke_width=T(0.5)),
    ]
    lg = pd.LinearGradient(begin=T([28.0, 20.0]), end=T([55.0, 40.0]), offsets=T([0.0, 0.4, 1.0]),
                           stop_colors=T([[1.0, 1.0, 0.0, 1.0], [0.0, 1.0, 1.0, 0.5], [102 / 255, 0, 1.0, 0.25]]))
    paints = [(c255(51, 102, 153, 0.75), c255(255, 0, 0, 0.5)), (None, c255(0, 0, 255, 1.0)), (lg, None),
              (c255(0, 153, 51, 0.5), c255(17, 34, 51, 1.0)), (c255(204, 51, 0, 0.625), None)]
    groups = [pd.ShapeGroup(shape_ids=torch.tensor([i]), fill_color=f, stroke_color=s, use_even_odd_rule=False)
              for i, (f, s) in enumerate(paints)]
    return shapes, groups


def test_saved_and_reloaded_scene_matches_oracle(tmp_path):
    """save_svg -> svg_to_scene -> GPU render equals the oracle's render of the scene as
    it was before saving (the reloaded description is identical array for array):
    forward, prefiltered forward, backward with d_translation, and the SDF."""
    print('svg save/load round trip', flush=True)
    shapes, groups = _saved_scene()
    path = tmp_path / 'saved.svg'
    pd.save_svg(str(path), 100, 100, shapes, groups)
    w, h, s2, g2 = pd.svg_to_scene(str(path))
    assert (w, h) == (100, 100)
    orig, back = flatten(100, 100, shapes, groups, 0, 0.5), flatten(w, h, s2, g2, 0, 0.5)
    for k, v in orig.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, back[k]), k
        else:
            assert v == back[k], k
    o, g = OracleRunner(orig), GpuRunner(back, _lib.MODE_TILED)
    for pre in (False, True):
        _compare(g.render(w, h, 2, 2, 4, prefilter=pre), o.render(w, h, 2, 2, 4, prefilter=pre), ('saved', pre))
    d = np.random.default_rng(6).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    kw = dict(d_image=d, translation=True)
    _compare(g.render(w, h, 2, 2, 4, **kw), o.render(w, h, 2, 2, 4, **kw), ('saved', 'bwd'))
    _compare(g.render(w, h, 1, 1, 0, output='sdf'), o.render(w, h, 1, 1, 0, output='sdf'), ('saved', 'sdf'))
