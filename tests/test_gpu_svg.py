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
