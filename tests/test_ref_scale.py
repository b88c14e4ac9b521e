"""The oracle against the reference's zero-filled CPU build (oracle/_ref/zeroinit: the
reference sources with every uninitialised local zero-filled, so its one uninitialised
interval-slot read behaves as the absent slot the oracle defines) on scenes of unusual
scale: 12,000 groups; a tile crowded by 3,138 groups (more than one candidate pass of
the tiled kernels holds); a closed path of 1,500 cubic segments and an open stroke of
600.  Images within 1e-5, every gradient array within 1e-3 of its largest magnitude,
prefiltering off and on.  The reference as built differs from both only on the 600-
segment spiral stroke's point and width gradients (1.9e-2 / 1.2e-2 relative, prefilter
off), through that uninitialised read."""
import numpy as np
import pytest

import scenes
from refdrive import OracleRunner, flatten, ref_build_exists, ref_render_in_child

pytestmark = pytest.mark.skipif(not ref_build_exists('zeroinit'),
                                reason='reference CPU build not present (make -C oracle ref)')

GRAD_KEYS = ('shape_params', 'points', 'thickness', 'stroke_width', 'color_params', 'shape_to_canvas',
             'filter_radius')
SCALE = {'many_groups': scenes.many_groups, 'crowded_tile': scenes.crowded_tile, 'long_paths': scenes.long_paths}


@pytest.mark.parametrize('pre', [False, True], ids=['nopre', 'pre'])
@pytest.mark.parametrize('name', sorted(SCALE))
def test_oracle_matches_zeroinit_reference_at_scale(name, pre):
    sc = SCALE[name]()
    w, h = sc['canvas']
    f = flatten(w, h, sc['shapes'], sc['groups'], 0, 0.5)
    d = np.random.default_rng(2).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    o = OracleRunner(f, threads=8)
    img = o.render(w, h, 2, 2, 3, prefilter=pre)['image']
    g = o.render(w, h, 2, 2, 3, prefilter=pre, d_image=d)['grads']
    r = ref_render_in_child(f, 'zeroinit', w, h, 2, 2, 3, [dict(prefilter=pre), dict(prefilter=pre, d_image=d)])
    assert float(np.abs(r[0]['image'] - img).max()) <= 1e-5
    for k in GRAD_KEYS:
        if g[k].size:
            ref = r[1]['grads'][k]
            assert float(np.abs(ref - g[k]).max()) <= 1e-3 * max(float(np.abs(ref).max()), 1e-12), k
