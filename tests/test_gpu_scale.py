"""The HIP render path against the oracle on scenes of unusual scale (the oracle itself
is checked against the reference's CPU build on the same scenes in test_ref_scale.py):
12,000 groups over a 256x256 canvas; a 16x16 canvas whose tiles each overlap 3,138 groups
(more candidates than one pass of the tiled kernels' LDS list holds, so the candidate
loop runs in several passes); a closed path of 1,500 cubic segments and an open stroke of
600.  Forward and backward, prefiltering off and on, at the parity bars (images within
1e-5, gradients within 1e-3 relative), no exemptions."""
import numpy as np
import pytest

import scenes
from gpudrive import GpuRunner
from refdrive import OracleRunner, flatten
from test_gpu_parity import MODES, _compare

pytestmark = pytest.mark.gpu

SCALE = {'many_groups': scenes.many_groups, 'crowded_tile': scenes.crowded_tile, 'long_paths': scenes.long_paths}


@pytest.mark.parametrize('mode', sorted(MODES))
@pytest.mark.parametrize('name', sorted(SCALE))
def test_gpu_matches_oracle_at_scale(name, mode):
    print('scale', name, mode, flush=True)
    sc = SCALE[name]()
    w, h = sc['canvas']
    f = flatten(w, h, sc['shapes'], sc['groups'], 0, 0.5)
    o, g = OracleRunner(f, threads=8), GpuRunner(f, MODES[mode])
    d = np.random.default_rng(2).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    tol = 1e-5 if mode == 'tiled' else 5e-5
    for pre in (False, True):
        _compare(g.render(w, h, 2, 2, 3, prefilter=pre), o.render(w, h, 2, 2, 3, prefilter=pre), (name, pre, 'fwd'),
                 tol)
        _compare(g.render(w, h, 2, 2, 3, prefilter=pre, d_image=d), o.render(w, h, 2, 2, 3, prefilter=pre, d_image=d),
                 (name, pre, 'bwd'))
