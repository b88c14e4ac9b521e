"""The continuous-batching server on the GPU: every request's image / SDF and gradients
are the same whether it is served alone or in a mix of concurrent requests of other
sizes and kinds, and agree with the oracle."""
import numpy as np
import pytest
import torch

import diffvg_amd as pd
from diffvg_amd import _lib, serve, workloads
from diffvg_amd.render_pytorch import flatten_args
from refdrive import OracleRunner
from test_oracle_pins import GRAD_KEYS, _rel

pytestmark = pytest.mark.gpu
DEV = torch.device('cuda:0')


def _specs(n=14, seed=31):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        res = int(rng.choice([32, 48, 64, 96]))
        npaths = int(rng.integers(1, 40))
        shapes, groups = workloads.closed_paths(rng, npaths, float(res), 2, 6, res / 8.0, 'cpu', stroke_frac=0.3,
                                                grad_frac=0.1)
        args = pd.RenderFunction.serialize_scene(res, res, shapes, groups)
        flat = flatten_args(args).desc
        sdf = i % 4 == 3
        pre = i % 3 == 1 and not sdf
        nsx, nsy = (1, 1) if sdf else (int(rng.integers(1, 4)), int(rng.integers(1, 4)))
        target = torch.from_numpy(rng.uniform(0, 1, (res, res, 1 if sdf else 4)).astype(np.float32)).to(DEV)
        out.append(dict(flat=flat, res=res, sdf=sdf, pre=pre, nsx=nsx, nsy=nsy, target=target, seed=i))
    return out


def _requests(specs, scenes):
    return [serve.Request(i, scenes[i], s['res'], s['res'], s['nsx'], s['nsy'], s['seed'], s['sdf'], s['pre'],
                          s['target'], 0.0) for i, s in enumerate(specs)]


def test_requests_same_alone_and_in_the_mix():
    specs = _specs()
    scenes = [_lib.DeviceScene(s['flat'], 0) for s in specs]
    # in the mix: everything arrives at once, small budget and batch size so that several
    # batches are in flight on different streams
    mixed = serve.Server(device=0, hbm_budget=64 << 20, max_batch=4, streams=3)
    done = mixed.run(_requests(specs, scenes), [0.0] * len(specs))
    assert len(done) == len(specs)
    mix = {r.rid: dict(out=r.out.cpu().numpy(), grad=r.grad.cpu().numpy()) for r in done}
    for i, s in enumerate(specs):
        print('request', i, s['res'], 'sdf' if s['sdf'] else 'colour', 'pre' if s['pre'] else '', flush=True)
        alone_srv = serve.Server(device=0, hbm_budget=1 << 30, max_batch=1, streams=1)
        r = alone_srv.run(_requests([s], [scenes[i]]), [0.0])[0]
        a, m = r.out.cpu().numpy(), mix[i]
        assert np.array_equal(a, m['out']), i                      # same kernels: bit-identical forward
        assert _rel(m['grad'], r.grad.cpu().numpy()) <= 1e-4, i    # atomic summation order only


def test_served_results_match_oracle():
    specs = _specs(n=6, seed=32)
    scenes = [_lib.DeviceScene(s['flat'], 0) for s in specs]
    srv = serve.Server(device=0, hbm_budget=1 << 30, max_batch=8, streams=2)
    done = {r.rid: r for r in srv.run(_requests(specs, scenes), [0.0] * len(specs))}
    for i, s in enumerate(specs):
        r = done[i]
        o = OracleRunner(s['flat'])
        w = s['res']
        if s['sdf']:
            ref = o.render(w, w, 1, 1, s['seed'], output='sdf')['sdf']
        else:
            ref = o.render(w, w, s['nsx'], s['nsy'], s['seed'], prefilter=s['pre'])['image']
        out = r.out.cpu().numpy()
        assert np.abs(out - ref).max() <= 1e-5, i
        d = (2.0 * (out - s['target'].cpu().numpy())).astype(np.float32)
        kw = dict(output='sdf', d_sdf=d) if s['sdf'] else dict(prefilter=s['pre'], d_image=d)
        g = o.render(w, w, s['nsx'] if not s['sdf'] else 1, s['nsy'] if not s['sdf'] else 1, s['seed'], **kw)['grads']
        layout, total = r.scene.grad_layout()
        block = r.grad.cpu().numpy()
        for k in GRAD_KEYS:
            n = g[k].size
            if n:
                assert _rel(block[layout[k]:layout[k] + n], g[k]) <= 1e-3, (i, k)
