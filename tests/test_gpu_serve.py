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


# the serving stream's request mix, with pinned requests: two 1024^2 colour renders
# (prefiltering off and on), SDFs of 2,500 and 2,100 paths, and prefiltered colour renders
STREAM_PINS = {0: dict(res=1024, sdf=False, prefilter=False),
               3: dict(res=1024, n_paths=60, sdf=False, prefilter=True),
               5: dict(res=128, n_paths=2500, sdf=True),
               8: dict(res=256, n_paths=2100, sdf=True),
               10: dict(res=256, sdf=False, prefilter=True),
               12: dict(res=128, sdf=False, prefilter=False)}


def _stream_specs(n=16, seed=51):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        r = workloads.serving_request(rng, device='cpu', **STREAM_PINS.get(i, {}))
        w, h = r['canvas']
        args = pd.RenderFunction.serialize_scene(w, h, r['shapes'], r['groups'], r['filter'])
        target = torch.from_numpy(rng.uniform(0, 1, (h, w, 1 if r['sdf'] else 4)).astype(np.float32)).to(DEV)
        out.append(dict(flat=flatten_args(args).desc, res=w, sdf=r['sdf'], pre=r['prefilter'] and not r['sdf'],
                        nsx=r['nsx'], nsy=r['nsy'], target=target, seed=i, paths=len(r['shapes'])))
    return out


def test_serving_stream_workload():
    """The serving workload's requests (1-4,096 paths, 128-1,024 px, colour or SDF,
    prefiltering on or off; pinned: 1024^2 colour with and without prefiltering, SDFs of
    more than 2,000 paths) served in a mix arriving at once and at Poisson times: every
    request's output is bit-identical to the request served alone and its gradients agree
    up to atomic summation order; the requests of at most 256^2 match the oracle."""
    specs = _stream_specs()
    kinds = [(s['res'], s['paths'], s['sdf'], s['pre']) for s in specs]
    assert any(r == 1024 and not sdf and not pre for r, _, sdf, pre in kinds)
    assert any(r == 1024 and pre for r, _, sdf, pre in kinds)
    assert sum(1 for _, p, sdf, _ in kinds if sdf and p >= 2000) >= 2
    scenes = [_lib.DeviceScene(s['flat'], 0) for s in specs]
    alone = {}
    for i, s in enumerate(specs):
        print('stream request', i, s['res'], s['paths'], 'paths', 'sdf' if s['sdf'] else 'colour',
              'pre' if s['pre'] else '', flush=True)
        r = serve.Server(device=0, hbm_budget=8 << 30, max_batch=1, streams=1).run(
            _requests([s], [scenes[i]]), [0.0])[0]
        assert r.error is None, r.error
        alone[i] = (r.out.cpu().numpy(), r.grad.cpu().numpy())
    rng = np.random.default_rng(52)
    for label, arr in (('at once', [0.0] * len(specs)),
                       ('poisson', list(np.cumsum(rng.exponential(0.004, len(specs))) - 0.0))):
        print('mixed stream', label, flush=True)
        srv = serve.Server(device=0, hbm_budget=2 << 30, max_batch=6, streams=3)
        done = srv.run(_requests(specs, scenes), sorted(arr))
        assert len(done) == len(specs)
        for r in done:
            assert r.error is None, r.error
            a_out, a_grad = alone[r.rid]
            assert np.array_equal(r.out.cpu().numpy(), a_out), (label, r.rid)
            assert _rel(r.grad.cpu().numpy(), a_grad) <= 1e-4, (label, r.rid)
    for i, s in enumerate(specs):
        if s['res'] > 256:
            continue
        print('stream request', i, 'vs oracle', flush=True)
        o = OracleRunner(s['flat'], threads=8)
        w = s['res']
        out, block = alone[i]
        if s['sdf']:
            ref = o.render(w, w, 1, 1, s['seed'], output='sdf')['sdf']
        else:
            ref = o.render(w, w, s['nsx'], s['nsy'], s['seed'], prefilter=s['pre'])['image']
        assert np.abs(out - ref).max() <= 1e-5, i
        d = (2.0 * (out - s['target'].cpu().numpy())).astype(np.float32)
        kw = dict(output='sdf', d_sdf=d) if s['sdf'] else dict(prefilter=s['pre'], d_image=d)
        g = o.render(w, w, 1 if s['sdf'] else s['nsx'], 1 if s['sdf'] else s['nsy'], s['seed'], **kw)['grads']
        layout, _ = scenes[i].grad_layout()
        for k in GRAD_KEYS:
            n = g[k].size
            if n:
                assert _rel(block[layout[k]:layout[k] + n], g[k]) <= 1e-3, (i, k)
