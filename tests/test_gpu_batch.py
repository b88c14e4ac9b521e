"""Batched rendering (dvg_batch_*): many scenes of different sizes, sample counts,
filters, outputs and options in one launch per stage, checked job by job against
single-scene renders of the same scenes (same device code, so forward images are
bit-identical and gradients agree up to atomic summation order) and against the oracle;
and replayed from a captured graph."""
import numpy as np
import pytest
import torch

from diffvg_amd import _lib
from diffvg_amd.batch import Batch
import scenes
from refdrive import OracleRunner
from test_oracle_pins import GRAD_KEYS, _rel

pytestmark = pytest.mark.gpu

DEV = torch.device('cuda:0')

# (case, filter override, options)
SPECS = [
    ('single_circle_64', None, dict(nsx=2, nsy=2, seed=1)),
    ('curve_fill', None, dict(nsx=3, nsy=2, seed=2, prefilter=True)),
    ('strokes', ('gaussian', 1.5), dict(nsx=2, nsy=2, seed=3)),
    ('overlap_eo', ('tent', 1.0), dict(nsx=1, nsy=1, seed=4, background=True)),
    ('gradients', None, dict(nsx=2, nsy=2, seed=5, size=(40, 24))),
    ('mixed_segments', None, dict(nsx=2, nsy=2, seed=6, prefilter=True, background=True)),
    ('thick_stroke', None, dict(nsx=4, nsy=4, seed=7)),
    ('transformed', None, dict(sdf=True)),
    ('curve_fill', None, dict(sdf=True, eval=200)),
    ('primitives', ('hann', 1.0), dict(nsx=2, nsy=3, seed=8, size=(70, 33))),
]


def _inputs(f, opt, k):
    w, h = opt.get('size', (f['canvas_width'], f['canvas_height']))
    rng = np.random.default_rng(100 + k)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(DEV)
    bg = t(rng.uniform(0, 1, (h, w, 4))) if opt.get('background') else None
    if opt.get('sdf'):
        ev = t(rng.uniform(0, min(w, h), (opt['eval'], 2))) if opt.get('eval') else None
        n = opt['eval'] if opt.get('eval') else None
        d = t(rng.uniform(-1, 1, (n, 1) if n else (h, w, 1)))
    else:
        ev = None
        d = t(rng.uniform(-1, 1, (h, w, 4)))
    return w, h, bg, ev, d


def _fwd_kw(f, opt, k, out):
    w, h, bg, ev, d = _inputs(f, opt, k)
    if opt.get('sdf'):
        kw = dict(width=w, height=h, nsx=1, nsy=1, seed=0, render_sdf=out, eval_positions=ev,
                  num_eval_positions=0 if ev is None else ev.shape[0])
    else:
        kw = dict(width=w, height=h, nsx=opt['nsx'], nsy=opt['nsy'], seed=opt['seed'], render_image=out,
                  background=bg, use_prefiltering=opt.get('prefilter', False))
    return kw, (bg, ev, d)


def _bwd_kw(f, opt, k, d_tr, d_bg):
    w, h, bg, ev, d = _inputs(f, opt, k)
    if opt.get('sdf'):
        kw = dict(width=w, height=h, nsx=1, nsy=1, seed=0, d_render_sdf=d, eval_positions=ev,
                  num_eval_positions=0 if ev is None else ev.shape[0])
    else:
        kw = dict(width=w, height=h, nsx=opt['nsx'], nsy=opt['nsy'], seed=opt['seed'], d_render_image=d,
                  background=bg, d_background=d_bg if bg is not None else None, d_translation=d_tr,
                  use_prefiltering=opt.get('prefilter', False))
    return kw, (bg, ev, d)


def _outputs(f, opt):
    w, h = opt.get('size', (f['canvas_width'], f['canvas_height']))
    if opt.get('sdf'):
        out = torch.zeros((opt['eval'], 1) if opt.get('eval') else (h, w, 1), device=DEV)
        return out, None, None
    return (torch.zeros(h, w, 4, device=DEV), torch.zeros(h, w, 2, device=DEV),
            torch.zeros(h, w, 4, device=DEV))


def _flats():
    return [scenes.flat_case(c, fl) for c, fl, _ in SPECS]


def _single(flats):
    res = []
    for k, (f, (_, _, opt)) in enumerate(zip(flats, SPECS)):
        sc = _lib.DeviceScene(f, 0)
        out, d_tr, d_bg = _outputs(f, opt)
        kw, keep = _fwd_kw(f, opt, k, out)
        sc.render(**kw)
        sc.zero_grads(None)
        kw, keep2 = _bwd_kw(f, opt, k, d_tr, d_bg)
        sc.render(**kw)
        torch.cuda.synchronize()
        res.append(dict(out=out.cpu().numpy(), grads=sc.grads(),
                        d_tr=None if d_tr is None else d_tr.cpu().numpy(),
                        d_bg=None if d_bg is None else d_bg.cpu().numpy()))
    return res


def _batched(flats, graph=False, replays=1, persistent=False):
    scs = [_lib.DeviceScene(f, 0) for f in flats]
    outs, fj, bj, keep = [], [], [], []
    for k, (f, (_, _, opt)) in enumerate(zip(flats, SPECS)):
        out, d_tr, d_bg = _outputs(f, opt)
        kw, kp = _fwd_kw(f, opt, k, out)
        fj.append((scs[k], kw))
        kw2, kp2 = _bwd_kw(f, opt, k, d_tr, d_bg)
        bj.append((scs[k], kw2))
        outs.append((out, d_tr, d_bg))
        keep += [kp, kp2]
    fb = Batch(fj, zero_grads=False, persistent=persistent)
    bb = Batch(bj, zero_grads=True, persistent=persistent)

    def run(stream):
        for out, d_tr, d_bg in outs:
            out.zero_()
            if d_tr is not None:
                d_tr.zero_()
                d_bg.zero_()
        fb.render(stream)
        bb.render(stream)

    results = []
    if not graph:
        run(None)
        fb.status()
        bb.status()
        replays = 1
    else:
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            run(s.cuda_stream)  # warm-up outside the capture
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            run(torch.cuda.current_stream().cuda_stream)
    for r in range(replays):
        if graph:
            junk = torch.randn(1 << 16, device=DEV).sum()  # eager work between replays
            g.replay()
            torch.cuda.synchronize()
            fb.status()
            bb.status()
            del junk
        results.append([dict(out=o.cpu().numpy(), grads=sc.grads(),
                             d_tr=None if t is None else t.cpu().numpy(),
                             d_bg=None if b is None else b.cpu().numpy())
                        for sc, (o, t, b) in zip(scs, outs)])
    return results


def _same(a, b, label, exact=True):
    # gradients: same arithmetic, float atomics summed in a different order (up to
    # ~1e-5 of the largest entry with 16 samples per pixel); the bar is 1e-3
    if exact:
        assert np.array_equal(a['out'], b['out']), label
    else:
        assert np.abs(a['out'] - b['out']).max() <= 1e-6, label
    for k in GRAD_KEYS:
        assert _rel(a['grads'][k], b['grads'][k]) <= 1e-4, (label, k)
    for k in ('d_tr', 'd_bg'):
        if b[k] is not None:
            assert _rel(a[k], b[k]) <= 1e-4, (label, k)


def test_batch_matches_single_scene_renders():
    flats = _flats()
    single = _single(flats)
    batched = _batched(flats)[0]
    for k, (a, b) in enumerate(zip(batched, single)):
        print('job', k, SPECS[k][0], SPECS[k][2], flush=True)
        _same(a, b, (k, SPECS[k][0]))


def test_batch_matches_oracle():
    flats = _flats()
    batched = _batched(flats)[0]
    for k, (f, (case, _, opt)) in enumerate(zip(flats, SPECS)):
        if opt.get('sdf') or opt.get('background'):
            continue
        w, h = opt.get('size', (f['canvas_width'], f['canvas_height']))
        o = OracleRunner(f)
        ref = o.render(w, h, opt['nsx'], opt['nsy'], opt['seed'], prefilter=opt.get('prefilter', False))['image']
        assert np.abs(batched[k]['out'] - ref).max() <= 1e-5, case
        d = _inputs(f, opt, k)[4].cpu().numpy()
        g = o.render(w, h, opt['nsx'], opt['nsy'], opt['seed'], prefilter=opt.get('prefilter', False), d_image=d,
                     translation=True)
        for key in GRAD_KEYS:
            assert _rel(batched[k]['grads'][key], g['grads'][key]) <= 1e-3, (case, key)


def test_batch_graph_replay():
    """Forward and backward batches captured in one graph, replayed three times with
    eager work in between; every replay equals the eager batch."""
    flats = _flats()
    eager = _batched(flats)[0]
    reps = _batched(flats, graph=True, replays=3)
    for r, res in enumerate(reps):
        for k, (a, b) in enumerate(zip(res, eager)):
            _same(a, b, ('replay', r, k), exact=True)


@pytest.mark.skip(reason='the persistent batch kernel did not finish on the GPU (cause not known); the path is disabled')
def test_persistent_batch_matches_single_scene_renders():
    """The same jobs as one persistent launch per batch (every stage of every job taken
    from a device queue, each task waiting for its own job's earlier stages)."""
    flats = _flats()
    single = _single(flats)
    batched = _batched(flats, persistent=True)[0]
    for k, (a, b) in enumerate(zip(batched, single)):
        print('persistent job', k, SPECS[k][0], SPECS[k][2], flush=True)
        _same(a, b, ('persistent', k, SPECS[k][0]))


@pytest.mark.skip(reason='the persistent batch kernel did not finish on the GPU (cause not known); the path is disabled')
def test_persistent_batch_graph_replay():
    """Persistent forward and backward launches captured in one graph and replayed three
    times (the last workgroup of each launch resets the task counters for the next):
    every replay equals the eager staged batch."""
    flats = _flats()
    eager = _batched(flats)[0]
    reps = _batched(flats, graph=True, replays=3, persistent=True)
    for r, res in enumerate(reps):
        for k, (a, b) in enumerate(zip(res, eager)):
            _same(a, b, ('persistent replay', r, k), exact=True)


@pytest.mark.parametrize('persistent', [False, pytest.param(True, marks=pytest.mark.skip(
    reason='the persistent batch kernel did not finish on the GPU (cause not known); the path is disabled'))])
def test_batch_job_with_forward_and_backward(persistent):
    """Jobs that render and differentiate in one call (render_image and d_render_image
    both set, as dvg_render allows): the backward replays the fragment lists its own
    forward keeps in the same launch, so it must wait for that forward."""
    cases = ['overlap_eo', 'strokes', 'gradients', 'mixed_segments']
    flats = [scenes.flat_case(c) for c in cases]
    jobs, outs, want = [], [], []
    for k, f in enumerate(flats):
        w, h = f['canvas_width'], f['canvas_height']
        d = torch.from_numpy(np.random.default_rng(30 + k).uniform(-1, 1, (h, w, 4)).astype(np.float32)).to(DEV)
        one = _lib.DeviceScene(f, 0)
        img1 = torch.zeros(h, w, 4, device=DEV)
        one.render(w, h, 2, 2, k, render_image=img1, d_render_image=d)
        torch.cuda.synchronize()
        want.append((img1.cpu().numpy(), one.grads()))
        sc = _lib.DeviceScene(f, 0)
        img = torch.zeros(h, w, 4, device=DEV)
        jobs.append((sc, dict(width=w, height=h, nsx=2, nsy=2, seed=k, render_image=img, d_render_image=d)))
        outs.append((sc, img, d))
    b = Batch(jobs, zero_grads=True, persistent=persistent)
    b.render(None)
    b.status()
    for k, ((sc, img, _), (w_img, w_g)) in enumerate(zip(outs, want)):
        assert np.array_equal(img.cpu().numpy(), w_img), k
        g = sc.grads()
        for key in GRAD_KEYS:
            assert _rel(g[key], w_g[key]) <= 1e-4, (k, key)


def test_batch_rejects_invalid_jobs():
    f = scenes.flat_case('single_circle_64')
    sc = _lib.DeviceScene(f, 0)
    out = torch.zeros(8, 8, 4, device=DEV)
    with pytest.raises(RuntimeError, match='PERSISTENT is disabled'):
        Batch([(sc, dict(width=8, height=8, nsx=1, nsy=1, seed=0, render_image=out))], persistent=True)
    with pytest.raises(RuntimeError, match='256 samples'):
        Batch([(sc, dict(width=8, height=8, nsx=17, nsy=16, seed=0, render_image=out))])
    with pytest.raises(RuntimeError, match='either colour or SDF'):
        Batch([(sc, dict(width=8, height=8, nsx=1, nsy=1, seed=0))])


def test_batched_prefiltered_circle_stroke_is_reported():
    """The reference's prefiltered forward asserts on a stroked circle (diffvg.cpp:868).
    A batch flags it in the scene's condition word and the batch's status raises it: for
    a forward-only batch, for a forward batch whose backward batch has run since (the
    backward replays kept lists and would not flag it again), and through render_batch
    at the next synchronize().  The status read clears the condition."""
    import diffvg_amd as pd
    from diffvg_amd import render_pytorch as rp
    from diffvg_amd.batch_render import render_batch
    c = pd.Circle(radius=scenes.T(10.0), center=scenes.T([32.0, 32.0]))
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=scenes.T([0.2, 0.4, 0.6, 1.0]),
                      stroke_color=scenes.T([0.0, 0.0, 0.0, 1.0]))
    f = scenes.flat_case(dict(canvas=(64, 64), shapes=[c], groups=[g], filter=('box', 0.5)))
    sc = _lib.DeviceScene(f, 0)
    img = torch.zeros(64, 64, 4, device=DEV)
    fwd = Batch([(sc, dict(width=64, height=64, nsx=2, nsy=2, seed=0, render_image=img, use_prefiltering=True))],
                zero_grads=False)
    fwd.render(None)
    with pytest.raises(RuntimeError, match='prefiltered stroke'):
        fwd.status()
    fwd.status()  # cleared by the read
    d = torch.ones(64, 64, 4, device=DEV)
    bwd = Batch([(sc, dict(width=64, height=64, nsx=2, nsy=2, seed=0, d_render_image=d, use_prefiltering=True))])
    fwd.render(None)
    bwd.render(None)
    with pytest.raises(RuntimeError, match='prefiltered stroke'):
        fwd.status()
    bwd.status()
    pd.set_use_gpu(True)
    pd.set_device(DEV)
    rp.synchronize()
    c.radius.requires_grad_(True)
    args = pd.RenderFunction.serialize_scene(64, 64, [c], [g], use_prefiltering=True)
    out = render_batch([args])
    out[0].sum().backward()
    with pytest.raises(RuntimeError, match='prefiltered stroke'):
        pd.synchronize()
    pd.synchronize()
    out = render_batch([args])  # forward only
    with pytest.raises(RuntimeError, match='prefiltered stroke'):
        pd.synchronize()
    pd.synchronize()


def _uneven_subset(items):
    """About ten scenes of the 256 spanning the workload: every resolution, 1 to ~500
    paths, and gradient fills (linear and radial)."""
    from diffvg_amd.color import LinearGradient, RadialGradient
    info = []
    for k, it in enumerate(items):
        lin = sum(isinstance(g.fill_color, LinearGradient) for g in it['groups'])
        rad = sum(isinstance(g.fill_color, RadialGradient) for g in it['groups'])
        info.append((k, it['canvas'][0], len(it['shapes']), lin, rad))
    pick = []
    for res in (28, 64, 128, 256):
        at = [x for x in info if x[1] == res]
        pick.append(min(at, key=lambda x: (x[2], x[0]))[0])              # fewest paths
        pick.append(max(at, key=lambda x: (x[2], -x[0]))[0])             # most paths
    pick.append(max(info, key=lambda x: (x[3], -x[0]))[0])               # most linear gradients
    pick.append(max(info, key=lambda x: (x[4], -x[0]))[0])               # most radial gradients
    both = [x for x in info if x[3] and x[4]]
    pick.append(min(both, key=lambda x: (x[2], x[0]))[0])                # both kinds, fewest paths
    mid = [x for x in info if x[1] == 256 and 40 <= x[2] <= 120]
    pick.append(min(mid, key=lambda x: x[0])[0])                         # a mid-sized 256 px scene
    return sorted(set(pick))


def test_uneven_scenes_256_graph():
    """The batched workload at full size: 256 scenes of 1-512 closed cubic paths (2-8
    segments, 10% linear or radial gradient fills) at 28-256 px, 2x2 samples, one forward
    batch + L2 loss gradient + backward batch captured in one HIP graph and replayed.
    Every scene's image equals its single-scene render bitwise and its gradients agree
    up to atomic summation order; a subset spanning resolutions, path counts and
    gradient fills matches the oracle (forward 1e-5, gradients 1e-3).  The graph is
    replayed three times with eager work between the replays, each replay checked."""
    from diffvg_amd import workloads
    import diffvg_amd as pd
    from diffvg_amd.render_pytorch import flatten_args
    items = workloads.uneven_scenes(device='cpu')
    n = len(items)
    sizes = [it['canvas'] for it in items]
    flats, scs = [], []
    for it in items:
        w, h = it['canvas']
        f = flatten_args(pd.RenderFunction.serialize_scene(w, h, it['shapes'], it['groups'], it['filter'])).desc
        flats.append(f)
        scs.append(_lib.DeviceScene(f, 0))
    npix = [w * h for w, h in sizes]
    offs = np.concatenate([[0], np.cumsum(npix)]).astype(np.int64)
    rng = np.random.default_rng(4)
    target = torch.from_numpy(rng.uniform(0, 1, (int(offs[-1]), 4)).astype(np.float32)).to(DEV)
    img = torch.zeros(int(offs[-1]), 4, device=DEV)
    d = torch.zeros(int(offs[-1]), 4, device=DEV)
    view = lambda t, i: t[offs[i]:offs[i + 1]].view(sizes[i][1], sizes[i][0], 4)
    fb = Batch([(scs[i], dict(width=sizes[i][0], height=sizes[i][1], nsx=2, nsy=2, seed=0,
                              render_image=view(img, i))) for i in range(n)], zero_grads=False)
    bb = Batch([(scs[i], dict(width=sizes[i][0], height=sizes[i][1], nsx=2, nsy=2, seed=0,
                              d_render_image=view(d, i))) for i in range(n)], zero_grads=True)
    stream = torch.cuda.Stream()

    def body(st):
        img.zero_()
        fb.render(st)
        torch.mul(img - target, 2.0, out=d)
        bb.render(st)

    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        body(stream.cuda_stream)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        body(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    replays = []
    for r in range(3):
        print('uneven_scenes_256 replay', r, flush=True)
        g.replay()
        torch.cuda.synchronize()
        fb.status()
        bb.status()
        replays.append((img.cpu().numpy(), d.cpu().numpy(), [sc.grads() for sc in scs]))
        junk = torch.rand(1 << 20, device=DEV)  # eager work between replays
        (junk * 2).sum().item()
    img0, d0, g0 = replays[0]
    for img_r, d_r, g_r in replays[1:]:
        assert np.array_equal(img_r, img0)
        for i in range(n):
            for k in GRAD_KEYS:
                if g0[i][k].size:
                    assert _rel(g_r[i][k], g0[i][k]) <= 1e-4, (i, k)
    # every scene against its single-scene render
    for i in range(n):
        w, h = sizes[i]
        one = torch.zeros(h, w, 4, device=DEV)
        scs[i].render(w, h, 2, 2, 0, render_image=one)
        scs[i].zero_grads(None)
        scs[i].render(w, h, 2, 2, 0, d_render_image=torch.from_numpy(d0[offs[i]:offs[i + 1]].reshape(h, w, 4)).to(DEV))
        torch.cuda.synchronize()
        assert np.array_equal(one.cpu().numpy().reshape(-1, 4), img0[offs[i]:offs[i + 1]]), i
        gs = scs[i].grads()
        for k in GRAD_KEYS:
            if gs[k].size:
                assert _rel(g0[i][k], gs[k]) <= 1e-4, (i, k)
    # a subset against the oracle
    for i in _uneven_subset(items):
        w, h = sizes[i]
        print('uneven scene', i, 'vs oracle:', w, 'px', flats[i]['num_shapes'], 'paths', flush=True)
        orc = OracleRunner(flats[i], threads=8)
        want = orc.render(w, h, 2, 2, 0)['image']
        assert np.abs(img0[offs[i]:offs[i + 1]].reshape(h, w, 4) - want).max() <= 1e-5, i
        gw = orc.render(w, h, 2, 2, 0, d_image=d0[offs[i]:offs[i + 1]].reshape(h, w, 4))['grads']
        for k in GRAD_KEYS:
            if gw[k].size:
                assert _rel(g0[i][k], gw[k]) <= 1e-3, (i, k)
