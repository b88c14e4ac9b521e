"""GPU parity of the HIP render path (libdiffvg_amd.so, called through its C ABI).

Bars (from the render path's contract): forward colour and SDF within 1e-5 max abs of
the reference; gradients within 1e-3 relative (max |a - b| / max |b| per gradient array)
of the reference at equal seeds.  Checked against

* the reference's own outputs (tests/golden/*.npz, written by its CPU build);
* the CPU oracle on the scene zoo, on filters and options the fixtures do not hold
  (Gaussian filter, backgrounds, odd canvas and sample counts, both schedules);
* the oracle on the full 512x512 workloads (300 closed paths; 2,048 open strokes).
"""
import math

import numpy as np
import pytest
import torch

import diffvg_amd as pd
from diffvg_amd import _lib, workloads
import scenes
from gpudrive import GpuRunner, ShimRunner
from refdrive import OracleRunner, flatten
from test_oracle_pins import GRAD_KEYS, _rel, check_against_fixture, fixture_names, load_fixture

pytestmark = pytest.mark.gpu

MODES = {'tiled': _lib.MODE_TILED, 'plain': _lib.MODE_PLAIN}
# The plain schedule (the comparator built like the reference's GPU path) splats every
# sample into its filter footprint with float atomics in arrival order; with a wide
# filter (a 5x5 Gaussian footprint at 2x2 samples sums ~100 terms per pixel) the order
# changes the rounding of the sum by up to ~n * 2^-24 of the value.  The tiled schedule
# gathers in a fixed order and is held to 1e-5.
IMAGE_TOL = {'tiled': 1e-5, 'plain': 5e-5}


def _runner_fn(g):
    def run(kw):
        kw = dict(kw)
        w, h, nsx, nsy, seed = (kw.pop(k) for k in ('width', 'height', 'nsx', 'nsy', 'seed'))
        return g.render(w, h, nsx, nsy, seed, **kw)
    return run


@pytest.mark.parametrize('mode', sorted(MODES))
@pytest.mark.parametrize('name', fixture_names())
def test_gpu_matches_reference_fixture(name, mode):
    print('fixture', name, mode, flush=True)
    flat, z = load_fixture(name)
    check_against_fixture(_runner_fn(GpuRunner(flat, MODES[mode])), flat, z, name + '/' + mode)


def _compare(a, b, label, image_tol=1e-5):
    for k in ('image', 'sdf'):
        if b.get(k) is not None:
            assert np.abs(a[k] - b[k]).max() <= image_tol, (label, k)
    for k in ('d_translation', 'd_background'):
        if b.get(k) is not None:
            assert _rel(a[k], b[k]) <= 1e-3, (label, k)
    if 'grads' in b:
        for k in GRAD_KEYS:
            assert _rel(a['grads'][k], b['grads'][k]) <= 1e-3, (label, k)


ZOO = [(c, f) for c in sorted(scenes.CASES) for f in [None, ('gaussian', 1.5)]
       if c != 'single_circle_256' or f is None]


@pytest.mark.parametrize('mode', sorted(MODES))
@pytest.mark.parametrize('case, filt', ZOO)
def test_gpu_matches_oracle_zoo(case, filt, mode):
    print('zoo', case, filt, mode, flush=True)
    f = scenes.flat_case(case, filt)
    w, h = f['canvas_width'], f['canvas_height']
    o, g = OracleRunner(f), GpuRunner(f, MODES[mode])
    bg = np.random.default_rng(1).uniform(0, 1, (h, w, 4)).astype(np.float32)
    d = np.random.default_rng(2).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    pres = [False] if case in ('primitives', 'single_circle_64', 'single_circle_256') else [False, True]
    for pre in pres:
        kw = dict(prefilter=pre)
        tol = IMAGE_TOL[mode]
        _compare(g.render(w, h, 2, 2, 5, **kw), o.render(w, h, 2, 2, 5, **kw), (case, pre, 'fwd'), tol)
        kw.update(background=bg)
        _compare(g.render(w, h, 2, 2, 5, **kw), o.render(w, h, 2, 2, 5, **kw), (case, pre, 'fwd bg'), tol)
        kw.update(d_image=d, translation=True, d_background=True)
        _compare(g.render(w, h, 2, 2, 5, **kw), o.render(w, h, 2, 2, 5, **kw), (case, pre, 'bwd bg'), tol)


@pytest.mark.parametrize('mode', sorted(MODES))
@pytest.mark.parametrize('w, h, nsx, nsy', [(37, 23, 1, 1), (37, 23, 3, 2), (5, 70, 2, 3), (1, 1, 4, 4),
                                            (33, 31, 17, 16)])
def test_gpu_odd_canvases_and_sample_counts(w, h, nsx, nsy, mode):
    """Partial tiles, non-square canvases, odd strata, one pixel, and more than 256
    samples per pixel (which the tiled schedule hands to the per-sample kernels)."""
    print('odd', w, h, nsx, nsy, mode, flush=True)
    sc = scenes.overlap(size=64, n=8)
    f = flatten(w, h, sc['shapes'], sc['groups'], 1, 1.0)
    o, g = OracleRunner(f), GpuRunner(f, MODES[mode])
    # beyond 256 samples per pixel the tiled schedule uses the per-sample kernels
    tol = IMAGE_TOL['plain'] if nsx * nsy > 256 else IMAGE_TOL[mode]
    _compare(g.render(w, h, nsx, nsy, 7), o.render(w, h, nsx, nsy, 7), 'fwd', tol)
    d = np.random.default_rng(3).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    _compare(g.render(w, h, nsx, nsy, 7, d_image=d), o.render(w, h, nsx, nsy, 7, d_image=d), 'bwd')


def test_gpu_render_size_differs_from_canvas():
    """The render resolution scales the canvas (diffvg.cpp:1188-1192 normalises by the
    render size and rescales by the canvas size)."""
    sc = scenes.overlap(size=64, n=6)
    f = flatten(64, 64, sc['shapes'], sc['groups'], 0, 0.5)
    o, g = OracleRunner(f), GpuRunner(f)
    _compare(g.render(96, 40, 2, 2, 1), o.render(96, 40, 2, 2, 1), 'fwd')
    d = np.random.default_rng(4).uniform(-1, 1, (40, 96, 4)).astype(np.float32)
    _compare(g.render(96, 40, 2, 2, 1, d_image=d), o.render(96, 40, 2, 2, 1, d_image=d), 'bwd')


def test_gpu_empty_scene():
    f = flatten(16, 16, [], [], 0, 0.5)
    g = GpuRunner(f)
    assert np.all(g.render(16, 16, 2, 2, 0)['image'] == 0)
    bg = np.random.default_rng(0).uniform(0, 1, (16, 16, 4)).astype(np.float32)
    assert np.array_equal(g.render(16, 16, 2, 2, 0, background=bg)['image'], bg)


@pytest.mark.parametrize('cfg', [2, 3])
def test_gpu_full_workloads_vs_oracle(cfg):
    """The full 512x512 benchmark scenes (300 closed paths with the tent twin of the
    Gaussian filter; 2,048 strokes), prefiltering off and on: forward images and the
    backward (edge sampling, or the prefiltered adjoint) against the oracle at the same
    seed.  Every pixel is held to 1e-5 and every gradient entry to 1e-3 of its array's
    largest magnitude -- no pixel or shape is exempt.  (A sample whose ray passes through
    a cubic segment joint is decided by the last bit of the double acos/cos/pow of the
    cubic solve; the kernels re-solve such roots with correctly rounded functions, so
    they agree with the oracle there too.)"""
    print('full workload', cfg, flush=True)
    c = workloads.closed_paths_300(device='cpu', filter_type='tent') if cfg == 2 else workloads.strokes_2048(device='cpu')
    w, h = c['canvas']
    f = flatten(w, h, c['shapes'], c['groups'], c['filter'].type, float(c['filter'].radius))
    o, g = OracleRunner(f), GpuRunner(f)
    owners = _grad_owners(f)
    d = np.random.default_rng(0).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    for pre in (False, True):
        a, b = g.render(w, h, 2, 2, 0, prefilter=pre)['image'], o.render(w, h, 2, 2, 0, prefilter=pre)['image']
        err = np.abs(a - b).max(-1)
        bad = [(int(x), int(y)) for y, x in zip(*np.nonzero(err > 1e-5))]
        assert not bad, ('forward', pre, bad[:8], float(err.max()))
        print('  prefilter', pre, 'forward max abs err %.2e' % float(err.max()), flush=True)
        ga = g.render(w, h, 2, 2, 0, d_image=d, prefilter=pre)['grads']
        gb = o.render(w, h, 2, 2, 0, d_image=d, prefilter=pre)['grads']
        for k in GRAD_KEYS:
            x, y = np.asarray(ga[k], np.float64), np.asarray(gb[k], np.float64)
            if y.size == 0:
                continue
            e = np.abs(x - y) / max(np.abs(y).max(), 1e-8)
            over = sorted(set(owners[k][e > 1e-3].tolist()))
            assert not over, ('backward', pre, k, float(e.max()), over[:8])


def _grad_owners(f):
    """Which shape (or group) each entry of each gradient array belongs to."""
    ns, ng = f['num_shapes'], f['num_groups']
    pts = np.zeros(f['num_points_total'], np.int64)
    par = np.zeros(f['num_shape_params'], np.int64)
    for s in range(ns):
        if f['shape_type'][s] == 2:
            o, n = f['path_point_offset'][s], f['path_num_points'][s]
            pts[o:o + n] = s
        else:
            o = f['shape_param_offset'][s]
            par[o:o + (3 if f['shape_type'][s] == 0 else 4)] = s
    col = np.zeros(f['num_color_params'], np.int64)
    for g in range(ng):
        for kind in ('fill', 'stroke'):
            t = f[kind + '_type'][g]
            if t >= 0:
                o = f[kind + '_offset'][g]
                col[o:o + (4 if t == 0 else 4 + 5 * f[kind + '_num_stops'][g])] = g
    return dict(shape_params=par, points=np.repeat(pts, 2), thickness=pts, stroke_width=np.arange(ns),
                color_params=col, shape_to_canvas=np.repeat(np.arange(ng), 9), filter_radius=np.zeros(1, np.int64))


def test_gpu_schedules_agree_bitwise_and_repeat():
    """Both schedules take every coverage decision on identical bits, so their forward
    images are identical, and a render repeated is identical (no atomics in the forward
    colour of the tiled schedule)."""
    f = scenes.flat_case('strokes')
    w, h = f['canvas_width'], f['canvas_height']
    t, p = GpuRunner(f, _lib.MODE_TILED), GpuRunner(f, _lib.MODE_PLAIN)
    a = t.render(w, h, 2, 2, 9)['image']
    assert np.array_equal(a, t.render(w, h, 2, 2, 9)['image'])
    assert np.abs(a - p.render(w, h, 2, 2, 9)['image']).max() <= 1e-6


# ---------------------------------------------------------------- through torch.autograd

def _params(shapes, groups):
    ps = []
    for s in shapes:
        for n in ('points', 'radius', 'center', 'p_min', 'p_max', 'stroke_width'):
            t = getattr(s, n, None)
            if isinstance(t, torch.Tensor) and t.is_floating_point():
                t.requires_grad_(True)
                ps.append(t)
    for g in groups:
        for col in (g.fill_color, g.stroke_color):
            if isinstance(col, torch.Tensor):
                col.requires_grad_(True)
                ps.append(col)
            elif col is not None:
                for n in ('begin', 'end', 'center', 'radius', 'offsets', 'stop_colors'):
                    t = getattr(col, n, None)
                    if isinstance(t, torch.Tensor):
                        t.requires_grad_(True)
                        ps.append(t)
        # a fresh leaf per group: pydiffvg's default transform is one tensor shared by
        # every ShapeGroup, which must not be turned into a leaf for the other tests
        g.shape_to_canvas = g.shape_to_canvas.detach().clone().requires_grad_(True)
        ps.append(g.shape_to_canvas)
    return ps


@pytest.mark.parametrize('case', ['primitives', 'gradients', 'transformed', 'thick_stroke', 'overlap_eo'])
def test_render_function_autograd_vs_oracle(case):
    """RenderFunction.apply + backward on cuda:0: the forward image and every parameter's
    .grad against the oracle's gradient arrays."""
    print('autograd', case, flush=True)
    pd.set_use_gpu(True)
    pd.set_device(torch.device('cuda:0'))
    sc = scenes.CASES[case]()
    w, h = sc['canvas']
    f = scenes.flat_case(case)
    ps = _params(sc['shapes'], sc['groups'])
    filt = pd.PixelFilter(type=pd.diffvg.FilterType.box, radius=torch.tensor(0.5, requires_grad=True))
    args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'], filt)
    img = pd.RenderFunction.apply(w, h, 2, 2, 4, None, *args)
    W = torch.from_numpy(np.random.default_rng(6).uniform(-1, 1, (h, w, 4)).astype(np.float32))
    (img.cpu() * W).sum().backward()
    o = OracleRunner(f)
    ref_img = o.render(w, h, 2, 2, 4)['image']
    assert np.abs(img.detach().cpu().numpy() - ref_img).max() <= 1e-5
    gref = o.render(w, h, 2, 2, 4, d_image=W.numpy())['grads']
    # expected .grad per tensor: the oracle's slices for every argument that is that
    # tensor, summed (ShapeGroup's default shape_to_canvas is one tensor shared by all
    # groups, as in pydiffvg, so autograd accumulates every group's share into it)
    from diffvg_amd.render_pytorch import flatten_args
    fl = flatten_args(args)
    expect, tensors = {}, {}
    for (ai, kind, off, cnt, shape) in fl.grad_map:
        t = args[ai]
        if not (isinstance(t, torch.Tensor) and t.requires_grad) or kind == 'filter_radius':
            continue
        e = gref[kind][off:off + cnt].astype(np.float64)
        expect[id(t)] = expect.get(id(t), 0) + e
        tensors[id(t)] = t
    scale = {k: max(np.abs(v).max(), 1e-8) for k, v in gref.items() if v.size}
    for key, t in tensors.items():
        assert t.grad is not None
        kind = [k for (ai, k, o, c, sh) in fl.grad_map if args[ai] is t][0]
        err = np.abs(t.grad.detach().cpu().numpy().reshape(-1).astype(np.float64) - expect[key]).max()
        assert err <= 1e-3 * scale[kind] * (1 + sum(1 for (ai, *_r) in fl.grad_map if args[ai] is t)), kind
    assert filt.radius.grad is not None
    assert _rel(filt.radius.grad.reshape(1).numpy(), gref['filter_radius']) <= 1e-3


def test_render_function_sdf_and_eval_positions():
    pd.set_use_gpu(True)
    sc = scenes.curve_fill()
    w, h = sc['canvas']
    f = scenes.flat_case('curve_fill')
    o = OracleRunner(f)
    args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'], output_type=pd.OutputType.sdf)
    sdf = pd.RenderFunction.apply(w, h, 1, 1, 0, None, *args)
    assert np.abs(sdf.detach().cpu().numpy() - o.render(w, h, 1, 1, 0, output='sdf')['sdf']).max() <= 1e-5
    ev = np.random.default_rng(8).uniform(0, 64, (300, 2)).astype(np.float32)
    args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'], output_type=pd.OutputType.sdf,
                                             eval_positions=torch.from_numpy(ev))
    sdf = pd.RenderFunction.apply(w, h, 1, 1, 0, None, *args)
    ref = o.render(w, h, 1, 1, 0, output='sdf', eval_positions=ev)['sdf']
    assert np.abs(sdf.detach().cpu().numpy().reshape(-1) - ref.reshape(-1)).max() <= 1e-5


def test_render_grad_translation_vs_oracle():
    """render_grad (render_pytorch.py:639-717): the d_translation image of a backward."""
    pd.set_use_gpu(True)
    sc = scenes.mixed_segments()
    w, h = sc['canvas']
    f = scenes.flat_case('mixed_segments')
    args = pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'])
    d = np.random.default_rng(9).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    out = pd.RenderFunction.render_grad(torch.from_numpy(d).cuda(), w, h, 2, 2, 0, None, *args)
    ref = OracleRunner(f).render(w, h, 2, 2, 0, d_image=d, translation=True)['d_translation']
    assert _rel(out.detach().cpu().numpy(), ref) <= 1e-3


def test_gpu_prefiltered_gradient_vs_finite_differences():
    """Central differences (the finite_difference_comp.py pattern) through RenderFunction
    on a prefiltered polyline stroke, where the image is smooth in the points."""
    pd.set_use_gpu(True)
    pts0 = torch.tensor([[10.0, 10.0], [50.0, 14.0], [44.0, 50.0]])
    W = torch.from_numpy(np.random.default_rng(12).uniform(-1, 1, (64, 64, 4)).astype(np.float32)).cuda()

    def loss(pts, grad=False):
        p = pts.clone().requires_grad_(grad)
        s = pd.Polygon(points=p, is_closed=False, stroke_width=torch.tensor(2.0))
        g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=None, stroke_color=torch.tensor([0.8, 0.2, 0.1, 1.0]))
        args = pd.RenderFunction.serialize_scene(64, 64, [s], [g], use_prefiltering=True)
        L = (pd.RenderFunction.apply(64, 64, 1, 1, 0, None, *args) * W).sum()
        if grad:
            L.backward()
            return p.grad.numpy().reshape(-1)
        return float(L)

    an = loss(pts0, True)
    fd = []
    for k in range(6):
        a, b = pts0.clone().reshape(-1), pts0.clone().reshape(-1)
        a[k] += 0.02
        b[k] -= 0.02
        fd.append((loss(a.reshape(3, 2)) - loss(b.reshape(3, 2))) / 0.04)
    fd = np.array(fd)
    assert np.linalg.norm(an - fd) / np.linalg.norm(fd) < 3e-3


def test_gpu_edge_sampling_unbiased():
    """Seed-averaged edge-sampled d(loss)/d(centre) of a circle against the exact
    derivative of the coverage loss (the non-prefiltered gradient is stochastic)."""
    W = np.zeros((48, 48, 4), np.float32)
    W[..., 3] = 1.0
    W[:, 24:, 3] = 2.0
    c = pd.Circle(radius=torch.tensor(10.0), center=torch.tensor([21.0, 24.0]))
    g = pd.ShapeGroup(shape_ids=torch.tensor([0]), fill_color=torch.tensor([1.0, 1, 1, 1]))
    r = GpuRunner(flatten(48, 48, [c], [g], 0, 0.5))
    gs = [r.render(48, 48, 4, 4, s, d_image=W)['grads']['shape_params'][1] for s in range(64)]
    assert np.mean(gs) == pytest.approx(2 * math.sqrt(100 - 9), rel=0.03)


@pytest.mark.parametrize('name', ['gradients', 'thick_stroke', 'transformed', 'curve_fill', 'primitives'])
def test_drop_in_diffvg_module_matches_reference_fixture(name):
    """The `diffvg` module shim (Scene / render / get_d_shape / get_d_shape_group /
    get_d_filter_radius), driven the way render_pytorch.py drives the reference's pybind11
    module, against the reference's own outputs."""
    print('shim', name, flush=True)
    flat, z = load_fixture(name)
    check_against_fixture(_runner_fn(ShimRunner(flat)), flat, z, 'shim/' + name)


@pytest.mark.parametrize('case', ['sdf_many', 'primitives', 'transformed', 'quad_approx'])
def test_gpu_sdf_tile_culling_vs_oracle(case):
    """The tiled SDF culls groups per tile by distance bounds; the oracle scans every group
    (sample_distance's linear loop).  Results and gradients must agree, including scenes
    with rects, circles, a non-uniform transform and approximate quadratic distances."""
    print('sdf', case, flush=True)
    if case == 'sdf_many':
        import diffvg_amd as pd_
        rng = np.random.default_rng(21)
        shapes, groups = workloads.closed_paths(rng, 300, 128.0, 2, 8, 8.0, 'cpu', stroke_frac=0.3)
        f = flatten(128, 128, shapes, groups, 0, 0.5)
    else:
        f = scenes.flat_case(case)
    w, h = f['canvas_width'], f['canvas_height']
    o, g = OracleRunner(f), GpuRunner(f)
    _compare(g.render(w, h, 1, 1, 0, output='sdf'), o.render(w, h, 1, 1, 0, output='sdf'), (case, 'sdf'))
    d = np.random.default_rng(5).uniform(-1, 1, (h, w, 1)).astype(np.float32)
    _compare(g.render(w, h, 1, 1, 0, output='sdf', d_sdf=d, translation=True),
             o.render(w, h, 1, 1, 0, output='sdf', d_sdf=d, translation=True), (case, 'sdf bwd'))
    _compare(g.render(w, h, 2, 2, 3, output='sdf'), o.render(w, h, 2, 2, 3, output='sdf'), (case, 'sdf 2x2'))


@pytest.mark.parametrize('case', ['strokes', 'overlap_nz', 'gradients'])
def test_gpu_tile_shape_does_not_change_results(case):
    """The autotuner's knob (tile width) changes only float summation order: forward
    images are identical and gradients agree to atomic-order noise."""
    f = scenes.flat_case(case)
    w, h = f['canvas_width'], f['canvas_height']
    d = np.random.default_rng(6).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    base = GpuRunner(f)
    ref_img = base.render(w, h, 2, 2, 3)['image']
    ref_g = base.render(w, h, 2, 2, 3, d_image=d)['grads']
    for tw in (1, 2, 4, 16, 32, 64):
        g = GpuRunner(f)
        g.scene.set_tiling(tw)
        assert np.array_equal(g.render(w, h, 2, 2, 3)['image'], ref_img), tw
        gg = g.render(w, h, 2, 2, 3, d_image=d)['grads']
        for k in GRAD_KEYS:
            assert _rel(gg[k], ref_g[k]) <= 1e-4, (tw, k)


@pytest.mark.parametrize('case', ['strokes', 'overlap_nz', 'gradients'])
def test_gpu_edge_samples_per_wave_does_not_change_results(case):
    """The autotuner's second knob (boundary samples per wave of the edge kernel: one or
    both colour queries of a sample per lane) changes only float summation order, also
    combined with non-square tiles; forward images are untouched by it."""
    f = scenes.flat_case(case)
    w, h = f['canvas_width'], f['canvas_height']
    d = np.random.default_rng(7).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    base = GpuRunner(f)
    ref_g = base.render(w, h, 2, 2, 3, d_image=d, translation=True)
    o = OracleRunner(f).render(w, h, 2, 2, 3, d_image=d, translation=True)
    for tw in (0, 4, 32):
        g = GpuRunner(f)
        g.scene.set_tiling(tw)
        g.scene.set_samples_per_wave(64)
        gg = g.render(w, h, 2, 2, 3, d_image=d, translation=True)
        assert _rel(gg['d_translation'], ref_g['d_translation']) <= 1e-4, tw
        for k in GRAD_KEYS:
            assert _rel(gg['grads'][k], ref_g['grads'][k]) <= 1e-4, (tw, k)
        _compare(gg, o, (case, tw, 'spw 64 vs oracle'))


def test_autotuner_picks_a_candidate():
    from diffvg_amd.autotune import Tuner
    f = scenes.flat_case('overlap_eo')
    sc = _lib.DeviceScene(f, 0)
    t = Tuner(reps=1)
    best = t.tune(sc, 64, 64, 2, 2)
    assert best in t.candidates(4)
    assert {spw for _, spw in t.candidates(4)} == {32, 64}
    assert t.apply(_lib.DeviceScene(f, 0), 64, 64, 2, 2) == best
