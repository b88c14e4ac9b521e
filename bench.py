"""Benchmark: fwd+bwd sample throughput of the render path on one MI355X.

Default workload: the 300-path, 512x512, 2x2, Gaussian-filter scene (prefiltering off),
one step = RenderFunction.apply + L2 loss + backward until the gradients are on the
device, scene marshalling and construction included.  Prints one JSON line.

The cpu_baseline leg runs in a subprocess: it times the reference's own CPU path
(oracle/_ref, built from the reference sources) on the same scene with the tent-filter
twin (the reference has no Gaussian filter) and counts the algorithmic work of each
pass with the oracle's instrumented traversal on a strided sample of the samples; the
counts, priced with the cost table below, give the roofline's algorithmic FLOPs.
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# FP32-equivalent FLOPs per counted query (an FP64 operation counts 2, a transcendental
# 16, a comparison 1).  Counter order of orc_counters.
COST = {
    'samples': 60,                # PCG position, canvas transform, colour gather/splat bookkeeping
    'top_box_tests': 8,
    'group_box_tests': 8,
    'seg_box_tests': 6,
    'line_tests': 16,
    'quad_tests': 120,            # float cubic solve (3 transcendentals in double)
    'cubic_winding_tests': 330,   # double cubic solve + up to 3 root evaluations in double
    'cubic_distance_tests': 420,  # endpoint tests, quintic set-up, isolator cubic in double
    'newton_iterations': 72,      # one double quintic + derivative evaluation and update
    'closest_point_tests': 12,
    'fragments': 30,              # paint lookup + over compositing
    'boundary_samples': 160,      # CDF searches, boundary point, normal, transforms
}
COUNTERS = list(COST)
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def cpu_baseline_main(cfg):
    """Subprocess body of the cpu_baseline leg (no GPU use)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    from diffvg_amd import workloads
    from refdrive import flatten, load_ref, OracleRunner, RefRunner, load_oracle
    c = workloads.closed_paths_300(device='cpu', filter_type='tent') if cfg == 2 else workloads.WORKLOADS[cfg](device='cpu')
    w, h = c['canvas']
    nsx, nsy = c['nsx'], c['nsy']
    f = flatten(w, h, c['shapes'], c['groups'], c['filter'].type, float(c['filter'].radius))
    out = {}
    d_img = np.random.default_rng(0).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    if load_ref() is not None:
        r = RefRunner(f)
        t0 = time.time()
        r.render(w, h, nsx, nsy, 0)
        t1 = time.time()
        r.render(w, h, nsx, nsy, 0, d_image=d_img)
        t2 = time.time()
        out.update(kind='reference', fwd_s=t1 - t0, bwd_s=t2 - t1, cores=os.cpu_count())
    else:
        o = OracleRunner(f)
        t0 = time.time()
        o.render(w, h, nsx, nsy, 0)
        t1 = time.time()
        o.render(w, h, nsx, nsy, 0, d_image=d_img)
        t2 = time.time()
        out.update(kind='port', fwd_s=t1 - t0, bwd_s=t2 - t1, cores=load_oracle().orc_get_num_threads())
    out['samples'] = w * h * nsx * nsy
    # algorithmic work of each pass, counted on every `stride`-th sample and scaled
    lib = load_oracle()
    stride = 4
    o = OracleRunner(f)
    lib.orc_set_counting(1)
    lib.orc_set_sample_stride(stride)
    cnt = (ctypes.c_int64 * 12)()
    o.render(w, h, nsx, nsy, 0)
    lib.orc_get_counters(0, ctypes.byref(cnt))
    out['counts_fwd'] = [int(x) * stride for x in cnt]
    o.render(w, h, nsx, nsy, 0, d_image=d_img)
    lib.orc_get_counters(0, ctypes.byref(cnt))
    out['counts_bwd'] = [int(x) * stride for x in cnt]
    lib.orc_get_counters(1, ctypes.byref(cnt))
    out['counts_edge'] = [int(x) * stride for x in cnt]
    lib.orc_set_counting(0)
    lib.orc_set_sample_stride(1)
    print(json.dumps(out))


def flops(counts, adjoint=False):
    f = sum(COST[k] * counts[i] for i, k in enumerate(COUNTERS))
    if adjoint:  # reverse compositing sweep and paint adjoints
        f += 40 * counts[COUNTERS.index('fragments')]
    return float(f)


def run_cpu_baseline(cfg):
    try:
        p = subprocess.run([sys.executable, os.path.abspath(__file__), '--cpu-baseline-only', '--config', str(cfg)],
                           capture_output=True, text=True, timeout=600)
        line = [l for l in p.stdout.strip().splitlines() if l.startswith('{')][-1]
        return json.loads(line)
    except Exception as e:  # the baseline is reported, never required
        return {'error': repr(e)}


def make_params(c, device):
    import torch
    params = []
    for s in c['shapes']:
        for name in ('points', 'radius', 'center', 'p_min', 'p_max'):
            t = getattr(s, name, None)
            if isinstance(t, torch.Tensor) and t.is_floating_point():
                t.requires_grad_(True)
                params.append(t)
        if isinstance(s.stroke_width, torch.Tensor):
            s.stroke_width.requires_grad_(True)
            params.append(s.stroke_width)
    for g in c['groups']:
        for col in (g.fill_color, g.stroke_color):
            if isinstance(col, torch.Tensor):
                col.requires_grad_(True)
                params.append(col)
    return params


def _scene_params(shapes, groups):
    import torch
    ps = []
    for sh in shapes:
        for name in ('points', 'radius', 'center', 'p_min', 'p_max', 'stroke_width'):
            t = getattr(sh, name, None)
            if isinstance(t, torch.Tensor) and t.is_floating_point():
                ps.append(t.requires_grad_(True))
    for g in groups:
        for col in (g.fill_color, g.stroke_color):
            if isinstance(col, torch.Tensor):
                ps.append(col.requires_grad_(True))
            elif col is not None:
                for name in ('begin', 'end', 'center', 'radius', 'offsets', 'stop_colors'):
                    t = getattr(col, name, None)
                    if isinstance(t, torch.Tensor):
                        ps.append(t.requires_grad_(True))
    return ps


def bench_batched(a):
    """256 uneven scenes (path counts log-uniform in [1, 512], 2-8 cubics per path, 10%
    gradient fills, resolution in {28, 64, 128, 256}, 2x2 spp, box filter), one fwd+bwd
    each with an L2 loss against a per-scene target.  Three measurements:
      graph    -- scenes resident; forward batch, loss gradient and backward batch
                  captured in one HIP graph and replayed (value);
      batched  -- per step, the user-level API: serialize every scene, render_batch
                  (pooled scenes rebuilt on the device, one batched forward and one
                  batched backward), L2 loss, backward into every parameter tensor;
      plain    -- per step: one RenderFunction call per scene on the PLAIN schedule
                  (the reference's one-scene-per-launch structure), loss, backward."""
    import numpy as np
    import torch
    import diffvg_amd as pydiffvg
    from diffvg_amd import workloads, _lib
    from diffvg_amd.batch import Batch
    from diffvg_amd.render_pytorch import flatten_args

    dev = torch.device('cuda:0')
    pydiffvg.set_use_gpu(True)
    pydiffvg.set_device(dev)
    items = workloads.uneven_scenes(device=dev)
    n = len(items)
    params = [_scene_params(it['shapes'], it['groups']) for it in items]
    sizes = [it['canvas'] for it in items]
    npix = [w * h for w, h in sizes]
    offs = np.concatenate([[0], np.cumsum(npix)]).astype(np.int64)
    samples = int(sum(npix)) * 4
    rng = np.random.default_rng(4)
    target_flat = torch.from_numpy(rng.uniform(0, 1, (int(offs[-1]), 4)).astype(np.float32)).to(dev)
    img_flat = torch.zeros(int(offs[-1]), 4, device=dev)
    d_flat = torch.zeros(int(offs[-1]), 4, device=dev)

    def view(flat, i):
        w, h = sizes[i]
        return flat[offs[i]:offs[i + 1]].view(h, w, 4)

    # ---- graph: resident scenes, one captured fwd + loss + bwd
    def build(i):
        it = items[i]
        w, h = sizes[i]
        args = pydiffvg.RenderFunction.serialize_scene(w, h, it['shapes'], it['groups'], it['filter'])
        fl = flatten_args(args)
        return args, fl, _lib.DeviceScene(fl.desc, 0)

    built = [build(i) for i in range(n)]
    tuned = {}
    if a.autotune:
        # tile width per workload signature, timed once per signature on the first scene
        from diffvg_amd.autotune import Tuner
        tuner = Tuner()
        for i, (_, _, ds) in enumerate(built):
            tuner.tune(ds, sizes[i][0], sizes[i][1], 2, 2)
        tuned = {str(k): v for k, v in tuner.best.items()}
    fjobs = [(ds, dict(width=sizes[i][0], height=sizes[i][1], nsx=2, nsy=2, seed=0, render_image=view(img_flat, i)))
             for i, (_, _, ds) in enumerate(built)]
    bjobs = [(ds, dict(width=sizes[i][0], height=sizes[i][1], nsx=2, nsy=2, seed=0, d_render_image=view(d_flat, i)))
             for i, (_, _, ds) in enumerate(built)]
    fb, bb = Batch(fjobs, zero_grads=False), Batch(bjobs, zero_grads=True)
    stream = torch.cuda.Stream()

    def graph_body(st):
        img_flat.zero_()
        fb.render(st)
        torch.mul(img_flat - target_flat, 2.0, out=d_flat)
        bb.render(st)

    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        graph_body(stream.cuda_stream)
    torch.cuda.synchronize()
    fb.status()
    bb.status()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        graph_body(torch.cuda.current_stream().cuda_stream)
    for _ in range(a.warmup):
        g.replay()
    torch.cuda.synchronize()
    reps = max(a.steps, 5)
    L = _lib.lib()
    L.dvg_profile_enable(0)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / reps
    fb.status()
    bb.status()
    # eager kernel breakdown of the same batches
    names = ctypes.create_string_buffer(4096)
    ms = (ctypes.c_double * 32)()
    cnts = (ctypes.c_int * 32)()
    L.dvg_profile_read(names, 4096, ms, cnts, 32)
    L.dvg_profile_enable(1)
    graph_body(None)
    torch.cuda.synchronize()
    L.dvg_profile_enable(0)
    nk = L.dvg_profile_read(names, 4096, ms, cnts, 32)
    kern = {nm: round(ms[i], 4) for i, nm in enumerate(names.value.decode().strip().split('\n')[:nk])}

    # ---- batched end to end: the user-level autograd API over all scenes per step
    # (serialize_scene per scene, pooled scenes rebuilt on the device, one batched
    # forward and one batched backward, gradients delivered to every parameter tensor)
    from diffvg_amd.batch_render import render_batch
    scene_list = [(sizes[i][0], sizes[i][1], items[i]['shapes'], items[i]['groups']) for i in range(n)]
    targets = [view(target_flat, i) for i in range(n)]

    def batched_step():
        for ps in params:
            for p in ps:
                p.grad = None
        imgs = render_batch(scene_list, 2, 2, [0] * n)
        loss = sum(((im - t) ** 2).sum() for im, t in zip(imgs, targets))
        loss.backward()

    for _ in range(2):
        batched_step()
    torch.cuda.synchronize()
    nb = max(2, a.steps // 4)
    t0 = time.perf_counter()
    for _ in range(nb):
        batched_step()
    torch.cuda.synchronize()
    t_batched = (time.perf_counter() - t0) / nb

    # ---- plain: one RenderFunction call per scene (PLAIN schedule)
    pydiffvg.set_render_mode('plain')

    def plain_step():
        for ps in params:
            for p in ps:
                p.grad = None
        for i, it in enumerate(items):
            w, h = sizes[i]
            args = pydiffvg.RenderFunction.serialize_scene(w, h, it['shapes'], it['groups'], it['filter'])
            img = pydiffvg.RenderFunction.apply(w, h, 2, 2, 0, None, *args)
            ((img - view(target_flat, i)) ** 2).sum().backward()

    plain_step()
    torch.cuda.synchronize()
    np_ = max(1, a.steps // 10)
    t0 = time.perf_counter()
    for _ in range(np_):
        plain_step()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / np_
    pydiffvg.set_render_mode('tiled')

    out = {
        'metric': 'fwd+bwd Msamples/s on 1 GPU; p50/p99 latency, batched/served vs plain',
        'value': round(samples / t_graph / 1e6, 4), 'unit': 'Msamples/s', 'n_gpus': 1, 'steps': reps,
        'warmup': a.warmup, 'ms_per_step': round(t_graph * 1e3, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32 (f64 root solves)', 'data': 'synthetic',
        'config': {'workload': 'uneven_scenes_256', 'scenes': n, 'samples': samples,
                   'paths': int(sum(len(it['shapes']) for it in items)), 'spp': [2, 2], 'filter': 'box r=0.5',
                   'mode': 'batched, one HIP graph per fwd+bwd'},
        'batched_end_to_end': {'ms_per_step': round(t_batched * 1e3, 3),
                               'value': round(samples / t_batched / 1e6, 4)},
        'plain': {'ms_per_step': round(t_plain * 1e3, 3), 'value': round(samples / t_plain / 1e6, 4)},
        'speedup_graph_vs_plain': round(t_plain / t_graph, 2),
        'speedup_batched_vs_plain': round(t_plain / t_batched, 2),
        'kernels_ms_eager': kern,
        'autotuned_tile_widths': tuned,
    }
    print(json.dumps(out))


def bench_serving(a):
    """Serving stream: Poisson arrivals of render + gradient requests on resident scenes
    (path counts log-uniform in [1, 4096], resolution in {128, 256, 512, 1024}, colour or
    SDF, prefiltering on or off).  The arrival rate is 0.8x the plain path's sustained
    request rate measured in the same run (one request at a time, PLAIN schedule); the
    same arrival sequence is then served by the continuous-batching server and by the
    plain path, and p50/p99 latency (arrival to gradients ready) and throughput are
    reported for both."""
    import numpy as np
    import torch
    import diffvg_amd as pydiffvg
    from diffvg_amd import workloads, _lib, serve
    from diffvg_amd.render_pytorch import flatten_args

    dev = torch.device('cuda:0')
    pydiffvg.set_use_gpu(True)
    pydiffvg.set_device(dev)
    rng = np.random.default_rng(5)
    n_req = a.requests
    specs = []
    for i in range(n_req):
        r = workloads.serving_request(rng, device='cpu')
        w, h = r['canvas']
        args = pydiffvg.RenderFunction.serialize_scene(w, h, r['shapes'], r['groups'], r['filter'])
        scene = _lib.DeviceScene(flatten_args(args).desc, 0)
        ch = 1 if r['sdf'] else 4
        target = torch.from_numpy(rng.uniform(0, 1, (h, w, ch)).astype(np.float32)).to(dev)
        specs.append((scene, w, h, r['nsx'], r['nsy'], r['sdf'], r['prefilter'], target, len(r['shapes'])))

    def make(i):
        sc, w, h, nsx, nsy, sdf, pre, target, _ = specs[i]
        return serve.Request(i, sc, w, h, nsx, nsy, 0, sdf, pre, target, 0.0)

    # plain service rate, back to back
    cal = [make(i) for i in range(n_req)]
    serve.serve_plain(cal[:4], [0.0] * 4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    serve.serve_plain(cal, [0.0] * n_req)
    t_plain_all = time.perf_counter() - t0
    mu = n_req / t_plain_all
    lam = 0.8 * mu
    arrivals = np.cumsum(rng.exponential(1.0 / lam, n_req))
    arrivals -= arrivals[0]

    # batched server (warm-up on the first requests, then the timed stream)
    srv = serve.Server(device=0, hbm_budget=a.hbm_budget << 30, max_batch=a.max_batch, streams=3)
    srv.run([make(i) for i in range(min(8, n_req))], [0.0] * min(8, n_req))
    torch.cuda.synchronize()
    srv = serve.Server(device=0, hbm_budget=a.hbm_budget << 30, max_batch=a.max_batch, streams=3)
    done = srv.run([make(i) for i in range(n_req)], list(arrivals))
    st_b = serve.latency_stats(done)
    plain_done = serve.serve_plain([make(i) for i in range(n_req)], list(arrivals))
    st_p = serve.latency_stats(plain_done)
    # each request alone on an idle device (its service time without queueing), and the
    # requests with the longest served latency
    solo = []
    for i in range(n_req):
        one = serve.Server(device=0, hbm_budget=a.hbm_budget << 30, max_batch=1, streams=1)
        r = one.run([make(i)], [0.0])[0]
        solo.append((r.done - r.arrival) * 1e3)
    slowest = sorted(done, key=lambda r: r.arrival - r.done)[:5]
    tail = [{'request': r.rid, 'latency_ms': round((r.done - r.arrival) * 1e3, 2),
             'alone_ms': round(solo[r.rid], 2), 'size': specs[r.rid][1], 'paths': specs[r.rid][8],
             'sdf': bool(specs[r.rid][5]), 'prefilter': bool(specs[r.rid][6])} for r in slowest]
    samples = sum(w * h * (1 if sdf else nsx * nsy) for (_, w, h, nsx, nsy, sdf, _, _, _) in specs)
    span_b = max(r.done for r in done) - min(r.arrival for r in done)
    out = {
        'metric': 'fwd+bwd Msamples/s on 1 GPU; p50/p99 latency, batched/served vs plain',
        'value': round(samples / span_b / 1e6, 4), 'unit': 'Msamples/s', 'n_gpus': 1, 'steps': n_req,
        'warmup': 8, 'ms_per_step': round(span_b / n_req * 1e3, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32 (f64 root solves)', 'data': 'synthetic',
        'config': {'workload': 'serving_stream', 'requests': n_req, 'samples': samples,
                   'mean_paths': float(np.mean([s_[8] for s_ in specs])),
                   'sdf_fraction': float(np.mean([s_[5] for s_ in specs])),
                   'arrival_rate_rps': round(lam, 3), 'plain_service_rate_rps': round(mu, 3),
                   'hbm_budget_gib': a.hbm_budget, 'max_batch': a.max_batch, 'streams': 3},
        'served': {k: round(v, 3) for k, v in st_b.items()},
        'plain': {k: round(v, 3) for k, v in st_p.items()},
        'alone': {'p50_ms': round(float(np.percentile(solo, 50)), 3),
                  'p99_ms': round(float(np.percentile(solo, 99)), 3), 'max_ms': round(max(solo), 3)},
        'slowest_served': tail,
    }
    print(json.dumps(out))


def copy_calibration(mib=256, reps=3):
    """A plain device copy of a known size (reads and writes `mib` MiB per dispatch), run
    under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.rand(n, device='cuda:0')
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({'copy_mib': mib, 'dispatches': reps}))


def load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed PMC summary
    (profiles/hbm_traffic.md rows: | kernel | workload | ... | bytes per launch |)."""
    path = os.path.join(ROOT, 'profiles', 'hbm_traffic.md')
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip().strip('`') for c in line.strip().strip('|').split('|')]
        if len(cells) >= 3 and cells[0] == kernel and cells[1] == workload:
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None,
                    help='timed steps (default 50 for one scene, 20 for the 256-scene batch)')
    ap.add_argument('--warmup', type=int, default=None,
                    help='untimed steps first (default 20 for one scene: clocks and pools settle; 3 for the batch)')
    ap.add_argument('--config', type=int, default=2)
    ap.add_argument('--mode', default='tiled', choices=['tiled', 'plain'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-baseline-only', action='store_true')
    ap.add_argument('--copy-calibration', action='store_true', help='run a plain copy kernel of known size')
    ap.add_argument('--requests', type=int, default=120, help='serving stream length (config 5)')
    ap.add_argument('--autotune', action='store_true', help='pick tile widths per workload (config 4)')
    ap.add_argument('--hbm-budget', type=int, default=64, help='GiB that requests in flight may hold (config 5)')
    ap.add_argument('--max-batch', type=int, default=32, help='requests per batch (config 5)')
    a = ap.parse_args()
    if a.steps is None:
        a.steps = 20 if a.config == 4 else 50
    if a.warmup is None:
        a.warmup = 3 if a.config == 4 else 20
    if a.cpu_baseline_only:
        cpu_baseline_main(a.config)
        return
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print('bench.py measures one GPU: the render path is built and measured on one MI355X, so --gpus %d '
              'is not run (how it would shard is a design note, not a measured path)' % a.gpus, file=sys.stderr)
        sys.exit(2)

    if a.config == 4:
        bench_batched(a)
        return
    if a.config == 5:
        bench_serving(a)
        return
    cpu = None if a.no_cpu_baseline else run_cpu_baseline(a.config)

    import numpy as np
    import torch
    import diffvg_amd as pydiffvg
    from diffvg_amd import workloads, _lib

    dev = torch.device('cuda:0')
    pydiffvg.set_use_gpu(True)
    pydiffvg.set_device(dev)
    pydiffvg.set_render_mode(a.mode)
    c = workloads.WORKLOADS[a.config](device=dev)
    w, h = c['canvas']
    nsx, nsy = c['nsx'], c['nsy']
    params = make_params(c, dev)
    if 'target' in c:
        target = c['target']
    else:
        target = torch.from_numpy(np.random.default_rng(a.config).uniform(0, 1, (h, w, 4)).astype(np.float32)).to(dev)
    render = pydiffvg.RenderFunction.apply

    def step(prefilter=False):
        for p in params:
            p.grad = None
        args = pydiffvg.RenderFunction.serialize_scene(w, h, c['shapes'], c['groups'], c['filter'],
                                                      use_prefiltering=prefilter)
        img = render(w, h, nsx, nsy, 0, None, *args)
        loss = ((img - target) ** 2).sum()
        loss.backward()
        return loss

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    L = _lib.lib()
    L.dvg_profile_enable.argtypes = [ctypes.c_int]
    L.dvg_profile_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    names = ctypes.create_string_buffer(4096)
    ms = (ctypes.c_double * 32)()
    cnts = (ctypes.c_int * 32)()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # per-kernel times: the same steps again with the library's events around each launch
    # (kept out of the timed steps above, whose launches they would space out)
    L.dvg_profile_read(names, 4096, ms, cnts, 32)
    L.dvg_profile_enable(1)
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    L.dvg_profile_enable(0)
    nk = L.dvg_profile_read(names, 4096, ms, cnts, 32)
    kern = {}
    for i, nm in enumerate(names.value.decode().strip().split('\n')[:nk]):
        kern[nm] = {'total_ms': ms[i], 'launches': cnts[i], 'avg_ms': ms[i] / max(cnts[i], 1)}
    samples = w * h * nsx * nsy
    sec = (t1 - t0) / a.steps
    value = samples / sec / 1e6
    kernel_ms = sum(v['total_ms'] for v in kern.values()) / a.steps

    # prefiltered variant of the same scene (reported alongside)
    pre = None
    if a.config == 2:
        for _ in range(2):
            step(True)
        torch.cuda.synchronize()
        n_pre = max(3, a.steps // 4)
        tp0 = time.perf_counter()
        for _ in range(n_pre):
            step(True)
        torch.cuda.synchronize()
        tp1 = time.perf_counter()
        pre = {'value': samples / ((tp1 - tp0) / n_pre) / 1e6, 'ms_per_step': (tp1 - tp0) / n_pre * 1e3}

    roofline = None
    if cpu and 'counts_fwd' in cpu and kern:
        per_kernel_flops = {'k_color': flops(cpu['counts_fwd']), 'kp_render': flops(cpu['counts_fwd']),
                            'k_color_bwd': flops(cpu['counts_bwd'], adjoint=True),
                            'k_edge': flops(cpu['counts_edge']), 'kp_edge': flops(cpu['counts_edge'])}
        dom = max((k for k in kern if k in per_kernel_flops), key=lambda k: kern[k]['total_ms'], default=None)
        if dom:
            avg_s = kern[dom]['avg_ms'] / 1e3
            achieved = per_kernel_flops[dom] / avg_s / 1e12
            roofline = {'bound': 'mfma', 'achieved': round(achieved, 4), 'peak': PEAK_FP32_TFLOPS,
                        'unit': 'TFLOP/s', 'frac': round(achieved / PEAK_FP32_TFLOPS, 6),
                        'traffic': load_traffic(dom, c['name']),
                        'kernel': dom, 'kernel_avg_ms': round(kern[dom]['avg_ms'], 4),
                        'flops_per_launch': per_kernel_flops[dom],
                        'note': 'vector FP32 peak; FLOPs are FP32-equivalent algorithmic counts'}
    cpu_baseline = None
    if cpu and 'fwd_s' in cpu:
        cpu_sec = cpu['fwd_s'] + cpu['bwd_s']
        cpu_baseline = {'value': round(cpu['samples'] / cpu_sec / 1e6, 4), 'unit': 'Msamples/s',
                        'cores': cpu['cores'], 'kind': cpu['kind'],
                        'sample': 'the full %dx%d %dx%d-spp scene, one forward and one backward (edge sampling), %s'
                                  % (w, h, nsx, nsy, 'tent r=1 twin of the Gaussian filter' if a.config == 2
                                     else 'the same box r=0.5 filter')}
    out = {
        'metric': 'fwd+bwd Msamples/s on 1 GPU; p50/p99 latency, batched/served vs plain',
        'value': round(value, 4), 'unit': 'Msamples/s', 'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_step': round(sec * 1e3, 4), 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 (f64 root solves)', 'data': 'synthetic',
        'config': {'workload': c['name'], 'canvas': [w, h], 'spp': [nsx, nsy],
                   'filter': 'gaussian r=1.5' if a.config == 2 else 'box r=0.5',
                   'paths': len(c['shapes']), 'mode': a.mode, 'prefilter': False},
        'kernel_ms_per_step': round(kernel_ms, 4),
        'kernels': {k: {'avg_ms': round(v['avg_ms'], 4), 'launches': v['launches']} for k, v in kern.items()},
        'prefilter_on': pre,
        'roofline': roofline,
        'cpu_baseline': cpu_baseline,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
