"""Benchmark of the render path on one MI355X; prints one JSON line.

Default line (`python bench.py`): 256 uneven scenes (path counts log-uniform in [1, 512],
2-8 cubics per path, 10% gradient fills, 28-256 px, 2x2 spp, box filter), one forward +
backward each with an L2 loss, rendered as one forward batch, the loss gradient and one
backward batch replayed from a HIP graph with the scenes resident in HBM (`value`,
Msamples/s), against the plain path (one RenderFunction call per scene on the PLAIN
schedule, the reference's launch structure) and the user-level `render_batch` end to
end.  Extra keys: the serving stream (Poisson arrivals at 0.8x the plain service rate,
p50/p99 latency served and plain, an arrival-rate sweep to the rate where the served
p99 passes 50 ms), the two 512x512 single scenes (300 closed paths with a Gaussian
filter; 2,048 strokes) with their PLAIN ratios, the roofline of the dominant batched
kernel and the CPU baseline.  `--config 2|3|4|5` runs one workload alone.

The cpu_baseline leg runs in a subprocess: it times the reference's own CPU path
(oracle/_ref, built from the reference sources) on a bounded sample of the same
workload, and counts the algorithmic work of each pass with the oracle's instrumented
traversal on a strided sample of the samples; the counts, priced with the cost table
below, give the roofline's algorithmic FLOPs.
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
# uneven scenes timed on the host cores for the batched workload's CPU baseline
CPU_SAMPLE_SCENES = 64
PEAK_HBM_GBS = 8000.0


def progress(msg):
    """One line per finished part of the run, on stderr (the JSON line is stdout's only
    output)."""
    print('[bench %s] %s' % (time.strftime('%H:%M:%S'), msg), file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return None


def _count_passes(lib, runner, w, h, nsx, nsy, d_img, stride):
    """Oracle query counts of one forward and one backward on every `stride`-th sample,
    scaled to all samples: (forward, backward interior, backward edge)."""
    lib.orc_set_counting(1)
    lib.orc_set_sample_stride(stride)
    cnt = (ctypes.c_int64 * 12)()
    runner.render(w, h, nsx, nsy, 0)
    lib.orc_get_counters(0, ctypes.byref(cnt))
    fwd = [int(x) * stride for x in cnt]
    runner.render(w, h, nsx, nsy, 0, d_image=d_img)
    lib.orc_get_counters(0, ctypes.byref(cnt))
    bwd = [int(x) * stride for x in cnt]
    lib.orc_get_counters(1, ctypes.byref(cnt))
    edge = [int(x) * stride for x in cnt]
    lib.orc_set_counting(0)
    lib.orc_set_sample_stride(1)
    return fwd, bwd, edge


def cpu_baseline_main(cfg):
    """Subprocess body of the cpu_baseline leg (no GPU use).  Times the reference's CPU
    build (oracle/_ref; the oracle when that is absent) and counts the algorithmic work
    of each pass with the oracle's instrumented traversal."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    from diffvg_amd import workloads
    from refdrive import flatten, load_ref, OracleRunner, RefRunner, load_oracle
    lib = load_oracle()
    threads = os.cpu_count() or 1
    lib.orc_set_num_threads(threads)
    if cfg == 4:
        items = workloads.uneven_scenes(device='cpu')
        scenes = []
        for k, it in enumerate(items):
            w, h = it['canvas']
            f = flatten(w, h, it['shapes'], it['groups'], it['filter'].type, float(it['filter'].radius))
            scenes.append((f, w, h, np.random.default_rng(k).uniform(-1, 1, (h, w, 4)).astype(np.float32)))
        nsx = nsy = 2
        timed = scenes[:CPU_SAMPLE_SCENES]
        sample = ('the first %d of the 256 uneven scenes (%d samples), one forward and one backward (edge '
                  'sampling) each' % (len(timed), sum(w * h * 4 for _, w, h, _ in timed)))
        stride = 8
    else:
        c = (workloads.closed_paths_300(device='cpu', filter_type='tent') if cfg == 2
             else workloads.WORKLOADS[cfg](device='cpu'))
        w, h = c['canvas']
        nsx, nsy = c['nsx'], c['nsy']
        f = flatten(w, h, c['shapes'], c['groups'], c['filter'].type, float(c['filter'].radius))
        scenes = timed = [(f, w, h, np.random.default_rng(0).uniform(-1, 1, (h, w, 4)).astype(np.float32))]
        sample = ('the full %dx%d %dx%d-spp scene, one forward and one backward (edge sampling), %s'
                  % (w, h, nsx, nsy, 'tent r=1 twin of the Gaussian filter' if cfg == 2 else 'the same box filter'))
        stride = 4
    out = {'sample': sample, 'cpu_model': cpu_model()}
    have_ref = load_ref() is not None
    t_fwd = t_bwd = 0.0
    for f, w, h, d_img in timed:
        r = RefRunner(f) if have_ref else OracleRunner(f, threads=threads)
        t0 = time.time()
        r.render(w, h, nsx, nsy, 0)
        t1 = time.time()
        r.render(w, h, nsx, nsy, 0, d_image=d_img)
        t2 = time.time()
        t_fwd += t1 - t0
        t_bwd += t2 - t1
    out.update(kind='reference' if have_ref else 'port', fwd_s=t_fwd, bwd_s=t_bwd,
               cores=os.cpu_count() if have_ref else lib.orc_get_num_threads(),
               samples=sum(w * h * nsx * nsy for _, w, h, _ in timed))
    tot = [np.zeros(12, np.int64) for _ in range(3)]
    for f, w, h, d_img in scenes:
        o = OracleRunner(f, threads=threads)
        for acc, c_ in zip(tot, _count_passes(lib, o, w, h, nsx, nsy, d_img, stride)):
            acc += np.array(c_, np.int64)
    out['counts_fwd'], out['counts_bwd'], out['counts_edge'] = [t.tolist() for t in tot]
    out['count_stride'] = stride
    print(json.dumps(out))


def flops(counts, adjoint=False):
    f = sum(COST[k] * counts[i] for i, k in enumerate(COUNTERS))
    if adjoint:  # reverse compositing sweep and paint adjoints
        f += 40 * counts[COUNTERS.index('fragments')]
    return float(f)


def run_cpu_baseline(cfg):
    progress('cpu baseline (subprocess)')
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


METRIC = 'fwd+bwd Msamples/s on 1 GPU; p50/p99 latency, batched/served vs plain'
DTYPE = 'f32 (f64 root solves)'


def _profile_lib():
    from diffvg_amd import _lib
    L = _lib.lib()
    names = ctypes.create_string_buffer(4096)
    ms = (ctypes.c_double * 32)()
    cnts = (ctypes.c_int * 32)()

    def read():
        nk = L.dvg_profile_read(names, 4096, ms, cnts, 32)
        return {nm: {'total_ms': ms[i], 'launches': cnts[i], 'avg_ms': ms[i] / max(cnts[i], 1)}
                for i, nm in enumerate(names.value.decode().strip().split('\n')[:nk])}
    return L, read


def leg_batched(steps, warmup, autotune=False, e2e_steps=4, plain_steps=1):
    """256 uneven scenes (path counts log-uniform in [1, 512], 2-8 cubics per path, 10%
    gradient fills, resolution in {28, 64, 128, 256}, 2x2 spp, box filter), one fwd+bwd
    each with an L2 loss against a per-scene target.  Three measurements:
      graph    -- scenes resident; forward batch, loss gradient and backward batch
                  captured in one HIP graph and replayed (`steps` timed replays);
      batched  -- per step, the user-level API: serialize every scene, render_batch
                  (pooled scenes rebuilt on the device, one batched forward and one
                  batched backward), L2 loss, backward into every parameter tensor;
      plain    -- per step: one RenderFunction call per scene on the PLAIN schedule
                  (the reference's one-scene-per-launch structure), loss, backward.
    Per-kernel times come from HIP events on the launch stream over `steps` eager
    renders of the same two batches right after the timed replays."""
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
        return _lib.DeviceScene(flatten_args(args).desc, 0)

    built = [build(i) for i in range(n)]
    progress('batch: %d scenes built' % n)
    fjobs = [(ds, dict(width=sizes[i][0], height=sizes[i][1], nsx=2, nsy=2, seed=0, render_image=view(img_flat, i)))
             for i, ds in enumerate(built)]
    bjobs = [(ds, dict(width=sizes[i][0], height=sizes[i][1], nsx=2, nsy=2, seed=0, d_render_image=view(d_flat, i)))
             for i, ds in enumerate(built)]

    def make_batches(persistent):
        return (Batch(fjobs, zero_grads=False, persistent=persistent),
                Batch(bjobs, zero_grads=True, persistent=persistent))

    def eager_step(batches, st=None):
        img_flat.zero_()
        batches[0].render(st)
        torch.mul(img_flat - target_flat, 2.0, out=d_flat)
        batches[1].render(st)

    tuned = None
    persistent = False
    if autotune:
        # per scene signature: tile width and edge samples per wave; then the batch's
        # launch structure (one launch per stage, or one persistent launch per pass)
        from diffvg_amd.autotune import Tuner
        tuner = Tuner()
        t_tune = time.perf_counter()
        for i, ds in enumerate(built):
            tuner.tune(ds, sizes[i][0], sizes[i][1], 2, 2)
        persistent, _ = tuner.tune_batch(make_batches, eager_step)
        torch.cuda.synchronize()
        progress('batch: autotuned in %.1f s' % (time.perf_counter() - t_tune))
        tuned = {'choices': {str(k): list(v) for k, v in tuner.best.items()},
                 'schedule': 'persistent' if persistent else 'staged',
                 'schedule_ms': {k: (None if v is None else round(v, 3)) for k, v in tuner.batch_timings.items()},
                 'tune_s': round(time.perf_counter() - t_tune, 2)}
    fb, bb = make_batches(persistent)
    # per-kernel timing (below) runs the one-launch-per-stage batches
    fb_k, bb_k = (fb, bb) if not persistent else make_batches(False)
    stream = torch.cuda.Stream()

    def graph_body(st):
        eager_step((fb, bb), st)

    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        graph_body(stream.cuda_stream)
    torch.cuda.synchronize()
    fb.status()
    bb.status()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        graph_body(torch.cuda.current_stream().cuda_stream)
    for _ in range(warmup):
        g.replay()
    torch.cuda.synchronize()
    L, read_profile = _profile_lib()
    L.dvg_profile_enable(0)
    read_profile()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / steps
    progress('batch: graph replay %.2f ms' % (t_graph * 1e3))
    fb.status()
    bb.status()
    # per-kernel times: the same jobs rendered eagerly, one launch per stage, with events
    # around each launch
    L.dvg_profile_enable(1)
    for _ in range(steps):
        eager_step((fb_k, bb_k))
    torch.cuda.synchronize()
    L.dvg_profile_enable(0)
    kern = read_profile()
    fb_k.status()
    bb_k.status()

    # ---- batched end to end through the user-level autograd API
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

    t_batched = None
    if e2e_steps:
        for _ in range(2):
            batched_step()
        pydiffvg.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(e2e_steps):
            batched_step()
        pydiffvg.synchronize()
        torch.cuda.synchronize()
        t_batched = (time.perf_counter() - t0) / e2e_steps
        progress('batch: render_batch end to end %.1f ms' % (t_batched * 1e3))
    t_direct = None
    if e2e_steps:
        # the same steps with the gradients written straight into the parameters' .grad
        pydiffvg.set_direct_leaf_gradients(True)
        try:
            for _ in range(2):
                batched_step()
            pydiffvg.synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(e2e_steps):
                batched_step()
            pydiffvg.synchronize()
            torch.cuda.synchronize()
            t_direct = (time.perf_counter() - t0) / e2e_steps
        finally:
            pydiffvg.set_direct_leaf_gradients(False)
        progress('batch: render_batch end to end, direct leaf gradients %.1f ms' % (t_direct * 1e3))

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

    t_plain = None
    if plain_steps:
        plain_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(plain_steps):
            plain_step()
        pydiffvg.synchronize()
        torch.cuda.synchronize()
        t_plain = (time.perf_counter() - t0) / plain_steps
        progress('batch: plain %.1f ms' % (t_plain * 1e3))
    pydiffvg.set_render_mode('tiled')
    return dict(samples=samples, scenes=n, paths=int(sum(len(it['shapes']) for it in items)), t_graph=t_graph,
                t_batched=t_batched, t_direct=t_direct, t_plain=t_plain, kernels=kern, tuned=tuned, steps=steps,
                e2e_steps=e2e_steps, plain_steps=plain_steps)


def leg_serving(n_req, hbm_budget_gib=64, max_batch=32, sweep=True, p99_bar_ms=50.0):
    """Serving stream: Poisson arrivals of render + gradient requests on resident scenes
    (path counts log-uniform in [1, 4096], resolution in {128, 256, 512, 1024}, colour or
    SDF, prefiltering on or off).  The arrival rate is 0.8x the plain path's sustained
    request rate measured in the same run (one request at a time, PLAIN schedule); the
    same arrival sequence is then served by the continuous-batching server and by the
    plain path, and p50/p99 latency (arrival to gradients ready) and throughput are
    reported for both.  The sweep then doubles the arrival rate until the served p99
    passes `p99_bar_ms`; the highest rate that stays under it is the saturation rate."""
    import numpy as np
    import torch
    import diffvg_amd as pydiffvg
    from diffvg_amd import workloads, _lib, serve
    from diffvg_amd.render_pytorch import flatten_args

    dev = torch.device('cuda:0')
    pydiffvg.set_use_gpu(True)
    pydiffvg.set_device(dev)
    rng = np.random.default_rng(5)
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

    def server():
        return serve.Server(device=0, hbm_budget=hbm_budget_gib << 30, max_batch=max_batch, streams=3)

    samples = sum(w * h * (1 if sdf else nsx * nsy) for (_, w, h, nsx, nsy, sdf, _, _, _) in specs)
    # plain service rate, back to back
    cal = [make(i) for i in range(n_req)]
    serve.serve_plain(cal[:4], [0.0] * 4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    serve.serve_plain(cal, [0.0] * n_req)
    t_plain_all = time.perf_counter() - t0
    mu = n_req / t_plain_all
    progress('serving: plain service rate %.2f requests/s' % mu)
    lam = 0.8 * mu
    gaps = rng.exponential(1.0, n_req)  # unit-rate Poisson gaps, scaled per arrival rate

    def arrivals(rate):
        t = np.cumsum(gaps / rate)
        return list(t - t[0])

    server().run([make(i) for i in range(min(8, n_req))], [0.0] * min(8, n_req))  # warm-up
    torch.cuda.synchronize()
    done = server().run([make(i) for i in range(n_req)], arrivals(lam))
    st_b = serve.latency_stats(done)
    progress('serving: served p99 %.1f ms' % st_b['p99_ms'])
    plain_done = serve.serve_plain([make(i) for i in range(n_req)], arrivals(lam))
    st_p = serve.latency_stats(plain_done)
    # each request alone on an idle device (its service time without queueing)
    solo = []
    for i in range(n_req):
        r = serve.Server(device=0, hbm_budget=hbm_budget_gib << 30, max_batch=1, streams=1).run([make(i)], [0.0])[0]
        solo.append((r.done - r.arrival) * 1e3)
    slowest = sorted(done, key=lambda r: r.arrival - r.done)[:3]
    tail = [{'request': r.rid, 'latency_ms': round((r.done - r.arrival) * 1e3, 2),
             'alone_ms': round(solo[r.rid], 2), 'size': specs[r.rid][1], 'paths': specs[r.rid][8],
             'sdf': bool(specs[r.rid][5]), 'prefilter': bool(specs[r.rid][6])} for r in slowest]
    span_b = max(r.done for r in done) - min(r.arrival for r in done)
    sweep_rows, saturation = [], None
    if sweep:
        rate = 2 * lam
        while True:
            d = server().run([make(i) for i in range(n_req)], arrivals(rate))
            s_ = serve.latency_stats(d)
            span = max(r.done for r in d) - min(r.arrival for r in d)
            row = {'arrival_rps': round(rate, 2), 'p50_ms': round(s_['p50_ms'], 2), 'p99_ms': round(s_['p99_ms'], 2),
                   'throughput_rps': round(s_['throughput_rps'], 2), 'msamples_s': round(samples / span / 1e6, 2)}
            sweep_rows.append(row)
            progress('serving: sweep %s' % row)
            if s_['p99_ms'] > p99_bar_ms or rate > 64 * mu:
                break
            saturation = row
            rate *= 2
        # everything at once: the server's peak throughput
        d = server().run([make(i) for i in range(n_req)], [0.0] * n_req)
        span = max(r.done for r in d) - min(r.arrival for r in d)
        peak = {'throughput_rps': round(n_req / span, 2), 'msamples_s': round(samples / span / 1e6, 2)}
    return dict(samples=samples, n_req=n_req, lam=lam, mu=mu, span_b=span_b, st_b=st_b, st_p=st_p, solo=solo,
                tail=tail, sweep=sweep_rows, saturation=saturation, peak=peak if sweep else None,
                mean_paths=float(np.mean([s_[8] for s_ in specs])),
                sdf_fraction=float(np.mean([s_[5] for s_ in specs])),
                prefilter_fraction=float(np.mean([s_[6] and not s_[5] for s_ in specs])),
                max_size=int(max(s_[1] for s_ in specs)), max_paths=int(max(s_[8] for s_ in specs)),
                hbm_budget_gib=hbm_budget_gib, max_batch=max_batch)


def leg_single(cfg, steps, warmup, mode='tiled', prefilter_steps=0, direct_steps=0):
    """One 512x512 scene (2: 300 closed paths, Gaussian filter; 3: 2,048 strokes, box
    filter): a step is serialize_scene + RenderFunction.apply + L2 loss + backward with
    the gradients on the device.  Per-kernel times come from the library's HIP events
    over the same number of steps run again after the timed ones.  `direct_steps` more
    steps are timed with the gradients written straight into the parameters' .grad
    (set_direct_leaf_gradients)."""
    import numpy as np
    import torch
    import diffvg_amd as pydiffvg
    from diffvg_amd import workloads

    dev = torch.device('cuda:0')
    pydiffvg.set_use_gpu(True)
    pydiffvg.set_device(dev)
    pydiffvg.set_render_mode(mode)
    c = workloads.WORKLOADS[cfg](device=dev)
    w, h = c['canvas']
    nsx, nsy = c['nsx'], c['nsy']
    params = make_params(c, dev)
    if 'target' in c:
        target = c['target']
    else:
        target = torch.from_numpy(np.random.default_rng(cfg).uniform(0, 1, (h, w, 4)).astype(np.float32)).to(dev)
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

    for _ in range(warmup):
        step()
    pydiffvg.synchronize()
    torch.cuda.synchronize()
    L, read_profile = _profile_lib()
    read_profile()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    pydiffvg.synchronize()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    progress('single %s (%s): %.2f ms per step' % (c['name'], mode, sec * 1e3))
    # per-kernel times: the same steps again with the library's events around each launch
    # (kept out of the timed steps above, whose launches they would space out)
    L.dvg_profile_enable(1)
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    L.dvg_profile_enable(0)
    kern = read_profile()
    pre = None
    if prefilter_steps:
        for _ in range(2):
            step(True)
        pydiffvg.synchronize()
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        for _ in range(prefilter_steps):
            step(True)
        pydiffvg.synchronize()
        torch.cuda.synchronize()
        tp = (time.perf_counter() - tp0) / prefilter_steps
        pre = {'value': round(w * h * nsx * nsy / tp / 1e6, 4), 'ms_per_step': round(tp * 1e3, 4)}
    direct = None
    if direct_steps and mode == 'tiled':
        pydiffvg.set_direct_leaf_gradients(True)
        try:
            for _ in range(2):
                step()
            pydiffvg.synchronize()
            torch.cuda.synchronize()
            td0 = time.perf_counter()
            for _ in range(direct_steps):
                step()
            pydiffvg.synchronize()
            torch.cuda.synchronize()
            td = (time.perf_counter() - td0) / direct_steps
        finally:
            pydiffvg.set_direct_leaf_gradients(False)
        direct = {'value': round(w * h * nsx * nsy / td / 1e6, 4), 'ms_per_step': round(td * 1e3, 4),
                  'what': 'the same step with the gradients written straight into the parameters\' .grad '
                          '(set_direct_leaf_gradients)'}
        progress('single %s direct leaf gradients: %.2f ms per step' % (c['name'], td * 1e3))
    pydiffvg.set_render_mode('tiled')
    return dict(name=c['name'], w=w, h=h, nsx=nsx, nsy=nsy, paths=len(c['shapes']), samples=w * h * nsx * nsy,
                sec=sec, kernels=kern, prefilter=pre, steps=steps, warmup=warmup, mode=mode, direct=direct)


def leg_scale(steps=10, plain_steps=2):
    """Scenes of unusual scale (12,000 groups at 256x256; 3,138 groups over a 16x16
    canvas; paths of 1,500 and 600 cubic segments at 128x128): a step is serialize_scene +
    RenderFunction.apply + L2 loss + backward (gradients reset as an optimizer's
    zero_grad does), on the tiled schedule, on it with direct leaf gradients, and on
    PLAIN."""
    import torch
    import diffvg_amd as pydiffvg
    from diffvg_amd import workloads

    dev = torch.device('cuda:0')
    pydiffvg.set_use_gpu(True)
    pydiffvg.set_device(dev)
    L, read_profile = _profile_lib()
    out = {}
    for make in workloads.SCALE:
        c = make(device=dev)
        w, h = c['canvas']
        params = make_params(c, dev)
        target = torch.rand(h, w, 4, device=dev, generator=torch.Generator(device=dev).manual_seed(1))

        def step():
            for p in params:
                p.grad = None
            args = pydiffvg.RenderFunction.serialize_scene(w, h, c['shapes'], c['groups'], c['filter'])
            img = pydiffvg.RenderFunction.apply(w, h, 2, 2, 0, None, *args)
            ((img - target) ** 2).sum().backward()

        res = {'paths': len(c['shapes']), 'canvas': [w, h]}
        for label, mode, k in (('tiled', 'tiled', steps), ('tiled_direct_leaf_grads', 'tiled', steps),
                               ('plain', 'plain', plain_steps)):
            pydiffvg.set_render_mode(mode)
            pydiffvg.set_direct_leaf_gradients(label == 'tiled_direct_leaf_grads')
            for _ in range(2):
                step()
            pydiffvg.synchronize()
            torch.cuda.synchronize()
            read_profile()
            L.dvg_profile_enable(1)
            t0 = time.perf_counter()
            for _ in range(k):
                step()
            pydiffvg.synchronize()
            torch.cuda.synchronize()
            sec = (time.perf_counter() - t0) / k
            L.dvg_profile_enable(0)
            kern = read_profile()
            res[label] = {'ms_per_step': round(sec * 1e3, 3),
                          'kernel_ms_per_step': round(sum(v['total_ms'] for v in kern.values()) / k, 3)}
        pydiffvg.set_render_mode('tiled')
        pydiffvg.set_direct_leaf_gradients(False)
        res['speedup_vs_plain'] = round(res['plain']['ms_per_step'] / res['tiled']['ms_per_step'], 2)
        res['kernel_speedup_vs_plain'] = round(res['plain']['kernel_ms_per_step'] /
                                               max(res['tiled']['kernel_ms_per_step'], 1e-9), 2)
        progress('scale %s: tiled %.2f ms (kernels %.2f), direct leaf grads %.2f ms, plain %.2f ms (kernels %.2f)' % (
            c['name'], res['tiled']['ms_per_step'], res['tiled']['kernel_ms_per_step'],
            res['tiled_direct_leaf_grads']['ms_per_step'], res['plain']['ms_per_step'],
            res['plain']['kernel_ms_per_step']))
        out[c['name']] = res
    return out


def leg_generative(steps=20, warmup=5, bs=32, n=16, size=32, loop_steps=5):
    """The generative-model front end at the GAN app's defaults (16 two-segment cubic
    strokes with colours, 32x32, 2x2 samples; batch 32): a step is bezier_render + L2 loss
    + backward into the stroke tensors.  Beside it the reference's structure: one
    RenderFunction call per batch item (rendering.py:262-290), on the same GPU."""
    import random
    import torch
    import diffvg_amd as pydiffvg
    from diffvg_amd import generative

    dev = torch.device('cuda:0')
    pydiffvg.set_use_gpu(True)
    pydiffvg.set_device(dev)
    g = torch.Generator(device=dev).manual_seed(3)
    pts = (torch.rand(bs, n, 7, 2, device=dev, generator=g) * 1.6 - 0.8).requires_grad_(True)
    widths = (torch.rand(bs, n, device=dev, generator=g) * 1.5 + 0.5).requires_grad_(True)
    alphas = (torch.rand(bs, n, device=dev, generator=g) * 0.8 + 0.2).requires_grad_(True)
    colors = torch.rand(bs, n, 3, device=dev, generator=g).requires_grad_(True)
    target = torch.rand(bs, 3, size, size, device=dev, generator=g)
    random.seed(0)

    def batched():
        out, _ = generative.bezier_render(pts, widths, alphas, canvas_size=size, colors=colors)
        ((out - target) ** 2).sum().backward()

    def loop():
        cs = size
        ap = 0.5 * (pts + 1.0) * cs
        ap = ap + 1e-4 * torch.randn_like(ap)
        outs = []
        for k in range(bs):
            shapes, groups = [], []
            for p in range(n):
                shapes.append(pydiffvg.Path(num_control_points=torch.zeros(2, dtype=torch.int32) + 2,
                                            points=ap[k, p].contiguous(), stroke_width=widths[k, p],
                                            is_closed=False))
                groups.append(pydiffvg.ShapeGroup(shape_ids=torch.tensor([len(shapes) - 1]), fill_color=None,
                                                  stroke_color=torch.cat([colors[k, p], alphas[k, p].view(1)])))
            raster = generative.render(cs, cs, shapes, groups, samples=2).permute(2, 0, 1)
            outs.append(raster[:3] * raster[3:4])
        ((torch.stack(outs) - target) ** 2).sum().backward()

    def timed(fn, k, w):
        for _ in range(w):
            fn()
        pydiffvg.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        pydiffvg.synchronize()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k

    tb = timed(batched, steps, warmup)
    tl = timed(loop, loop_steps, 2)
    samples = bs * size * size * 4
    progress('generative strokes: batched %.2f ms, per-item loop %.2f ms per step' % (tb * 1e3, tl * 1e3))
    return {'workload': 'bezier_strokes_b%d_n%d_%d' % (bs, n, size), 'batch': bs, 'strokes': n,
            'canvas': [size, size], 'spp': [2, 2],
            'value': round(samples / tb / 1e6, 4), 'unit': 'Msamples/s', 'ms_per_step': round(tb * 1e3, 4),
            'per_item_loop': {'ms_per_step': round(tl * 1e3, 4), 'value': round(samples / tl / 1e6, 4),
                              'what': 'one RenderFunction fwd+bwd per batch item (the reference structure)'},
            'speedup_vs_loop': round(tl / tb, 2),
            'what': 'bezier_render (one autograd node, one forward and one backward batch) + L2 loss + backward'}


def _roofline(cpu, kern, per_kernel, workload):
    """Roofline of the dominant kernel (most total time among those with a FLOP count)."""
    if not cpu or 'counts_fwd' not in cpu or not kern:
        return None
    per_kernel_flops = {k: flops(cpu[src], adjoint=adj) for k, (src, adj) in per_kernel.items()}
    dom = max((k for k in kern if k in per_kernel_flops), key=lambda k: kern[k]['total_ms'], default=None)
    if dom is None:
        return None
    avg_s = kern[dom]['avg_ms'] / 1e3
    achieved = per_kernel_flops[dom] / avg_s / 1e12
    return {'bound': 'valu', 'achieved': round(achieved, 4), 'peak': PEAK_FP32_TFLOPS, 'unit': 'TFLOP/s',
            'frac': round(achieved / PEAK_FP32_TFLOPS, 6), 'traffic': load_traffic(dom, workload),
            'kernel': dom, 'kernel_avg_ms': round(kern[dom]['avg_ms'], 4), 'kernel_launches': kern[dom]['launches'],
            'flops_per_launch': per_kernel_flops[dom],
            'note': 'vector FP32 peak (no MFMA on this path); FLOPs are FP32-equivalent algorithmic counts of the '
                    'reference algorithm (oracle traversal counts x cost table); kernel time from HIP events on the '
                    'launch stream'}


SINGLE_KERNELS = {'k_color': ('counts_fwd', False), 'kp_render': ('counts_fwd', False),
                  'k_color_bwd': ('counts_bwd', True), 'k_edge': ('counts_edge', False),
                  'kp_edge': ('counts_edge', False)}
BATCH_KERNELS = {'kb_color': ('counts_fwd', False), 'kb_color_bwd': ('counts_bwd', True),
                 'kb_edge': ('counts_edge', False)}


def _cpu_baseline(cpu):
    if not cpu or 'fwd_s' not in cpu:
        return None
    return {'value': round(cpu['samples'] / (cpu['fwd_s'] + cpu['bwd_s']) / 1e6, 4), 'unit': 'Msamples/s',
            'cores': cpu['cores'], 'kind': cpu['kind'], 'sample': cpu['sample'], 'cpu_model': cpu.get('cpu_model')}


def _kernels_out(kern):
    return {k: {'avg_ms': round(v['avg_ms'], 4), 'launches': v['launches']} for k, v in kern.items()}


def _batched_fields(b):
    if b['t_batched'] is None or b['t_plain'] is None:  # the graph replay alone (profiling runs)
        return {'kernels_ms_eager': {k: round(v['avg_ms'], 4) for k, v in b['kernels'].items()},
                'kernel_ms_per_step': round(sum(v['total_ms'] for v in b['kernels'].values()) / b['steps'], 4),
                'autotune': b['tuned']}
    return {
        'batched_end_to_end': {'ms_per_step': round(b['t_batched'] * 1e3, 3),
                               'value': round(b['samples'] / b['t_batched'] / 1e6, 4), 'steps': b['e2e_steps'],
                               'what': 'serialize_scene + render_batch (pooled device rebuilds, one batched fwd '
                                       'and bwd) + L2 loss + backward into every parameter tensor'},
        'plain': {'ms_per_step': round(b['t_plain'] * 1e3, 3), 'value': round(b['samples'] / b['t_plain'] / 1e6, 4),
                  'steps': b['plain_steps'],
                  'what': 'one RenderFunction fwd+bwd per scene on the PLAIN schedule (reference launch structure)'},
        'speedup_graph_vs_plain': round(b['t_plain'] / b['t_graph'], 2),
        'speedup_batched_vs_plain': round(b['t_plain'] / b['t_batched'], 2),
        'batched_end_to_end_direct_leaf_grads': None if b.get('t_direct') is None else {
            'ms_per_step': round(b['t_direct'] * 1e3, 3), 'value': round(b['samples'] / b['t_direct'] / 1e6, 4),
            'speedup_vs_plain': round(b['t_plain'] / b['t_direct'], 2),
            'what': 'the same step with set_direct_leaf_gradients(True): the backward writes each parameter '
                    'tensor\'s .grad itself'},
        'kernels_ms_eager': {k: round(v['avg_ms'], 4) for k, v in b['kernels'].items()},
        'kernel_ms_per_step': round(sum(v['total_ms'] for v in b['kernels'].values()) / b['steps'], 4),
        'autotune': b['tuned'],
    }


def _serving_fields(s):
    import numpy as np
    r3 = lambda d: {k: round(v, 3) for k, v in d.items()}
    return {
        'arrival_rate_rps': round(s['lam'], 3), 'plain_service_rate_rps': round(s['mu'], 3),
        'requests': s['n_req'], 'mean_paths': round(s['mean_paths'], 1), 'max_paths': s['max_paths'],
        'max_size': s['max_size'], 'sdf_fraction': round(s['sdf_fraction'], 3),
        'prefilter_fraction': round(s['prefilter_fraction'], 3),
        'served': r3(s['st_b']), 'plain': r3(s['st_p']),
        'served_msamples_s': round(s['samples'] / s['span_b'] / 1e6, 4),
        'alone': {'p50_ms': round(float(np.percentile(s['solo'], 50)), 3),
                  'p99_ms': round(float(np.percentile(s['solo'], 99)), 3), 'max_ms': round(max(s['solo']), 3)},
        'slowest_served': s['tail'],
        'sweep': s['sweep'], 'saturation_under_p99_50ms': s['saturation'], 'peak_all_at_once': s['peak'],
        'hbm_budget_gib': s['hbm_budget_gib'], 'max_batch': s['max_batch'],
    }


def _single_fields(t, p):
    out = {'value': round(t['samples'] / t['sec'] / 1e6, 4), 'ms_per_step': round(t['sec'] * 1e3, 4),
           'kernel_ms_per_step': round(sum(v['total_ms'] for v in t['kernels'].values()) / t['steps'], 4),
           'kernels': _kernels_out(t['kernels']), 'paths': t['paths'], 'canvas': [t['w'], t['h']],
           'spp': [t['nsx'], t['nsy']]}
    if t['prefilter']:
        out['prefilter_on'] = t['prefilter']
    if t.get('direct'):
        out['direct_leaf_grads'] = t['direct']
    if p is not None:
        out['plain'] = {'value': round(p['samples'] / p['sec'] / 1e6, 4), 'ms_per_step': round(p['sec'] * 1e3, 4),
                        'kernel_ms_per_step': round(sum(v['total_ms'] for v in p['kernels'].values()) / p['steps'], 4)}
        out['speedup_vs_plain'] = round(p['sec'] / t['sec'], 2)
        if t.get('direct'):
            out['direct_leaf_grads']['speedup_vs_plain'] = round(p['sec'] * 1e3 / t['direct']['ms_per_step'], 2)
    return out


def _between_legs():
    """Drops the pooled scenes an earlier leg left behind (scenes of another workload) and
    collects the garbage, so that each leg's host side is timed as when it runs alone."""
    import gc
    from diffvg_amd import render_pytorch
    render_pytorch.clear_scene_cache()
    gc.collect()


def run_headline(a):
    """The default line: the 256-scene batch (value), the serving stream and both 512x512
    single scenes."""
    cpu = None if a.no_cpu_baseline else run_cpu_baseline(4)
    b = leg_batched(a.steps, a.warmup, autotune=not a.no_autotune)
    _between_legs()
    s = leg_serving(a.requests, a.hbm_budget, a.max_batch)
    singles = {}
    for cfg, (st, wu, pst) in ((2, (30, 10, 8)), (3, (12, 4, 0))):
        _between_legs()
        t = leg_single(cfg, st, wu, 'tiled', prefilter_steps=pst, direct_steps=st)
        p = leg_single(cfg, 3, 1, 'plain')
        singles[t['name']] = _single_fields(t, p)
    out = {
        'metric': METRIC, 'value': round(b['samples'] / b['t_graph'] / 1e6, 4), 'unit': 'Msamples/s', 'n_gpus': 1,
        'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': round(b['t_graph'] * 1e3, 4), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': DTYPE, 'data': 'synthetic',
        'config': {'workload': 'uneven_scenes_256', 'scenes': b['scenes'], 'samples': b['samples'],
                   'paths': b['paths'], 'spp': [2, 2], 'filter': 'box r=0.5',
                   'mode': 'batched: forward batch + loss gradient + backward batch replayed from one HIP graph, '
                           'scenes resident in HBM',
                   'schedule': (b['tuned'] or {}).get('schedule', 'staged')},
    }
    out.update(_batched_fields(b))
    out['serving_stream'] = _serving_fields(s)
    out['single_scene'] = singles
    _between_legs()
    out['generative_strokes'] = leg_generative()
    _between_legs()
    out['scale_scenes'] = leg_scale()
    out['roofline'] = _roofline(cpu, b['kernels'], BATCH_KERNELS, 'uneven_scenes_256')
    out['cpu_baseline'] = _cpu_baseline(cpu)
    print(json.dumps(out))


def run_config(a):
    """One workload per line (profiling runs): 2, 3 single scene; 4 batch; 5 serving."""
    base = {'metric': METRIC, 'unit': 'Msamples/s', 'n_gpus': 1, 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': DTYPE, 'data': 'synthetic'}
    if a.config == 4:
        cpu = None if a.no_cpu_baseline else run_cpu_baseline(4)
        b = leg_batched(a.steps, a.warmup, autotune=not a.no_autotune, e2e_steps=0 if a.graph_only else 4,
                        plain_steps=0 if a.graph_only else 1)
        out = dict(base, value=round(b['samples'] / b['t_graph'] / 1e6, 4), steps=a.steps, warmup=a.warmup,
                   ms_per_step=round(b['t_graph'] * 1e3, 4),
                   config={'workload': 'uneven_scenes_256', 'scenes': b['scenes'], 'samples': b['samples'],
                           'paths': b['paths'], 'spp': [2, 2], 'filter': 'box r=0.5',
                           'mode': 'batched, one HIP graph per fwd+bwd',
                           'schedule': (b['tuned'] or {}).get('schedule', 'staged')})
        out.update(_batched_fields(b))
        out['roofline'] = _roofline(cpu, b['kernels'], BATCH_KERNELS, 'uneven_scenes_256')
        out['cpu_baseline'] = _cpu_baseline(cpu)
    elif a.config == 'scale':
        r = leg_scale()
        first = next(iter(r.values()))
        out = dict(base, value=None, steps=10, warmup=2, ms_per_step=first['tiled']['ms_per_step'],
                   config={'workload': 'scale_scenes', 'scenes': list(r)})
        out['scale_scenes'] = r
    elif a.config == 'gen':
        gst = a.steps if a.steps_given else 20
        gw = a.warmup if a.warmup_given else 5
        r = leg_generative(gst, gw)
        out = dict(base, value=r['value'], steps=gst, warmup=gw, ms_per_step=r['ms_per_step'],
                   config={'workload': r['workload'], 'batch': r['batch'], 'strokes': r['strokes'],
                           'canvas': r['canvas'], 'spp': r['spp']})
        out['generative_strokes'] = r
    elif a.config == 5:
        s = leg_serving(a.requests, a.hbm_budget, a.max_batch, sweep=not a.no_sweep)
        out = dict(base, value=round(s['samples'] / s['span_b'] / 1e6, 4), steps=s['n_req'], warmup=8,
                   ms_per_step=round(s['span_b'] / s['n_req'] * 1e3, 4),
                   config={'workload': 'serving_stream', 'requests': s['n_req'], 'samples': s['samples']})
        out['serving_stream'] = _serving_fields(s)
    else:
        cpu = None if a.no_cpu_baseline else run_cpu_baseline(a.config)
        t = leg_single(a.config, a.steps, a.warmup, a.mode, prefilter_steps=8 if a.config == 2 else 0,
                       direct_steps=a.steps)
        f = _single_fields(t, None)
        out = dict(base, value=f['value'], steps=a.steps, warmup=a.warmup, ms_per_step=f['ms_per_step'],
                   config={'workload': t['name'], 'canvas': [t['w'], t['h']], 'spp': [t['nsx'], t['nsy']],
                           'filter': 'gaussian r=1.5' if a.config == 2 else 'box r=0.5', 'paths': t['paths'],
                           'mode': a.mode, 'prefilter': False})
        out.update({k: v for k, v in f.items() if k not in ('value', 'ms_per_step')})
        out['roofline'] = _roofline(cpu, t['kernels'], SINGLE_KERNELS, t['name'])
        out['cpu_baseline'] = _cpu_baseline(cpu)
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
        # rows name the kernel as rocprof does (template arguments included); the first
        # instantiation listed stands for the bench's kernel name
        if len(cells) >= 3 and cells[0].split('<')[0] == kernel and cells[1] == workload:
            try:
                return float(cells[-1])
            except ValueError:
                return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None,
                    help='timed steps of the line\'s workload (default 20 graph replays of the 256-scene batch; '
                         '50 steps of a single scene)')
    ap.add_argument('--warmup', type=int, default=None, help='untimed steps first (default 5; 20 for a single scene)')
    ap.add_argument('--config', default='headline', choices=['headline', '2', '3', '4', '5', 'gen', 'scale'],
                    help='headline (default): the 256-scene batch with the serving stream and the single scenes '
                         'as extra keys; 2/3: one 512x512 scene; 4: the batch alone; 5: the serving stream alone; '
                         'gen: the generative stroke batch; scale: scenes of unusual scale')
    ap.add_argument('--mode', default='tiled', choices=['tiled', 'plain'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-baseline-only', action='store_true')
    ap.add_argument('--copy-calibration', action='store_true', help='run a plain copy kernel of known size')
    ap.add_argument('--requests', type=int, default=120, help='serving stream length')
    ap.add_argument('--no-sweep', action='store_true', help='serving: skip the arrival-rate sweep')
    ap.add_argument('--no-autotune', action='store_true',
                    help='batch: keep the default tile shape, samples per wave and launch structure instead of '
                         'timing the candidates per workload signature before the timed replays')
    ap.add_argument('--graph-only', action='store_true',
                    help='config 4: time the graph replay only (no end-to-end or PLAIN legs; for profiling runs)')
    ap.add_argument('--hbm-budget', type=int, default=64, help='GiB that requests in flight may hold (serving)')
    ap.add_argument('--max-batch', type=int, default=32, help='requests per batch (serving)')
    a = ap.parse_args()
    single = a.config in ('2', '3')
    a.steps_given, a.warmup_given = a.steps is not None, a.warmup is not None
    if a.steps is None:
        a.steps = 50 if single else 20
    if a.warmup is None:
        a.warmup = 20 if single else 5
    if a.cpu_baseline_only:
        cpu_baseline_main(4 if a.config == 'headline' else int(a.config))
        return
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print('bench.py measures one GPU: the render path is built and measured on one MI355X, so --gpus %d '
              'is not run (how it would shard is a design note, not a measured path)' % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.config == 'headline':
        run_headline(a)
    else:
        a.config = a.config if a.config in ('gen', 'scale') else int(a.config)
        run_config(a)


if __name__ == '__main__':
    main()
