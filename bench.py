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
    c = workloads.config2(device='cpu', filter_type='tent') if cfg == 2 else workloads.CONFIGS[cfg](device='cpu')
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--config', type=int, default=2)
    ap.add_argument('--mode', default='tiled', choices=['tiled', 'plain'])
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-baseline-only', action='store_true')
    a = ap.parse_args()
    if a.cpu_baseline_only:
        cpu_baseline_main(a.config)
        return
    if a.gpus != 1:
        print('bench.py measures one GPU: every box in this setup has one MI355X, so --gpus %d is not run '
              '(the multi-GPU sharding design is described in DESIGN.md)' % a.gpus, file=sys.stderr)
        sys.exit(2)

    cpu = None if a.no_cpu_baseline else run_cpu_baseline(a.config)

    import numpy as np
    import torch
    import diffvg_amd as pydiffvg
    from diffvg_amd import workloads, _lib

    dev = torch.device('cuda:0')
    pydiffvg.set_use_gpu(True)
    pydiffvg.set_device(dev)
    pydiffvg.set_render_mode(a.mode)
    c = workloads.CONFIGS[a.config](device=dev)
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
    L.dvg_profile_read(names, 4096, ms, cnts, 32)
    L.dvg_profile_enable(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
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
                        'unit': 'TFLOP/s', 'frac': round(achieved / PEAK_FP32_TFLOPS, 6), 'traffic': None,
                        'kernel': dom, 'kernel_avg_ms': round(kern[dom]['avg_ms'], 4),
                        'flops_per_launch': per_kernel_flops[dom],
                        'note': 'vector FP32 peak; FLOPs are FP32-equivalent algorithmic counts'}
    cpu_baseline = None
    if cpu and 'fwd_s' in cpu:
        cpu_sec = cpu['fwd_s'] + cpu['bwd_s']
        cpu_baseline = {'value': round(cpu['samples'] / cpu_sec / 1e6, 4), 'unit': 'Msamples/s',
                        'cores': cpu['cores'], 'kind': cpu['kind'],
                        'sample': 'the full %dx%d %dx%d-spp scene, one forward and one backward (edge sampling), '
                                  'tent r=1 twin of the Gaussian filter' % (w, h, nsx, nsy)}
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
