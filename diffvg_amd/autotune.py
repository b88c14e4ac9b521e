"""Autotuner of the tiled schedule: tile shape, samples per wave, batch launch structure.

Per scene (`Tuner.tune`), two knobs that never change results beyond float summation
order:

* the pixel tile width (a power of two).  A tile holds 256 samples; its width sets its
  aspect and how the four wavefronts of a workgroup cover it (pixel rows in order).
  Wider tiles keep a wave's samples in fewer rows, narrower ones in fewer columns; which
  candidate lists are shortest and which waves diverge least depends on the scene;
* the boundary samples per wave of the edge kernel: 32 (each lane shades one of the
  two colour queries of a sample, so a lane keeps one compositor live) or 64 (each lane
  shades both, with twice the state but half the lanes).

`tune` times one forward + backward per (width, samples per wave) on the caller's scene
(HIP events on the current stream) and remembers the fastest per workload signature --
render size bucket, samples per pixel, group-count bucket, prefiltering -- so later
scenes of the same kind reuse it without timing.

Per batch (`Tuner.tune_batch`): one launch per stage or one persistent launch
(`Batch(persistent=...)`), timed on the caller's batches; a structure the library
refuses (the persistent launch is disabled in this build) is skipped.
"""
import math

import torch

from . import _lib

EDGE_SAMPLES_PER_WAVE = (32, 64)


def signature(width, height, nsx, nsy, num_groups, prefilter):
    def bucket(v):
        return 1 << max(0, int(math.ceil(math.log2(max(v, 1)))))
    return (bucket(width), bucket(height), nsx * nsy, bucket(num_groups), bool(prefilter))


def _time_ms(fn, reps):
    """Smallest of `reps` timed runs of fn() after one untimed run (HIP events on the
    current stream)."""
    times = []
    for r in range(reps + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        if r > 0:
            times.append(e0.elapsed_time(e1))
    return min(times)


class Tuner:
    def __init__(self, reps=2):
        self.best = {}
        self.reps = reps
        self.timings = {}
        self.batch_timings = {}

    def candidates(self, spp, prefilter=False):
        """(tile width, edge samples per wave) pairs to time.  Prefiltered renders have no
        edge pass, so only the width varies for them."""
        px = max(1, 256 // spp)
        out, tw = [], 1
        while tw <= px:
            if px // tw >= 1:
                out.append(tw)
            tw *= 2
        # drop the extreme slivers
        widths = [t for t in out if px // t >= 2 or t == out[-1]][-6:]
        spws = (32,) if prefilter else EDGE_SAMPLES_PER_WAVE
        return [(tw, spw) for tw in widths for spw in spws]

    @staticmethod
    def _set(scene, choice):
        tw, spw = choice
        scene.set_tiling(tw)
        scene.set_samples_per_wave(spw)

    def tune(self, scene, width, height, nsx, nsy, seed=0, prefilter=False, num_groups=None):
        """Times each candidate on `scene` (one forward + backward each), sets the fastest
        and returns it as (tile width, edge samples per wave)."""
        ng = num_groups if num_groups is not None else scene.flat['num_groups']
        key = signature(width, height, nsx, nsy, ng, prefilter)
        if key in self.best:
            self._set(scene, self.best[key])
            return self.best[key]
        dev = torch.device('cuda', torch.cuda.current_device())
        img = torch.zeros(height, width, 4, device=dev)
        d_img = torch.rand(height, width, 4, device=dev) - 0.5
        st = torch.cuda.current_stream().cuda_stream

        def step():
            img.zero_()
            scene.render(width, height, nsx, nsy, seed, render_image=img, use_prefiltering=prefilter, stream=st)
            scene.zero_grads(st)
            scene.render(width, height, nsx, nsy, seed, d_render_image=d_img, use_prefiltering=prefilter, stream=st)

        res = {}
        for choice in self.candidates(nsx * nsy, prefilter):
            self._set(scene, choice)
            res[choice] = _time_ms(step, self.reps)
        best = min(res, key=res.get)
        self.best[key] = best
        self.timings[key] = res
        self._set(scene, best)
        return best

    def apply(self, scene, width, height, nsx, nsy, prefilter=False):
        """Sets the remembered choice for this workload signature, if any."""
        key = signature(width, height, nsx, nsy, scene.flat['num_groups'], prefilter)
        if key in self.best:
            self._set(scene, self.best[key])
            return self.best[key]
        return None

    def tune_batch(self, make_batches, run, reps=None):
        """Launch structure of a batched workload.  make_batches(persistent) returns the
        batches of one step built with that structure; run(batches) enqueues one step.
        Returns (persistent, batches) for the faster structure; the timings (ms per step)
        are kept in `batch_timings`."""
        reps = self.reps if reps is None else reps
        res, built = {}, {}
        for persistent in (False, True):
            try:
                b = make_batches(persistent)
            except RuntimeError:  # a structure this build refuses is not a candidate
                continue
            built[persistent] = b
            res[persistent] = _time_ms(lambda: run(b), reps)
            for x in b:
                x.status()
        best = min(res, key=res.get)
        self.batch_timings = {'staged_ms': res.get(False), 'persistent_ms': res.get(True)}
        return best, built[best]
