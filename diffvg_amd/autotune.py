"""Autotuner for the tiled schedule's pixel tile shape.

A tile holds 256 samples; its width (a power of two) sets its aspect and how the four
wavefronts of a workgroup cover it (pixel rows in order).  Wider tiles keep a wave's
samples in fewer rows, narrower ones in fewer columns; which candidate lists are
shortest and which waves diverge least depends on the scene.  `Tuner.tune` times one
forward + backward per candidate width on the caller's scene (HIP events on the current
stream) and remembers the fastest per workload signature -- render size bucket, samples
per pixel, group-count bucket, prefiltering -- so later scenes of the same kind reuse it
without timing.  The choice never changes results beyond float summation order.
"""
import math

import torch

from . import _lib


def signature(width, height, nsx, nsy, num_groups, prefilter):
    def bucket(v):
        return 1 << max(0, int(math.ceil(math.log2(max(v, 1)))))
    return (bucket(width), bucket(height), nsx * nsy, bucket(num_groups), bool(prefilter))


class Tuner:
    def __init__(self, reps=2):
        self.best = {}
        self.reps = reps
        self.timings = {}

    def candidates(self, spp):
        px = max(1, 256 // spp)
        out, tw = [], 1
        while tw <= px:
            if px // tw >= 1:
                out.append(tw)
            tw *= 2
        # drop the extreme slivers
        return [t for t in out if px // t >= 2 or t == out[-1]][-6:]

    def tune(self, scene, width, height, nsx, nsy, seed=0, prefilter=False, num_groups=None):
        """Times each candidate tile width on `scene` (one forward + backward each) and
        sets the fastest; returns it."""
        ng = num_groups if num_groups is not None else scene.flat['num_groups']
        key = signature(width, height, nsx, nsy, ng, prefilter)
        if key in self.best:
            scene.set_tiling(self.best[key])
            return self.best[key]
        dev = torch.device('cuda', torch.cuda.current_device())
        img = torch.zeros(height, width, 4, device=dev)
        d_img = torch.rand(height, width, 4, device=dev) - 0.5
        st = torch.cuda.current_stream().cuda_stream
        res = {}
        for tw in self.candidates(nsx * nsy):
            scene.set_tiling(tw)
            times = []
            for r in range(self.reps + 1):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                img.zero_()
                scene.render(width, height, nsx, nsy, seed, render_image=img, use_prefiltering=prefilter,
                             stream=st)
                scene.zero_grads(st)
                scene.render(width, height, nsx, nsy, seed, d_render_image=d_img, use_prefiltering=prefilter,
                             stream=st)
                e1.record()
                e1.synchronize()
                if r > 0:
                    times.append(e0.elapsed_time(e1))
            res[tw] = min(times)
        best = min(res, key=res.get)
        self.best[key] = best
        self.timings[key] = res
        scene.set_tiling(best)
        return best

    def apply(self, scene, width, height, nsx, nsy, prefilter=False):
        """Sets the remembered tile width for this workload signature, if any."""
        key = signature(width, height, nsx, nsy, scene.flat['num_groups'], prefilter)
        if key in self.best:
            scene.set_tiling(self.best[key])
            return self.best[key]
        return None
