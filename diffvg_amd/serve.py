"""Continuous-batching render server.

Scenes are registered once and stay resident on the device (their BVHs, sampling tables,
parameters and gradient blocks).  Requests -- render + gradient of one resident scene at
a given size, sample count, output (colour or SDF) and prefiltering, with an L2 loss
against a target -- arrive over time.  The server admits waiting requests whenever their
memory fits the HBM budget and a slot is free, launches each admitted group as one
forward batch, the loss gradient and one backward batch on one of several streams, and
retires batches as their completion events fire.  Requests that arrive while others are
on the GPU are admitted into the next batch on another stream, so their work overlaps
with the batches in flight.

The reference has no serving component: each of its requests would be one
RenderFunction forward + backward (render_pytorch.py:174-868) issued one at a time,
which is what `serve_plain` measures for comparison.
"""
import collections
import time

import numpy as np
import torch

from . import _lib
from .batch import Batch

EDGE_SAMPLE_BYTES = 64


def request_bytes(w, h, spp, sdf, n_eval=0):
    """Device bytes a request holds while in flight (outputs, loss gradient, workspace)."""
    npix = w * h
    if sdf:
        n = n_eval if n_eval else npix
        return 3 * 4 * n + 2 * 4 * npix
    nsamp = npix * spp
    return npix * 4 * 4 * 3 + npix * 4 + nsamp * 16 + nsamp * (EDGE_SAMPLE_BYTES + 8) + 4096


class Request:
    __slots__ = ('rid', 'scene', 'w', 'h', 'nsx', 'nsy', 'seed', 'sdf', 'prefilter', 'target', 'arrival',
                 'admitted', 'done', 'bytes', 'out', 'grad', 'loss_grad', 'error')

    def __init__(self, rid, scene, w, h, nsx, nsy, seed, sdf, prefilter, target, arrival):
        self.rid, self.scene, self.w, self.h = rid, scene, w, h
        self.nsx, self.nsy, self.seed, self.sdf, self.prefilter = nsx, nsy, seed, sdf, prefilter
        self.target, self.arrival = target, arrival
        self.admitted = self.done = None
        self.bytes = request_bytes(w, h, nsx * nsy, sdf)
        self.out = self.grad = self.loss_grad = None
        self.error = None


class _InFlight:
    def __init__(self, reqs, fb, bb, event, stream, grads):
        self.reqs, self.fb, self.bb, self.event, self.stream, self.grads = reqs, fb, bb, event, stream, grads


class Server:
    """hbm_budget: bytes that requests in flight may hold; max_batch: requests per batch;
    streams: batches that may be in flight at once."""

    def __init__(self, device=0, hbm_budget=8 << 30, max_batch=64, streams=3, copy_grads=True,
                 stream_factory=None):
        self.dev = torch.device('cuda', device)
        self.budget = int(hbm_budget)
        self.max_batch = max_batch
        make = stream_factory or (lambda: torch.cuda.Stream(device=self.dev))
        self.streams = [make() for _ in range(streams)]
        self.free_streams = list(self.streams)
        self.in_use = 0
        self.waiting = collections.deque()
        self.flight = []
        self.done = []
        self.copy_grads = copy_grads

    # -- admission --------------------------------------------------------------
    def submit(self, req):
        if req.bytes > self.budget:
            raise ValueError('request %d needs %d bytes, more than the whole budget' % (req.rid, req.bytes))
        self.waiting.append(req)

    def _admit(self, now):
        if not self.free_streams or not self.waiting:
            return
        # admit in arrival order while the budget allows; a request whose scene is busy
        # (in flight, or already in this group) waits for the next batch, because the
        # scene's gradient block is shared by its renders
        group, skipped = [], []
        used = self.in_use
        busy = {id(r.scene) for f in self.flight for r in f.reqs}
        while self.waiting and len(group) < self.max_batch:
            r = self.waiting[0]
            if used + r.bytes > self.budget:
                break
            self.waiting.popleft()
            if id(r.scene) in busy:
                skipped.append(r)
                continue
            busy.add(id(r.scene))
            used += r.bytes
            group.append(r)
        self.waiting.extendleft(reversed(skipped))
        if not group:
            return
        self.in_use = used
        stream = self.free_streams.pop()
        for r in group:
            r.admitted = now
        self._launch(group, stream)

    def _launch(self, group, stream):
        dev = self.dev
        fj, bj = [], []
        with torch.cuda.stream(stream):
            for r in group:
                if r.sdf:
                    r.out = torch.zeros(r.h, r.w, 1, device=dev)
                    r.loss_grad = torch.empty(r.h, r.w, 1, device=dev)
                    fj.append((r.scene, dict(width=r.w, height=r.h, nsx=1, nsy=1, seed=r.seed, render_sdf=r.out)))
                    bj.append((r.scene, dict(width=r.w, height=r.h, nsx=1, nsy=1, seed=r.seed,
                                             d_render_sdf=r.loss_grad)))
                else:
                    r.out = torch.zeros(r.h, r.w, 4, device=dev)
                    r.loss_grad = torch.empty(r.h, r.w, 4, device=dev)
                    fj.append((r.scene, dict(width=r.w, height=r.h, nsx=r.nsx, nsy=r.nsy, seed=r.seed,
                                             render_image=r.out, use_prefiltering=r.prefilter)))
                    bj.append((r.scene, dict(width=r.w, height=r.h, nsx=r.nsx, nsy=r.nsy, seed=r.seed,
                                             d_render_image=r.loss_grad, use_prefiltering=r.prefilter)))
        fb = Batch(fj, zero_grads=False)
        bb = Batch(bj, zero_grads=True)
        st = stream.cuda_stream
        with torch.cuda.stream(stream):
            fb.render(st)
            for r in group:
                torch.mul(r.out - r.target, 2.0, out=r.loss_grad)
            bb.render(st)
            grads = None
            if self.copy_grads:
                grads = [r.scene.grads_device(dev, st)[0] for r in group]
            ev = torch.cuda.Event()
            ev.record(stream)
        self.flight.append(_InFlight(group, fb, bb, ev, stream, grads))

    # -- retirement -------------------------------------------------------------
    def _retire(self, now_fn):
        still = []
        for f in self.flight:
            if f.event.query():
                t = now_fn()
                err = None
                try:
                    f.fb.status()
                    f.bb.status()
                except RuntimeError as e:  # a condition the reference asserts on
                    err = str(e)
                for k, r in enumerate(f.reqs):
                    r.error = err
                    r.done = t
                    if f.grads is not None:
                        r.grad = f.grads[k]
                    self.in_use -= r.bytes
                    self.done.append(r)
                self.free_streams.append(f.stream)
            else:
                still.append(f)
        self.flight = still

    def run(self, requests, arrivals, clock=time.perf_counter, poll=0.0002):
        """Serves `requests` arriving at offsets `arrivals` (seconds, ascending) from now;
        returns the completed requests with arrival / done times on one clock."""
        t0 = clock()
        k = 0
        n = len(requests)
        while len(self.done) < n:
            now = clock() - t0
            while k < n and arrivals[k] <= now:
                requests[k].arrival = arrivals[k]
                self.submit(requests[k])
                k += 1
            self._retire(lambda: clock() - t0)
            self._admit(clock() - t0)
            if not self.flight and not self.waiting and k < n:
                time.sleep(max(0.0, min(poll * 10, arrivals[k] - (clock() - t0))))
            else:
                time.sleep(poll)
        return self.done


def serve_plain(requests, arrivals, clock=time.perf_counter):
    """One request at a time in arrival order, each as a synchronous single-scene forward
    and backward on the PLAIN schedule (the reference's structure)."""
    t0 = clock()
    dev = None
    for req, a in zip(requests, arrivals):
        now = clock() - t0
        if now < a:
            time.sleep(a - now)
        req.arrival = a
        req.admitted = clock() - t0
        sc = req.scene
        dev = dev or torch.device('cuda', 0)
        sc.set_mode(_lib.MODE_PLAIN)
        try:
            _plain_one(req, sc, dev)
        except RuntimeError as e:  # a condition the reference asserts on
            req.error = str(e)
        torch.cuda.synchronize()
        sc.set_mode(_lib.MODE_TILED)
        req.done = clock() - t0
    return requests


def _plain_one(req, sc, dev):
    if req.sdf:
        out = torch.zeros(req.h, req.w, 1, device=dev)
        sc.render(req.w, req.h, 1, 1, req.seed, render_sdf=out)
        d = 2.0 * (out - req.target)
        sc.zero_grads(None)
        sc.render(req.w, req.h, 1, 1, req.seed, d_render_sdf=d)
    else:
        out = torch.zeros(req.h, req.w, 4, device=dev)
        sc.render(req.w, req.h, req.nsx, req.nsy, req.seed, render_image=out, use_prefiltering=req.prefilter)
        d = 2.0 * (out - req.target)
        sc.zero_grads(None)
        sc.render(req.w, req.h, req.nsx, req.nsy, req.seed, d_render_image=d, use_prefiltering=req.prefilter)
    req.grad = sc.grads_device(dev)[0]
    req.out = out


def latency_stats(reqs):
    failed = sum(1 for r in reqs if r.error)
    lat = np.array([r.done - r.arrival for r in reqs]) * 1e3
    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    return dict(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()), requests=len(reqs), failed=failed,
                throughput_rps=len(reqs) / span if span > 0 else float('inf'))
