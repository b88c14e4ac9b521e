"""Batched rendering of many scenes (the C ABI's dvg_batch_*).

A `Batch` holds a list of jobs, each a `DeviceScene` plus the keyword arguments of
`DeviceScene.render` (sizes, seed, output / gradient buffers, prefiltering, eval
positions).  `render()` enqueues one kernel per stage over the work items of all jobs:
nothing is allocated and nothing waits, so a render can be captured into a CUDA/HIP
graph (`torch.cuda.graph`) and replayed.  Every tensor named in a job must stay alive
and in place while the batch exists; the batch keeps references to them.

The reference renders such workloads one scene per call
(apps/generative_models/rendering.py:262-290 loops over items calling
RenderFunction.apply); here each job has the semantics of one `render(...)` call.
"""
import ctypes

from . import _lib

ZERO_GRADS = 1
PERSISTENT = 2


class Batch:
    """`persistent`: the whole render is one kernel launch whose workgroups take the
    batch's work items from a device queue (DVG_BATCH_PERSISTENT), instead of one launch
    per stage."""

    def __init__(self, jobs, zero_grads=True, persistent=False):
        L = _lib.lib()
        self._lib = L
        n = len(jobs)
        if n == 0:
            raise ValueError('a batch needs at least one job')
        scenes = (ctypes.c_void_p * n)()
        args = (_lib.RenderArgs * n)()
        keep = []
        for k, (scene, kw) in enumerate(jobs):
            scenes[k] = scene.handle
            args[k] = scene.make_args(**kw)
            keep.append(scene)
            keep.extend(v for v in kw.values() if hasattr(v, 'data_ptr'))
        self._keep = keep
        h = ctypes.c_void_p()
        flags = (ZERO_GRADS if zero_grads else 0) | (PERSISTENT if persistent else 0)
        _lib.check(L.dvg_batch_create(scenes, args, n, flags, ctypes.byref(h)))
        self.handle = h
        self.n = n
        self.persistent = bool(persistent)

    def render(self, stream=None):
        """Enqueues the batch on `stream` (a raw HIP stream handle; None = null stream)."""
        _lib.check(self._lib.dvg_batch_render(self.handle, stream))

    def status(self):
        """Waits for the last render; raises on the conditions the reference asserts on."""
        _lib.check(self._lib.dvg_batch_status(self.handle))

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h and _lib._lib is not None:
            _lib._lib.dvg_batch_destroy(h)
            self.handle = None
