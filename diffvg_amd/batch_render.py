"""Differentiable rendering of many scenes in one launch per stage.

`BatchRenderFunction` is the batched counterpart of `RenderFunction`: it takes the
serialize_scene argument lists of several scenes (any sizes, shape counts, sample
counts) and renders all of them with one dvg_batch forward and one dvg_batch backward,
where the reference renders such workloads with one RenderFunction call per scene
(apps/generative_models/rendering.py:262-290 loops over the batch).  Each scene keeps
the semantics of `RenderFunction.apply(width, height, nsx, nsy, seed, None, *args)`:
same image, same gradients for its argument tensors.

Scenes of a topology seen before are rebuilt on the device from their new values
(dvg_scene_update) in a pool, so a training loop over same-shaped scenes builds nothing
on the host after its first step.
"""
import torch

from . import _lib
from . import render_pytorch as _rp
from .batch import Batch
from .device import get_device
from .render_pytorch import (_KINDS, OutputType, _PENDING, _Lease, _Resized, _batch_params, _check_gpu,
                             _device_scene_for, _find_template, _pooled_scene, _stream, check_rebuild, synchronize)

__all__ = ['BatchRenderFunction', 'render_batch']

_HINTS = []


class BatchRenderFunction(torch.autograd.Function):
    """apply(spec, *args): spec is a tuple of (width, height, num_samples_x,
    num_samples_y, seed, num_args) per scene; args are the scenes' serialize_scene lists
    concatenated.  Returns one image tensor per scene.

    As RenderFunction.apply, the autograd node is taken over each scene's float arrays
    by kind and its filter radius, gathered from the argument tensors by differentiable
    concatenation, not over every argument tensor; with set_direct_leaf_gradients(True)
    a scene whose argument tensors allow it has its gradients written into their .grad
    by the backward instead (the node then holds a zero-size anchor leaf)."""

    @classmethod
    def apply(cls, spec, *args):
        _check_gpu(read_findings=False)
        dev = get_device()
        grad = torch.is_grad_enabled()
        scenes, flats, inputs = [], [], []
        direct = _rp._direct_leaf_grads
        # the argument list of each scene, and the pooled topology it matches (if any)
        arg_lists, k = [], 0
        for (w, h, nsx, nsy, seed, n) in spec:
            arg_lists.append(args[k:k + n])
            k += n
        tpls = [_find_template(a, _HINTS[i] if i < len(_HINTS) else None) for i, a in enumerate(arg_lists)]
        plans = [tp.direct_leaves(a, dev) if (tp is not None and grad and direct) else None
                 for tp, a in zip(tpls, arg_lists)]
        pooled = [i for i, tp in enumerate(tpls) if tp is not None]
        gathered = {}
        if len(pooled) > 1:
            # one gather per kind over all pooled scenes instead of one per scene
            try:
                need = grad and any(plans[i] is None for i in pooled)
                for i, vals in zip(pooled, _batch_params([tpls[i] for i in pooled],
                                                         [arg_lists[i] for i in pooled], dev, need)):
                    if plans[i] is not None:  # delivered directly: no edge into the graph
                        vals = {kd: None if v is None else v.detach() for kd, v in vals.items()}
                    gathered[i] = vals
            except _Resized:
                gathered = {}
        try:
            for i, (w, h, nsx, nsy, seed, n) in enumerate(spec):
                a = arg_lists[i]
                if i in gathered:
                    flat, scene, vals = _pooled_scene(tpls[i], a, dev, gathered[i], plans[i], False)
                else:
                    flat, scene, vals = _device_scene_for(a, dev, check=False, grad=grad, direct=direct,
                                                          hint=tpls[i])
                scenes.append(scene)
                flats.append(flat)
                if flat.output_type != OutputType.color or flat.eval_positions.shape[0] != 0:
                    raise NotImplementedError('BatchRenderFunction renders colour images (OutputType.color)')
                if nsx * nsy > 256:
                    raise ValueError('a batched render takes at most 256 samples per pixel')
                radius = a[flat.num_args - 1]
                inputs.append(radius if isinstance(radius, torch.Tensor) else None)
                inputs += [None] * len(_KINDS) if vals is None else [vals[kd] for kd in _KINDS]
            if k != len(args):
                raise ValueError('spec does not account for every argument')
            # the template each position used, tried first by the next batch (training
            # loops render the same scenes in the same order)
            _HINTS[:] = [getattr(sc, '_template', None) for sc in scenes]
            if any(getattr(f, 'direct', None) for f in flats):
                inputs.append(torch.empty(0, requires_grad=True))
            for sc in scenes:
                check_rebuild(sc)
            # findings of earlier backward passes, read once this call's scene work is on
            # the stream (the rebuild checks have waited for their kernels)
            synchronize()
        except BaseException:
            for sc in scenes:
                if isinstance(sc, _Lease):
                    sc.release()
            raise
        return _BatchNode.apply((tuple(spec), scenes, flats), *inputs)


class _BatchNode(torch.autograd.Function):
    """One batched forward and backward; inputs per scene: filter radius, then the float
    arrays in _KINDS order."""

    @staticmethod
    def forward(ctx, meta, *inputs):
        spec, scenes, flats = meta
        dev = get_device()
        st = _stream()
        images, jobs = [], []
        for (w, h, nsx, nsy, seed, n), scene, flat in zip(spec, scenes, flats):
            img = torch.zeros(h, w, 4, device=dev)
            images.append(img)
            jobs.append((scene, dict(width=w, height=h, nsx=nsx, nsy=nsy, seed=seed, render_image=img,
                                     use_prefiltering=flat.use_prefiltering)))
        fwd = Batch(jobs, zero_grads=False)
        fwd.render(st)
        if any(flat.use_prefiltering for flat in flats):
            # a prefiltered stroke of a circle group, on which the reference asserts in the
            # forward: read at the next render call or synchronize(), as RenderFunction does
            _PENDING.append((_BatchCheck(fwd, None), None))
        ctx.spec, ctx.scenes, ctx.flats = spec, scenes, flats
        ctx.shapes = [None if t is None else t.shape for t in inputs]
        ctx.fwd = fwd
        return tuple(images)

    @staticmethod
    def backward(ctx, *grad_imgs):
        # earlier findings are read at the next render call or synchronize(), not here
        # the forward batch holds the output images, whose graph holds this context:
        # dropping it here lets the scenes go back to their pool without a cycle collection
        ctx.fwd = None
        dev = get_device()
        st = _stream()
        jobs, keep = [], []
        for i, ((w, h, nsx, nsy, seed, n), sc, flat) in enumerate(zip(ctx.spec, ctx.scenes, ctx.flats)):
            g = grad_imgs[i]
            g = torch.zeros(h, w, 4, device=dev) if g is None else g.to(torch.float32).contiguous()
            keep.append(g)
            jobs.append((sc, dict(width=w, height=h, nsx=nsx, nsy=nsy, seed=seed, d_render_image=g,
                                  use_prefiltering=flat.use_prefiltering)))
        bwd = Batch(jobs, zero_grads=True)
        bwd.render(st)
        need = ctx.needs_input_grad
        per = 1 + len(_KINDS)
        out = [None] * len(ctx.shapes)
        for i, sc in enumerate(ctx.scenes):
            base = per * i
            direct = getattr(ctx.flats[i], 'direct', None)
            if direct is None and not any(need[1 + base + j] for j in range(per)):
                continue
            block, layout = sc.grads_device(dev, st)
            if direct is not None:
                direct[0].deliver(block, layout, direct[1])
            for j, name in enumerate(('filter_radius',) + _KINDS):
                shape = ctx.shapes[base + j]
                if shape is None or not need[1 + base + j]:
                    continue
                o = layout[name]
                cnt = 1
                for d in shape:
                    cnt *= d
                out[base + j] = block[o:o + cnt].view(shape)
        # as RenderFunction's backward: returns while the kernels run; the batch's findings
        # (and a non-finite grad_img) are read at the next entry or at synchronize()
        finite = torch.stack([g.isfinite().all() for g in keep]).all()
        _PENDING.append((_BatchCheck(bwd, ctx.scenes), finite))
        ctx.scenes = None
        return (None,) + tuple(out)


class _BatchCheck:
    """Keeps a batch (and, for a backward batch, its scene leases) until its findings are
    read."""

    def __init__(self, batch, scenes):
        self.batch, self.scenes = batch, scenes

    def status(self):
        self.batch.status()


def render_batch(items, num_samples_x=2, num_samples_y=2, seeds=None):
    """items: list of (canvas_width, canvas_height, shapes, shape_groups) or of
    serialize_scene argument lists; returns the list of rendered images (differentiable
    w.r.t. every parameter tensor of every scene)."""
    from .render_pytorch import RenderFunction
    spec, flat_args = [], []
    for i, it in enumerate(items):
        if isinstance(it, (list, tuple)) and len(it) == 4 and isinstance(it[2], (list, tuple)):
            w, h, shapes, groups = it
            a = RenderFunction.serialize_scene(w, h, shapes, groups)
        else:
            a = list(it)
            w, h = int(a[0]), int(a[1])
        seed = 0 if seeds is None else int(seeds[i])
        spec.append((int(w), int(h), int(num_samples_x), int(num_samples_y), seed, len(a)))
        flat_args += a
    return list(BatchRenderFunction.apply(tuple(spec), *flat_args))
