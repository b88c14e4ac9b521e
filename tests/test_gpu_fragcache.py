"""A colour backward that follows its forward (same seed, sizes and parameters) replays
the fragment lists the forward kept instead of shading its samples again.  Its results
must be those of a backward that shades: compared here against a fresh scene's backward
(no forward before it), with scenes whose samples have more fragments than are kept
(those samples shade again), and with forwards whose lists must not be used (another
seed, parameters changed by a device rebuild)."""
import numpy as np
import pytest
import torch

import diffvg_amd as pd
from gpudrive import GpuRunner
import scenes

pytestmark = pytest.mark.gpu

KEYS = ('shape_params', 'points', 'thickness', 'stroke_width', 'color_params', 'shape_to_canvas', 'filter_radius')


def layers(size=48, n=14, strokes=True):
    """n overlapping translucent circles: the centre samples have n fragments, more than
    a forward keeps (without strokes for prefiltering, where circle strokes are rejected
    as the reference rejects them)."""
    shapes = [pd.Circle(radius=scenes.T(14.0 - 0.5 * k), center=scenes.T([24.0 + 0.3 * k, 24.0 - 0.2 * k]))
              for k in range(n)]
    groups = [pd.ShapeGroup(shape_ids=torch.tensor([k]), fill_color=scenes.T([0.1 * (k % 7), 0.5, 0.9 - 0.05 * k, 0.3]),
                            stroke_color=scenes.T([0, 0, 0, 0.5]) if strokes and k % 3 == 0 else None)
              for k in range(n)]
    for k, s in enumerate(shapes):
        s.stroke_width = scenes.T(0.8)
    return dict(canvas=(size, size), shapes=shapes, groups=groups, filter=('box', 0.5))


def _flat(case):
    if case == 'layers':
        return scenes.flat_case(layers())
    if case == 'layers_fill':
        return scenes.flat_case(layers(strokes=False))
    return scenes.flat_case(case)


def _same(a, b, label):
    assert np.array_equal(a['d_translation'], b['d_translation']) or \
        np.abs(a['d_translation'] - b['d_translation']).max() <= 1e-3 * max(np.abs(b['d_translation']).max(), 1e-8), label
    for k in KEYS:
        x, y = a['grads'][k], b['grads'][k]
        if y.size == 0:
            continue
        assert np.abs(x - y).max() <= 1e-3 * max(np.abs(y).max(), 1e-8), (label, k)


@pytest.mark.parametrize('case,prefilter', [('overlap_eo', False), ('gradients', False), ('strokes', False),
                                             ('transformed', False), ('layers', False), ('overlap_eo', True),
                                             ('gradients', True), ('strokes', True), ('transformed', True),
                                             ('layers_fill', True)])
def test_replayed_backward_equals_shading_backward(case, prefilter):
    print('fragcache', case, prefilter, flush=True)
    f = _flat(case)
    w, h = f['canvas_width'], f['canvas_height']
    d = np.random.default_rng(3).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    a = GpuRunner(f)
    fwd = a.render(w, h, 2, 2, 5, prefilter=prefilter)
    kept = a.render(w, h, 2, 2, 5, d_image=d, translation=True, prefilter=prefilter)
    fresh = GpuRunner(f).render(w, h, 2, 2, 5, d_image=d, translation=True, prefilter=prefilter)
    _same(kept, fresh, case)
    # a second forward of the same render gives the same image
    assert np.array_equal(a.render(w, h, 2, 2, 5, prefilter=prefilter)['image'], fwd['image'])


def test_lists_of_another_seed_are_not_used():
    print('fragcache seed', flush=True)
    f = _flat('layers')
    w, h = f['canvas_width'], f['canvas_height']
    d = np.random.default_rng(4).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    a = GpuRunner(f)
    a.render(w, h, 2, 2, 1)
    other = a.render(w, h, 2, 2, 2, d_image=d, translation=True)
    fresh = GpuRunner(f).render(w, h, 2, 2, 2, d_image=d, translation=True)
    _same(other, fresh, 'seed')


def test_lists_are_dropped_by_a_rebuild():
    print('fragcache rebuild', flush=True)
    f1 = _flat('layers')
    w, h = f1['canvas_width'], f1['canvas_height']
    f2 = dict(f1)
    f2['shape_params'] = (f1['shape_params'] * 1.15).astype(np.float32)
    d = np.random.default_rng(5).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    a = GpuRunner(f1)
    a.render(w, h, 2, 2, 7)
    a.scene.update(f2['filter_radius'], None, shape_params=torch.from_numpy(f2['shape_params']).cuda())
    a.scene.status()
    after = a.render(w, h, 2, 2, 7, d_image=d, translation=True)
    fresh = GpuRunner(f2).render(w, h, 2, 2, 7, d_image=d, translation=True)
    _same(after, fresh, 'rebuild')


def _batch_grads(sc, w, h, seed, d, prefilter, forward_seed=None):
    """Batched backward on `sc` (after a batched forward with `forward_seed`, if given)."""
    from diffvg_amd.batch import Batch
    dev = torch.device('cuda:0')
    if forward_seed is not None:
        img = torch.zeros(h, w, 4, device=dev)
        fb = Batch([(sc, dict(width=w, height=h, nsx=2, nsy=2, seed=forward_seed, render_image=img,
                              use_prefiltering=prefilter))], zero_grads=False)
        fb.render(None)
    dt = torch.zeros(h, w, 2, device=dev)
    di = torch.from_numpy(d).to(dev)
    bb = Batch([(sc, dict(width=w, height=h, nsx=2, nsy=2, seed=seed, d_render_image=di, d_translation=dt,
                          use_prefiltering=prefilter))], zero_grads=True)
    bb.render(None)
    torch.cuda.synchronize()
    bb.status()
    return dict(grads=sc.grads(), d_translation=dt.cpu().numpy())


@pytest.mark.parametrize('prefilter', [False, True])
def test_batched_backward_replays_only_matching_lists(prefilter):
    """Batches keep lists too (graph replays included, as the batch tests check); the key
    a forward wrote is checked on the device: another seed, or a device update of the
    parameters after the forward, makes the backward shade."""
    from diffvg_amd import _lib
    print('fragcache batch', prefilter, flush=True)
    f1 = _flat('layers_fill')
    w, h = f1['canvas_width'], f1['canvas_height']
    d = np.random.default_rng(9).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    fresh = _batch_grads(_lib.DeviceScene(f1, 0), w, h, 5, d, prefilter)
    kept = _batch_grads(_lib.DeviceScene(f1, 0), w, h, 5, d, prefilter, forward_seed=5)
    _same(kept, fresh, 'batch kept')
    other = _batch_grads(_lib.DeviceScene(f1, 0), w, h, 5, d, prefilter, forward_seed=6)
    _same(other, fresh, 'batch other seed')
    # forward, then a device update, then the backward: the lists are stale
    f2 = dict(f1)
    f2['shape_params'] = (f1['shape_params'] * 1.15).astype(np.float32)
    from diffvg_amd.batch import Batch
    sc = _lib.DeviceScene(f1, 0)
    img = torch.zeros(h, w, 4, device='cuda')
    Batch([(sc, dict(width=w, height=h, nsx=2, nsy=2, seed=5, render_image=img, use_prefiltering=prefilter))],
          zero_grads=False).render(None)
    sc.update(f2['filter_radius'], None, shape_params=torch.from_numpy(f2['shape_params']).cuda())
    sc.status()
    after = _batch_grads(sc, w, h, 5, d, prefilter)
    fresh2 = _batch_grads(_lib.DeviceScene(f2, 0), w, h, 5, d, prefilter)
    _same(after, fresh2, 'batch after update')


@pytest.mark.parametrize('prefilter', [False, True])
def test_scene_twice_in_one_batch_keeps_no_lists(prefilter):
    """Two forward jobs of one scene in a batch would write one list buffer at once: such a
    scene keeps no lists, and a backward after that batch equals a fresh one."""
    from diffvg_amd import _lib
    from diffvg_amd.batch import Batch
    print('fragcache scene twice in a batch', prefilter, flush=True)
    f1 = _flat('layers_fill')
    w, h = f1['canvas_width'], f1['canvas_height']
    d = np.random.default_rng(11).uniform(-1, 1, (h, w, 4)).astype(np.float32)
    fresh = _batch_grads(_lib.DeviceScene(f1, 0), w, h, 5, d, prefilter)
    sc = _lib.DeviceScene(f1, 0)
    imgs = [torch.zeros(h, w, 4, device='cuda') for _ in range(2)]
    Batch([(sc, dict(width=w, height=h, nsx=2, nsy=2, seed=s, render_image=im, use_prefiltering=prefilter))
           for s, im in ((6, imgs[0]), (5, imgs[1]))], zero_grads=False).render(None)
    torch.cuda.synchronize()
    after = _batch_grads(sc, w, h, 5, d, prefilter)
    _same(after, fresh, 'scene twice')
