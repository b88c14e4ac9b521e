"""Writes the golden fixtures of the render path from the reference's own CPU build.

Run in the build container only (needs oracle/_ref/diffvg.so, made by `make -C oracle ref`
from the reference sources):

    python tests/golden/make_goldens.py

Each fixture `tests/golden/<case>.npz` holds the flattened scene (keys `desc_*`, the
dvg_scene_desc arrays), the render settings, the seeded adjoint inputs and the reference's
outputs:

    image, image_pre            forward colour, prefiltering off / on
    grads_*, grads_pre_*        scene gradients for d_image (d-buffers of diffvg.cpp:1626-1660)
    d_translation               render_grad's translation image (prefiltering off)
    sdf, sdf_grads_*            SDF output and its gradients (selected cases)
    sdf_eval, sdf_eval_grads_*  the same at 200 seeded eval positions

Gradients the reference computes but does not expose (stroke gradient paints,
scene.cpp:866-869) are stored as NaN and skipped by the tests.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from refdrive import RefRunner, load_ref  # noqa: E402
from scenes import flat_case  # noqa: E402

SEED = 3
NS = 2
# (case, filter override or None, prefilter variant, sdf variant)
CASES = [
    ('single_circle_256', None, False, False),
    ('single_circle_64', None, True, False),
    ('curve_fill', None, True, True),
    ('open_stroke', None, True, False),
    ('thick_stroke', None, True, False),
    ('mixed_segments', None, True, True),
    ('mixed_segments', ('tent', 1.0), False, False),
    ('mixed_segments', ('hann', 1.0), False, False),
    ('mixed_segments', ('parabolic', 1.0), False, False),
    ('quad_approx', None, True, False),
    ('primitives', None, False, False),  # prefiltered: the reference aborts (diffvg.cpp:867)
    ('gradients', None, True, False),
    ('transformed', None, True, True),
    ('overlap_eo', None, True, True),
    ('overlap_eo', ('tent', 1.0), True, False),
    ('overlap_nz', None, True, False),
    ('strokes', None, True, False),
]


def fixture_name(case, filt):
    return case if filt is None else '%s_%s' % (case, filt[0])


def single_circle_target():
    from refdrive import flatten
    from diffvg_amd import workloads
    c = workloads.single_circle()
    shapes, groups = c['target_scene']
    w, h = c['canvas']
    return flatten(w, h, shapes, groups, 0, 0.5)


def d_input(shape, seed):
    return np.random.default_rng(seed).uniform(-1, 1, shape).astype(np.float32)


def main():
    if load_ref() is None:
        sys.exit('oracle/_ref/diffvg.so is not built (make -C oracle ref)')
    index = {}
    for case, filt, pre, sdf in CASES:
        f = flat_case(case, filt)
        w, h = f['canvas_width'], f['canvas_height']
        r = RefRunner(f)
        out = {}
        for k, v in f.items():
            out['desc_' + k] = np.asarray(v)
        out['seed'] = np.int64(SEED)
        out['ns'] = np.int64(NS)
        out['image'] = r.render(w, h, NS, NS, SEED)['image']
        if case == 'single_circle_256':
            # the single_circle.py iteration: L2 loss against the circle at (100,120), r = 60
            tw = single_circle_target()
            target = RefRunner(tw).render(w, h, NS, NS, SEED)['image']
            out['target'] = target
            d_img = (2.0 * (out['image'] - target)).astype(np.float32)
        else:
            d_img = d_input((h, w, 4), 0)
        out['d_image'] = d_img
        g = r.render(w, h, NS, NS, SEED, d_image=d_img, translation=True)
        out['d_translation'] = g['d_translation']
        for k, v in g['grads'].items():
            out['grads_' + k] = v
        if pre:
            out['image_pre'] = r.render(w, h, NS, NS, SEED, prefilter=True)['image']
            g = r.render(w, h, NS, NS, SEED, d_image=d_img, prefilter=True)
            for k, v in g['grads'].items():
                out['grads_pre_' + k] = v
        if sdf:
            out['sdf'] = r.render(w, h, 1, 1, 0, output='sdf')['sdf']
            d_sdf = d_input((h, w, 1), 1)
            out['d_sdf'] = d_sdf
            g = r.render(w, h, 1, 1, 0, output='sdf', d_sdf=d_sdf)
            for k, v in g['grads'].items():
                out['sdf_grads_' + k] = v
            # eval-positions mode (render_pytorch.py:94-97): SDF at arbitrary points
            ev = np.random.default_rng(2).uniform(0, [w, h], (200, 2)).astype(np.float32)
            out['eval_positions'] = ev
            out['sdf_eval'] = r.render(w, h, 1, 1, 0, output='sdf', eval_positions=ev)['sdf']
            d_ev = d_input((200, 1), 3)
            out['d_sdf_eval'] = d_ev
            g = r.render(w, h, 1, 1, 0, output='sdf', d_sdf=d_ev, eval_positions=ev)
            for k, v in g['grads'].items():
                out['sdf_eval_grads_' + k] = v
        name = fixture_name(case, filt)
        np.savez_compressed(os.path.join(HERE, name + '.npz'), **out)
        index[name] = dict(case=case, filter=filt, prefilter=pre, sdf=sdf)
        print('wrote', name, flush=True)
    with open(os.path.join(HERE, 'index.json'), 'w') as fh:
        json.dump(index, fh, indent=1, sort_keys=True)


if __name__ == '__main__':
    main()
