"""Scene templates (the topology a pooled device scene is rebuilt for) on the CPU: a later
argument list matches when its topology is the template's; the same tensor objects skip
the shape comparison, and a tensor resized in place since is caught by the gathers."""
import torch

import diffvg_amd as pd
from diffvg_amd import render_pytorch as rp
import scenes


def _args(sc):
    w, h = sc['canvas']
    return pd.RenderFunction.serialize_scene(w, h, sc['shapes'], sc['groups'])


def test_template_matches_same_topology_and_rejects_other_shapes():
    sc = scenes.CASES['strokes']()
    a = _args(sc)
    tpl = rp._SceneTemplate(a, rp.flatten_args(a))
    assert tpl.matches(_args(sc)) and tpl.matches(_args(sc))
    # new tensors with the same shapes: still the topology
    for s in sc['shapes']:
        s.points = s.points.clone() + 1.0
    assert tpl.matches(_args(sc))
    # a path with one more point: another topology
    p = sc['shapes'][3]
    p.points = torch.cat([p.points, p.points[-1:]])
    p.num_control_points = torch.cat([p.num_control_points, torch.zeros(1, dtype=torch.int32)])
    assert not tpl.matches(_args(sc))


def test_in_place_resize_is_caught_by_the_gather():
    sc = scenes.CASES['strokes']()
    a = _args(sc)
    tpl = rp._SceneTemplate(a, rp.flatten_args(a))
    assert tpl.matches(_args(sc))
    vals = tpl.device_params(_args(sc), torch.device('cpu'))
    assert vals['points'].numel() == 2 * tpl.flat.desc['num_points_total']
    p = sc['shapes'][0].points
    p.data = torch.cat([p.data, p.data[:1]])  # same tensor object, one more point
    b = _args(sc)
    assert tpl.matches(b)  # the identity path does not compare shapes ...
    try:
        tpl.device_params(b, torch.device('cpu'))
        raise AssertionError('resize not caught')
    except rp._Resized:
        pass  # ... the gather's element count does
    tpl._float_seen = None
    assert not tpl.matches(b)
