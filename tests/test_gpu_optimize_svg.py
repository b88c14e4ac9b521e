"""OptimizableSvg on the GPU: the scenes its document tree builds render like the oracle
renders them (forward and backward), its `render` equals svg_to_scene's scene where the
two loaders read a document alike, and its optimizers recover a rotation and a
translation the way apps/simple_transform_svg.py trains one."""
import math

import numpy as np
import pytest
import torch

import diffvg_amd as pd
from diffvg_amd import _lib
from gpudrive import GpuRunner
from refdrive import OracleRunner, flatten
from test_gpu_parity import _compare

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings('ignore')]

DOC = """<svg xmlns="http://www.w3.org/2000/svg" xmlns:xlink="http://www.w3.org/1999/xlink"
     version="2.0" width="80" height="64">
  <defs>
    <linearGradient id="base">
      <stop offset="0" stop-color="#ff8000"/>
      <stop offset="0.5" stop-color="#0080ff" stop-opacity="0.6"/>
      <stop offset="1" stop-color="#20c040"/>
    </linearGradient>
    <linearGradient id="lin" xlink:href="#base" x1="0" y1="0" x2="30" y2="10"/>
  </defs>
  <g id="outer" transform="translate(6 4) rotate(12)" opacity="0.8">
    <g id="inner" transform="scale(1.2)" style="fill:#3366cc;stroke:#aa2211;stroke-width:1.5">
      <path id="blob" d="M 5 5 C 15 0 25 10 30 5 Q 35 20 20 28 L 4 20 Z"/>
      <circle id="dot" cx="40" cy="12" r="7"/>
    </g>
    <rect id="box" x="4" y="34" width="30" height="14" fill="url(#lin)"/>
  </g>
  <ellipse id="ell" cx="62" cy="44" rx="12" ry="7" fill="#118877" fill-opacity="0.7"/>
  <polygon id="poly" points="50,4 74,8 70,26 54,20" fill="#55aa33" fill-rule="evenodd"/>
</svg>
"""


@pytest.fixture(autouse=True)
def _gpu():
    pd.set_use_gpu(True)
    pd.set_device(torch.device('cuda:0'))


@pytest.fixture
def doc(tmp_path):
    f = tmp_path / 'doc.svg'
    f.write_text(DOC)
    return str(f)


@pytest.mark.parametrize('device', ['cpu', 'cuda'])
def test_optimizable_scene_matches_oracle(doc, device):
    print('optimizable svg vs oracle', device, flush=True)
    o = pd.OptimizableSvg(doc, device=torch.device(device))
    w, h, shapes, groups = o.build_scene()
    flat = flatten(int(w), int(h), shapes, groups, 0, 0.5)
    ora, gpu = OracleRunner(flat), GpuRunner(flat, _lib.MODE_TILED)
    _compare(gpu.render(int(w), int(h), 2, 2, 0), ora.render(int(w), int(h), 2, 2, 0), ('optsvg', 'fwd'))
    d = np.random.default_rng(2).uniform(-1, 1, (int(h), int(w), 4)).astype(np.float32)
    kw = dict(d_image=d, translation=True)
    _compare(gpu.render(int(w), int(h), 2, 2, 0, **kw), ora.render(int(w), int(h), 2, 2, 0, **kw), ('optsvg', 'bwd'))
    # and through the document's own render call
    img = o.render()
    ref = gpu.render(int(w), int(h), 2, 2, 0)['image']
    assert np.abs(img.detach().cpu().numpy() - ref).max() <= 1e-5


def test_render_equals_svg_to_scene_where_loaders_agree(tmp_path):
    """Fills without transforms or strokes: both loaders build the same tensors."""
    print('optimizable svg vs svg_to_scene', flush=True)
    f = tmp_path / 'a.svg'
    f.write_text('<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="48" height="40">'
                 '<path d="M 4 4 C 14 0 24 10 30 5 Q 35 20 20 28 L 4 20 Z" fill="#3366cc"/>'
                 '<circle cx="34" cy="26" r="9" fill="#aa2211" fill-opacity="0.5"/>'
                 '<polygon points="30,2 46,6 42,18" fill="#55aa33"/></svg>')
    a = pd.OptimizableSvg(str(f)).render(seed=4)
    w, h, shapes, groups = pd.svg_to_scene(str(f))
    b = pd.RenderFunction.apply(w, h, 2, 2, 4, None, *pd.RenderFunction.serialize_scene(w, h, shapes, groups))
    assert torch.equal(a.detach(), b.detach())


def test_training_recovers_rotation_and_translation(tmp_path):
    """A shape under a group transform trains (rigid mode) towards a target rendered
    with that group rotated by 0.2 rad and moved by (3, -2): the loss falls and the
    trained rotation and translation approach the target's."""
    print('optimizable svg training', flush=True)
    body = ('<path d="M -12 -8 L 14 -8 L 14 2 L 4 2 L 4 10 L -12 10 Z" fill="#3366cc"/>'
            '<circle cx="8" cy="6" r="3" fill="#aa2211"/>')
    head = '<svg xmlns="http://www.w3.org/2000/svg" version="2.0" width="64" height="64">'
    src = tmp_path / 'src.svg'
    src.write_text(head + '<g id="g" transform="translate(32 32)">' + body + '</g></svg>')
    tgt = tmp_path / 'tgt.svg'
    angle, shift = 0.2, (3.0, -2.0)
    tgt.write_text(head + '<g id="g" transform="translate({} {}) rotate({})">'.format(
        32 + shift[0], 32 + shift[1], math.degrees(angle)) + body + '</g></svg>')
    target = pd.OptimizableSvg(str(tgt)).render(seed=0).detach()

    s = pd.SvgOptimizationSettings()
    s.global_override(["optimize_color"], False)
    s.global_override(["paths", "optimize_points"], False)
    s.global_override(["transforms", "transform_lr"], 1e-2)
    o = pd.OptimizableSvg(str(src), s)
    node = o.node('g').transform
    losses = []
    for it in range(150):
        o.zero_grad()
        img = o.render(seed=it)
        loss = (img - target).pow(2).mean()
        loss.backward()
        losses.append(float(loss))
        o.step()
        if it % 25 == 0:
            print(it, losses[-1], float(node.theta), float(node.scale), node.translate.tolist(), flush=True)
    pd.synchronize()
    theta = float(node.theta)
    # the group's translation is applied before its rotation and scale: compare the
    # composed matrix's offset with the target's canvas offset
    M = node.matrix().detach().cpu().numpy()
    print('final', losses[-1], theta, M.tolist(), flush=True)
    assert np.mean(losses[-5:]) < 0.1 * np.mean(losses[:3]), losses
    assert abs(theta - angle) < 0.1 * angle, theta
    np.testing.assert_allclose(M[0:2, 2], [32 + shift[0], 32 + shift[1]], atol=0.5)
