"""GPU-side test driver: runs libdiffvg_amd.so through its C ABI (diffvg_amd._lib) with
the same call signature and outputs as refdrive.OracleRunner (test infrastructure)."""
import numpy as np
import torch

from diffvg_amd import _lib


class GpuRunner:
    def __init__(self, flat, mode=_lib.MODE_TILED, device=0):
        self.flat = flat
        self.dev = torch.device('cuda', device)
        self.scene = _lib.DeviceScene(flat, device, mode)

    def _t(self, a):
        return None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(self.dev)

    def render(self, width, height, nsx, nsy, seed, *, output='color', prefilter=False, background=None,
               d_image=None, d_sdf=None, translation=False, eval_positions=None, d_background=False):
        dev = self.dev
        eval_n = 0 if eval_positions is None else int(eval_positions.shape[0])
        img = sdf = None
        if d_image is None and d_sdf is None:
            if output == 'color':
                img = torch.zeros(height, width, 4, device=dev)
            else:
                sdf = torch.zeros((eval_n, 1) if eval_n else (height, width, 1), device=dev)
        d_tr = torch.zeros(height, width, 2, device=dev) if translation else None
        d_bg = torch.zeros(height, width, 4, device=dev) if (d_background and background is not None) else None
        bg, ev, di, ds = self._t(background), self._t(eval_positions), self._t(d_image), self._t(d_sdf)
        self.scene.zero_grads(None)
        self.scene.render(width, height, nsx, nsy, seed, background=bg, render_image=img, render_sdf=sdf,
                          d_background=d_bg, d_render_image=di, d_render_sdf=ds, d_translation=d_tr,
                          use_prefiltering=prefilter, eval_positions=ev, num_eval_positions=eval_n, stream=None)
        torch.cuda.synchronize()

        def h(t):
            return None if t is None else t.cpu().numpy()

        out = dict(image=h(img), sdf=h(sdf), d_translation=h(d_tr), d_background=h(d_bg))
        if d_image is not None or d_sdf is not None:
            out['grads'] = self.scene.grads()
        return out
