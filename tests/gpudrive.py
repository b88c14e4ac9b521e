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


class ShimRunner:
    """Drives the drop-in `diffvg` module (diffvg_amd.diffvg) exactly as the reference's
    render_pytorch.py drives its pybind11 module, with device image buffers."""

    def __init__(self, flat, device=0):
        from diffvg_amd import diffvg as dv
        from refdrive import RefRunner
        self.dv = dv
        self.dev = torch.device('cuda', device)
        self.inner = RefRunner(flat, dv=dv, use_gpu=True, gpu_index=device)

    def render(self, width, height, nsx, nsy, seed, *, output='color', prefilter=False, background=None,
               d_image=None, d_sdf=None, translation=False, eval_positions=None, d_background=False):
        dv, dev, r = self.dv, self.dev, self.inner
        eval_n = 0 if eval_positions is None else int(eval_positions.shape[0])
        img = sdf = None
        if d_image is None and d_sdf is None:
            if output == 'color':
                img = torch.zeros(height, width, 4, device=dev)
            else:
                sdf = torch.zeros((eval_n, 1) if eval_n else (height, width, 1), device=dev)
        d_tr = torch.zeros(height, width, 2, device=dev) if translation else None
        t = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
        bg, ev, di, ds = t(background), t(eval_positions), t(d_image), t(d_sdf)
        d_bg = torch.zeros(height, width, 4, device=dev) if (d_background and background is not None) else None
        P = lambda x: dv.float_ptr(0 if x is None else x.data_ptr())
        if d_image is not None or d_sdf is not None:
            r.scene = dv.Scene(r.flat['canvas_width'], r.flat['canvas_height'], r.shapes, r.groups, r.filt, True,
                               dev.index)
        dv.render(r.scene, P(bg), P(img), P(sdf), width, height, nsx, nsy, seed, P(d_bg), P(di), P(ds), P(d_tr),
                  bool(prefilter), P(ev), eval_n)
        torch.cuda.synchronize()
        h = lambda x: None if x is None else x.cpu().numpy()
        out = dict(image=h(img), sdf=h(sdf), d_translation=h(d_tr), d_background=h(d_bg))
        if d_image is not None or d_sdf is not None:
            out['grads'] = r._grads()
        return out
