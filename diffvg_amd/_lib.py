"""ctypes binding of libdiffvg_amd.so (the C ABI declared in include/diffvg_amd.h).

The library is loaded lazily and loudly: if it is missing or cannot be loaded, every
render raises; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libdiffvg_amd.so')

I32P = ctypes.POINTER(ctypes.c_int32)
F32P = ctypes.POINTER(ctypes.c_float)
U8P = ctypes.POINTER(ctypes.c_uint8)

MODE_TILED = 0
MODE_PLAIN = 1

DESC_FIELDS = [
    ('canvas_width', ctypes.c_int32), ('canvas_height', ctypes.c_int32),
    ('num_shapes', ctypes.c_int32),
    ('shape_type', I32P), ('stroke_width', F32P), ('shape_param_offset', I32P), ('shape_params', F32P),
    ('num_shape_params', ctypes.c_int32),
    ('path_ctrl_offset', I32P), ('path_num_base_points', I32P), ('path_point_offset', I32P),
    ('path_num_points', I32P), ('path_has_thickness', U8P), ('path_is_closed', U8P),
    ('path_use_distance_approx', U8P), ('num_control_points', I32P), ('num_ctrl_total', ctypes.c_int32),
    ('points', F32P), ('thickness', F32P), ('num_points_total', ctypes.c_int32),
    ('num_groups', ctypes.c_int32), ('group_shape_offset', I32P), ('group_num_shapes', I32P),
    ('group_shape_ids', I32P), ('num_group_shape_ids', ctypes.c_int32),
    ('fill_type', I32P), ('fill_offset', I32P), ('fill_num_stops', I32P),
    ('stroke_type', I32P), ('stroke_offset', I32P), ('stroke_num_stops', I32P),
    ('color_params', F32P), ('num_color_params', ctypes.c_int32),
    ('use_even_odd_rule', U8P), ('shape_to_canvas', F32P),
    ('filter_type', ctypes.c_int32), ('filter_radius', ctypes.c_float),
]


class SceneDesc(ctypes.Structure):
    _fields_ = DESC_FIELDS


class RenderArgs(ctypes.Structure):
    _fields_ = [('background_image', ctypes.c_void_p), ('render_image', ctypes.c_void_p),
                ('render_sdf', ctypes.c_void_p),
                ('width', ctypes.c_int32), ('height', ctypes.c_int32),
                ('num_samples_x', ctypes.c_int32), ('num_samples_y', ctypes.c_int32),
                ('seed', ctypes.c_uint64), ('d_background_image', ctypes.c_void_p),
                ('d_render_image', ctypes.c_void_p), ('d_render_sdf', ctypes.c_void_p),
                ('d_translation', ctypes.c_void_p), ('use_prefiltering', ctypes.c_int32),
                ('eval_positions', ctypes.c_void_p), ('num_eval_positions', ctypes.c_int32),
                ('stream', ctypes.c_void_p)]


class SceneParams(ctypes.Structure):
    _fields_ = [('shape_params', ctypes.c_void_p), ('points', ctypes.c_void_p), ('thickness', ctypes.c_void_p),
                ('stroke_width', ctypes.c_void_p), ('color_params', ctypes.c_void_p),
                ('shape_to_canvas', ctypes.c_void_p)]


class Grads(ctypes.Structure):
    _fields_ = [('d_shape_params', F32P), ('d_points', F32P), ('d_thickness', F32P),
                ('d_stroke_width', F32P), ('d_color_params', F32P), ('d_shape_to_canvas', F32P),
                ('d_filter_radius', F32P)]


_lib = None


def lib():
    """Returns the loaded library; raises RuntimeError when it is not available."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError('diffvg_amd: %s is not built; run `python -m diffvg_amd.build` '
                           '(or __graft_entry__.build())' % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.dvg_scene_create.argtypes = [ctypes.POINTER(SceneDesc), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.dvg_scene_destroy.argtypes = [ctypes.c_void_p]
    L.dvg_render.argtypes = [ctypes.c_void_p, ctypes.POINTER(RenderArgs)]
    L.dvg_render_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(RenderArgs)]
    L.dvg_scene_get_grads.argtypes = [ctypes.c_void_p, ctypes.POINTER(Grads)]
    L.dvg_scene_zero_grads.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.dvg_scene_set_mode.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dvg_scene_set_tiling.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dvg_scene_set_samples_per_wave.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dvg_scene_grad_layout.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64 * 7),
                                        ctypes.POINTER(ctypes.c_int64)]
    L.dvg_scene_copy_grads.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.dvg_scene_update.argtypes = [ctypes.c_void_p, ctypes.POINTER(SceneParams), ctypes.c_float, ctypes.c_void_p]
    L.dvg_scene_status.argtypes = [ctypes.c_void_p]
    L.dvg_scene_debug_block.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                        ctypes.POINTER(ctypes.c_int64)]
    L.dvg_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(RenderArgs), ctypes.c_int,
                                   ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.dvg_batch_render.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.dvg_batch_status.argtypes = [ctypes.c_void_p]
    L.dvg_batch_destroy.argtypes = [ctypes.c_void_p]
    L.dvg_profile_enable.argtypes = [ctypes.c_int]
    L.dvg_profile_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_int), ctypes.c_int]
    L.dvg_last_error.restype = ctypes.c_char_p
    L.dvg_abi_version.restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise RuntimeError('diffvg_amd: ' + lib().dvg_last_error().decode(errors='replace'))


def make_desc(flat):
    """Flat numpy scene description -> (SceneDesc, keep-alive list)."""
    d = SceneDesc()
    keep = []
    for name, ctype in DESC_FIELDS:
        v = flat[name]
        if isinstance(v, np.ndarray):
            arr = np.ascontiguousarray(v)
            keep.append(arr)
            setattr(d, name, arr.ctypes.data_as(ctype))
        else:
            setattr(d, name, v)
    return d, keep


def ptr(t):
    """Device (or host) address of a tensor, or None."""
    return None if t is None else t.data_ptr()


class DeviceScene:
    """Owns one dvg_scene handle."""

    def __init__(self, flat, gpu_index=-1, mode=MODE_TILED):
        self._lib = lib()
        self.flat = flat
        desc, keep = make_desc(flat)
        h = ctypes.c_void_p()
        check(self._lib.dvg_scene_create(ctypes.byref(desc), int(gpu_index), ctypes.byref(h)))
        self.handle = h
        if mode != MODE_TILED:
            check(self._lib.dvg_scene_set_mode(self.handle, int(mode)))

    def set_mode(self, mode):
        check(self._lib.dvg_scene_set_mode(self.handle, int(mode)))

    def set_tiling(self, tile_width):
        check(self._lib.dvg_scene_set_tiling(self.handle, int(tile_width)))

    def set_samples_per_wave(self, edge_samples_per_wave):
        check(self._lib.dvg_scene_set_samples_per_wave(self.handle, int(edge_samples_per_wave)))

    @staticmethod
    def make_args(width, height, nsx, nsy, seed, background=None, render_image=None, render_sdf=None,
                  d_background=None, d_render_image=None, d_render_sdf=None, d_translation=None,
                  use_prefiltering=False, eval_positions=None, num_eval_positions=0, stream=None):
        return RenderArgs(ptr(background), ptr(render_image), ptr(render_sdf), int(width), int(height), int(nsx),
                          int(nsy), int(seed) & 0xFFFFFFFFFFFFFFFF, ptr(d_background), ptr(d_render_image),
                          ptr(d_render_sdf), ptr(d_translation), 1 if use_prefiltering else 0, ptr(eval_positions),
                          int(num_eval_positions), stream)

    def render(self, width, height, nsx, nsy, seed, **kw):
        a = self.make_args(width, height, nsx, nsy, seed, **kw)
        check(self._lib.dvg_render(self.handle, ctypes.byref(a)))

    def update(self, filter_radius, stream=None, shape_params=None, points=None, thickness=None, stroke_width=None,
               color_params=None, shape_to_canvas=None):
        """Rebuilds the scene on the device from new parameter values (float32 device
        tensors in the description's layout; None keeps an array).  The tensors must stay
        alive until the stream has run the copies (stream-ordered allocation does that)."""
        p = SceneParams(ptr(shape_params), ptr(points), ptr(thickness), ptr(stroke_width), ptr(color_params),
                        ptr(shape_to_canvas))
        check(self._lib.dvg_scene_update(self.handle, ctypes.byref(p), float(filter_radius), stream))

    def status(self):
        """Waits for the scene's last stream; raises on the conditions its last update or
        render found."""
        check(self._lib.dvg_scene_status(self.handle))

    def debug_block(self):
        n = ctypes.c_int64()
        check(self._lib.dvg_scene_debug_block(self.handle, None, 0, ctypes.byref(n)))
        buf = np.zeros(int(n.value), np.uint8)
        if n.value:
            check(self._lib.dvg_scene_debug_block(self.handle, buf.ctypes.data, int(n.value), ctypes.byref(n)))
        return buf

    def zero_grads(self, stream=None):
        check(self._lib.dvg_scene_zero_grads(self.handle, stream))

    GRAD_KINDS = ('shape_params', 'points', 'thickness', 'stroke_width', 'color_params', 'shape_to_canvas',
                  'filter_radius')

    def grad_layout(self):
        offs = (ctypes.c_int64 * 7)()
        total = ctypes.c_int64()
        check(self._lib.dvg_scene_grad_layout(self.handle, ctypes.byref(offs), ctypes.byref(total)))
        return dict(zip(self.GRAD_KINDS, [int(o) for o in offs])), int(total.value)

    def grads_device(self, torch_device, stream=None):
        """Gradient block as one device tensor plus the offsets of its arrays."""
        import torch
        layout, total = self.grad_layout()
        block = torch.empty(total, dtype=torch.float32, device=torch_device)
        check(self._lib.dvg_scene_copy_grads(self.handle, ctypes.c_void_p(block.data_ptr()), stream))
        return block, layout

    def grads(self):
        f = self.flat
        out = dict(
            shape_params=np.zeros(max(f['num_shape_params'], 1), np.float32),
            points=np.zeros(max(2 * f['num_points_total'], 1), np.float32),
            thickness=np.zeros(max(f['num_points_total'], 1), np.float32),
            stroke_width=np.zeros(max(f['num_shapes'], 1), np.float32),
            color_params=np.zeros(max(f['num_color_params'], 1), np.float32),
            shape_to_canvas=np.zeros(max(9 * f['num_groups'], 1), np.float32),
            filter_radius=np.zeros(1, np.float32))
        g = Grads(*[out[k].ctypes.data_as(F32P) for k in ('shape_params', 'points', 'thickness', 'stroke_width',
                                                           'color_params', 'shape_to_canvas', 'filter_radius')])
        check(self._lib.dvg_scene_get_grads(self.handle, ctypes.byref(g)))
        out['shape_params'] = out['shape_params'][:f['num_shape_params']]
        out['points'] = out['points'][:2 * f['num_points_total']]
        out['thickness'] = out['thickness'][:f['num_points_total']]
        out['stroke_width'] = out['stroke_width'][:f['num_shapes']]
        out['color_params'] = out['color_params'][:f['num_color_params']]
        out['shape_to_canvas'] = out['shape_to_canvas'][:9 * f['num_groups']]
        return out

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h and _lib is not None:
            _lib.dvg_scene_destroy(h)
            self.handle = None
