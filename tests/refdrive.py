"""Test-side drivers for the two CPU checkers (test infrastructure only).

* ``OracleRunner`` -- ctypes front end of oracle/liboracle.so, the CPU restatement.
* ``RefRunner``    -- drives oracle/_ref/diffvg.so, the reference's own CPU path
  compiled from its sources by oracle/Makefile, through its pybind11 classes, built
  the way pydiffvg/render_pytorch.py:202-368 builds them.

Both take a scene in the pydiffvg object model (diffvg_amd.Circle, Path, ShapeGroup,
...), flattened here independently of the product's own marshalling so that a
marshalling mistake in the product cannot hide behind a shared helper.
"""
import ctypes
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, 'oracle', 'liboracle.so')
REF_DIR = os.path.join(ROOT, 'oracle', '_ref')

FILTERS = {'box': 0, 'tent': 1, 'parabolic': 2, 'hann': 3, 'gaussian': 4}


def _np(x, dtype=np.float32):
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(x, dtype=dtype))


def _filter_type_int(ftype):
    if isinstance(ftype, (int, np.integer)):
        return int(ftype)
    name = getattr(ftype, 'name', str(ftype)).split('.')[-1]
    return FILTERS[name]


def flatten(canvas_width, canvas_height, shapes, shape_groups, filter_type=0, filter_radius=0.5):
    """Scene -> dict of flat arrays in the orc_scene_desc / dvg_scene_desc layout."""
    import diffvg_amd as pd
    n = len(shapes)
    stype = np.zeros(n, np.int32)
    sw = np.zeros(n, np.float32)
    poff = np.full(n, -1, np.int32)
    params = []
    ctrl_off = np.zeros(n, np.int32)
    nbase = np.zeros(n, np.int32)
    pt_off = np.zeros(n, np.int32)
    npts = np.zeros(n, np.int32)
    has_thick = np.zeros(n, np.uint8)
    closed = np.zeros(n, np.uint8)
    approx = np.zeros(n, np.uint8)
    ncp_all, pts_all, thick_all = [], [], []
    n_ctrl = 0
    n_pts = 0
    for i, s in enumerate(shapes):
        if isinstance(s, (pd.Path, pd.Polygon)):
            stype[i] = 2
            pts = _np(s.points).reshape(-1, 2)
            if isinstance(s, pd.Polygon):
                nc = np.zeros(pts.shape[0] if s.is_closed else pts.shape[0] - 1, np.int32)
                use_approx = False
            else:
                nc = _np(s.num_control_points, np.int32).reshape(-1)
                use_approx = bool(s.use_distance_approx)
            w = _np(s.stroke_width).reshape(-1)
            if isinstance(s, pd.Path) and w.shape[0] > 1:
                has_thick[i] = 1
                thick_all.append(w)
                sw[i] = 0.0  # render_pytorch.py:95-98
            else:
                thick_all.append(np.zeros(pts.shape[0], np.float32))
                sw[i] = float(w[0])
            ctrl_off[i] = n_ctrl
            nbase[i] = nc.shape[0]
            pt_off[i] = n_pts
            npts[i] = pts.shape[0]
            closed[i] = 1 if s.is_closed else 0
            approx[i] = 1 if use_approx else 0
            ncp_all.append(nc)
            pts_all.append(pts)
            n_ctrl += nc.shape[0]
            n_pts += pts.shape[0]
            continue
        sw[i] = float(_np(s.stroke_width).reshape(-1)[0])
        poff[i] = sum(len(p) for p in params)
        if isinstance(s, pd.Circle):
            stype[i] = 0
            params.append(np.concatenate([_np(s.radius).reshape(1), _np(s.center).reshape(2)]))
        elif isinstance(s, pd.Ellipse):
            stype[i] = 1
            params.append(np.concatenate([_np(s.radius).reshape(2), _np(s.center).reshape(2)]))
        elif isinstance(s, pd.Rect):
            stype[i] = 3
            params.append(np.concatenate([_np(s.p_min).reshape(2), _np(s.p_max).reshape(2)]))
        else:
            raise TypeError(type(s))
    ng = len(shape_groups)
    g_off = np.zeros(ng, np.int32)
    g_n = np.zeros(ng, np.int32)
    ids_all = []
    ftype = np.full(ng, -1, np.int32)
    foff = np.zeros(ng, np.int32)
    fns = np.zeros(ng, np.int32)
    stype_c = np.full(ng, -1, np.int32)
    soff = np.zeros(ng, np.int32)
    sns = np.zeros(ng, np.int32)
    colors = []
    even_odd = np.zeros(ng, np.uint8)
    xf = np.zeros((ng, 9), np.float32)
    ncol = 0

    def paint(c):
        nonlocal ncol
        if c is None:
            return -1, 0, 0
        off = ncol
        if isinstance(c, pd.LinearGradient):
            arr = np.concatenate([_np(c.begin).reshape(2), _np(c.end).reshape(2),
                                  _np(c.offsets).reshape(-1), _np(c.stop_colors).reshape(-1)])
            t, ns_ = 1, _np(c.offsets).reshape(-1).shape[0]
        elif isinstance(c, pd.RadialGradient):
            arr = np.concatenate([_np(c.center).reshape(2), _np(c.radius).reshape(2),
                                  _np(c.offsets).reshape(-1), _np(c.stop_colors).reshape(-1)])
            t, ns_ = 2, _np(c.offsets).reshape(-1).shape[0]
        else:
            arr = _np(c).reshape(4)
            t, ns_ = 0, 0
        colors.append(arr)
        ncol += arr.shape[0]
        return t, off, ns_

    cur = 0
    for g, grp in enumerate(shape_groups):
        ids = _np(grp.shape_ids, np.int32).reshape(-1)
        g_off[g] = cur
        g_n[g] = ids.shape[0]
        ids_all.append(ids)
        cur += ids.shape[0]
        ftype[g], foff[g], fns[g] = paint(grp.fill_color)
        stype_c[g], soff[g], sns[g] = paint(grp.stroke_color)
        even_odd[g] = 1 if grp.use_even_odd_rule else 0
        xf[g] = _np(grp.shape_to_canvas).reshape(9)

    def cat(lst, dtype):
        return np.ascontiguousarray(np.concatenate(lst).astype(dtype)) if lst else np.zeros(1, dtype)

    return dict(
        canvas_width=int(canvas_width), canvas_height=int(canvas_height), num_shapes=n,
        shape_type=stype, stroke_width=sw, shape_param_offset=poff,
        shape_params=cat(params, np.float32), num_shape_params=int(sum(len(p) for p in params)),
        path_ctrl_offset=ctrl_off, path_num_base_points=nbase, path_point_offset=pt_off,
        path_num_points=npts, path_has_thickness=has_thick, path_is_closed=closed,
        path_use_distance_approx=approx, num_control_points=cat(ncp_all, np.int32), num_ctrl_total=n_ctrl,
        points=cat([p.reshape(-1) for p in pts_all], np.float32), thickness=cat(thick_all, np.float32),
        num_points_total=n_pts, num_groups=ng, group_shape_offset=g_off, group_num_shapes=g_n,
        group_shape_ids=cat(ids_all, np.int32), num_group_shape_ids=cur,
        fill_type=ftype, fill_offset=foff, fill_num_stops=fns,
        stroke_type=stype_c, stroke_offset=soff, stroke_num_stops=sns,
        color_params=cat(colors, np.float32), num_color_params=ncol,
        use_even_odd_rule=even_odd, shape_to_canvas=np.ascontiguousarray(xf.reshape(-1)),
        filter_type=_filter_type_int(filter_type), filter_radius=float(filter_radius))


I32P = ctypes.POINTER(ctypes.c_int32)
F32P = ctypes.POINTER(ctypes.c_float)
U8P = ctypes.POINTER(ctypes.c_uint8)

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


class Desc(ctypes.Structure):
    _fields_ = DESC_FIELDS


def make_desc(flat):
    d = Desc()
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


class RenderArgs(ctypes.Structure):
    _fields_ = [('background_image', F32P), ('render_image', F32P), ('render_sdf', F32P),
                ('width', ctypes.c_int32), ('height', ctypes.c_int32),
                ('num_samples_x', ctypes.c_int32), ('num_samples_y', ctypes.c_int32),
                ('seed', ctypes.c_uint64), ('d_background_image', F32P), ('d_render_image', F32P),
                ('d_render_sdf', F32P), ('d_translation', F32P), ('use_prefiltering', ctypes.c_int32),
                ('eval_positions', F32P), ('num_eval_positions', ctypes.c_int32),
                ('reserved', ctypes.c_void_p)]


class Grads(ctypes.Structure):
    _fields_ = [('d_shape_params', F32P), ('d_points', F32P), ('d_thickness', F32P),
                ('d_stroke_width', F32P), ('d_color_params', F32P), ('d_shape_to_canvas', F32P),
                ('d_filter_radius', F32P)]


def _fp(a):
    return None if a is None else a.ctypes.data_as(F32P)


_oracle = None


def load_oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError('oracle/liboracle.so is not built (run make -C oracle)')
        lib = ctypes.CDLL(ORACLE_SO)
        lib.orc_scene_create.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(ctypes.c_void_p)]
        lib.orc_render.argtypes = [ctypes.c_void_p, ctypes.POINTER(RenderArgs)]
        lib.orc_scene_get_grads.argtypes = [ctypes.c_void_p, ctypes.POINTER(Grads)]
        lib.orc_scene_destroy.argtypes = [ctypes.c_void_p]
        lib.orc_scene_zero_grads.argtypes = [ctypes.c_void_p]
        lib.orc_last_error.restype = ctypes.c_char_p
        lib.orc_set_num_threads.argtypes = [ctypes.c_int]
        lib.orc_pcg32_sample.argtypes = [ctypes.c_int, ctypes.c_uint64, F32P]
        lib.orc_pcg32_sample.restype = ctypes.c_uint32
        lib.orc_solve_cubic_d.argtypes = [ctypes.c_double] * 4 + [ctypes.POINTER(ctypes.c_double)]
        lib.orc_solve_cubic_f.argtypes = [ctypes.c_float] * 4 + [F32P]
        lib.orc_solve_quadratic_f.argtypes = [ctypes.c_float] * 3 + [F32P, F32P]
        lib.orc_filter_weight.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        lib.orc_filter_weight.restype = ctypes.c_float
        lib.orc_d_filter_weight.argtypes = [ctypes.c_int] + [ctypes.c_float] * 4
        lib.orc_d_filter_weight.restype = ctypes.c_float
        for fn in ('orc_winding_number', 'orc_is_inside', 'orc_within_distance'):
            getattr(lib, fn).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float]
        lib.orc_compute_distance.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                             ctypes.c_float, F32P, F32P, I32P, I32P, F32P]
        lib.orc_scene_bvh.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, I32P, F32P, F32P, ctypes.c_int]
        lib.orc_scene_sampling_tables.argtypes = [ctypes.c_void_p, F32P, F32P, F32P, I32P, I32P]
        lib.orc_morton2d.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int]
        lib.orc_morton2d.restype = ctypes.c_uint32
        lib.orc_sample_boundary.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64, F32P]
        lib.orc_set_counting.argtypes = [ctypes.c_int]
        lib.orc_set_sample_stride.argtypes = [ctypes.c_int]
        lib.orc_get_counters.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int64 * 12)]
        lib.orc_slot_samples.argtypes = [ctypes.c_int]
        lib.orc_slot_samples.restype = ctypes.c_int64
        lib.orc_slot_edge_samples_by_shape.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.c_int]
        _oracle = lib
    return _oracle


def grad_buffers(flat):
    return dict(
        shape_params=np.zeros(max(flat['num_shape_params'], 1), np.float32),
        points=np.zeros(max(2 * flat['num_points_total'], 1), np.float32),
        thickness=np.zeros(max(flat['num_points_total'], 1), np.float32),
        stroke_width=np.zeros(max(flat['num_shapes'], 1), np.float32),
        color_params=np.zeros(max(flat['num_color_params'], 1), np.float32),
        shape_to_canvas=np.zeros(max(9 * flat['num_groups'], 1), np.float32),
        filter_radius=np.zeros(1, np.float32))


def trim_grads(flat, g):
    return dict(
        shape_params=g['shape_params'][:flat['num_shape_params']],
        points=g['points'][:2 * flat['num_points_total']],
        thickness=g['thickness'][:flat['num_points_total']],
        stroke_width=g['stroke_width'][:flat['num_shapes']],
        color_params=g['color_params'][:flat['num_color_params']],
        shape_to_canvas=g['shape_to_canvas'][:9 * flat['num_groups']],
        filter_radius=g['filter_radius'])


class OracleRunner:
    """Runs the oracle on one flattened scene.

    One thread by default: with several, the reference's CPU schedule splats samples with
    float atomics in arrival order, so a pixel's sum (64 terms under a Gaussian filter of
    radius 1.5 at 2x2 samples) rounds differently from run to run by up to ~1e-5, the
    size of the image tolerance the GPU results are held to.  With one thread the order
    is the sample order and the oracle's image is the same in every run."""

    def __init__(self, flat, threads=1):
        self.lib = load_oracle()
        self.flat = flat
        self.lib.orc_set_num_threads(threads)
        self.desc, self._keep = make_desc(flat)
        h = ctypes.c_void_p()
        if self.lib.orc_scene_create(ctypes.byref(self.desc), ctypes.byref(h)) != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, 'h', None):
            self.lib.orc_scene_destroy(self.h)
            self.h = None

    def render(self, width, height, nsx, nsy, seed, *, output='color', prefilter=False, background=None,
               d_image=None, d_sdf=None, translation=False, eval_positions=None, d_background=False):
        eval_n = 0 if eval_positions is None else int(eval_positions.shape[0])
        img = None
        sdf = None
        if d_image is None and d_sdf is None:
            if output == 'color':
                img = np.zeros((height, width, 4), np.float32)
            else:
                sdf = np.zeros((eval_n, 1) if eval_n else (height, width, 1), np.float32)
        d_tr = np.zeros((height, width, 2), np.float32) if translation else None
        d_bg = np.zeros((height, width, 4), np.float32) if (d_background and background is not None) else None
        bg = None if background is None else np.ascontiguousarray(background, np.float32)
        ev = None if eval_positions is None else np.ascontiguousarray(eval_positions, np.float32)
        di = None if d_image is None else np.ascontiguousarray(d_image, np.float32)
        ds = None if d_sdf is None else np.ascontiguousarray(d_sdf, np.float32)
        a = RenderArgs(_fp(bg), _fp(img), _fp(sdf), width, height, nsx, nsy, seed, _fp(d_bg), _fp(di), _fp(ds),
                       _fp(d_tr), 1 if prefilter else 0, _fp(ev), eval_n, None)
        self.lib.orc_scene_zero_grads(self.h)
        if self.lib.orc_render(self.h, ctypes.byref(a)) != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())
        out = dict(image=img, sdf=sdf, d_translation=d_tr, d_background=d_bg)
        if d_image is not None or d_sdf is not None:
            g = grad_buffers(self.flat)
            gs = Grads(*[_fp(g[k]) for k in ('shape_params', 'points', 'thickness', 'stroke_width',
                                             'color_params', 'shape_to_canvas', 'filter_radius')])
            self.lib.orc_scene_get_grads(self.h, ctypes.byref(gs))
            out['grads'] = trim_grads(self.flat, g)
        return out


def load_ref(build=None):
    """Import the reference CPU build (oracle/_ref/diffvg.so, or oracle/_ref/<build>/diffvg.so)
    or return None.  Both builds define the same pybind11 module and classes, so one
    process can hold only one of them: run the other with ``ref_render_in_child``."""
    d = REF_DIR if build is None else os.path.join(REF_DIR, build)
    if not os.path.exists(os.path.join(d, 'diffvg.so')):
        return None
    if d not in sys.path:
        sys.path.insert(0, d)
    import diffvg as dv  # noqa: E402  (the reference module, compiled here)
    if os.path.dirname(os.path.abspath(dv.__file__)) != os.path.abspath(d):
        raise RuntimeError('another reference build is already loaded: %s' % dv.__file__)
    return dv


def ref_build_exists(build=None):
    d = REF_DIR if build is None else os.path.join(REF_DIR, build)
    return os.path.exists(os.path.join(d, 'diffvg.so'))


def ref_render_in_child(flat, build, width, height, nsx, nsy, seed, runs):
    """Render with the reference build ``build`` (None = oracle/_ref) in a child Python
    process.  ``runs`` is a list of keyword dicts for RefRunner.render; returns the list
    of their outputs (image, and grads for a backward)."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        arrays = {'flat__' + k: np.asarray(v) for k, v in flat.items()}
        for i, kw in enumerate(runs):
            for k, v in kw.items():
                arrays['run%d__%s' % (i, k)] = np.asarray(v)
        np.savez(os.path.join(td, 'in.npz'), **arrays)
        cmd = [sys.executable, os.path.abspath(__file__), os.path.join(td, 'in.npz'), os.path.join(td, 'out.npz'),
               '' if build is None else build, str(width), str(height), str(nsx), str(nsy), str(seed), str(len(runs))]
        subprocess.run(cmd, check=True)
        with np.load(os.path.join(td, 'out.npz')) as z:
            out = [dict() for _ in runs]
            for k in z.files:
                i, rest = k.split('__', 1)
                i = int(i[3:])
                if rest.startswith('grad_'):
                    out[i].setdefault('grads', {})[rest[5:]] = z[k]
                else:
                    out[i][rest] = z[k]
        return out


def _child_main(argv):
    src, dst, build, w, h, nsx, nsy, seed, nruns = argv
    w, h, nsx, nsy, seed, nruns = int(w), int(h), int(nsx), int(nsy), int(seed), int(nruns)
    with np.load(src) as z:
        flat = {}
        runs = [dict() for _ in range(nruns)]
        for k in z.files:
            head, name = k.split('__', 1)
            v = z[k]
            if v.ndim == 0:
                v = v.item()
            if head == 'flat':
                flat[name] = v
            else:
                runs[int(head[3:])][name] = v
    dv = load_ref(build or None)
    if dv is None:
        raise SystemExit('reference build %r not found' % build)
    r = RefRunner(flat, dv=dv)
    res = {}
    for i, kw in enumerate(runs):
        o = r.render(w, h, nsx, nsy, seed, **kw)
        for k in ('image', 'sdf'):
            if o.get(k) is not None:
                res['run%d__%s' % (i, k)] = o[k]
        for k, v in o.get('grads', {}).items():
            res['run%d__grad_%s' % (i, k)] = v
    np.savez(dst, **res)


class RefRunner:
    """Runs the reference CPU build on one flattened scene (builds pybind objects)."""

    def __init__(self, flat, dv=None, use_gpu=False, gpu_index=-1):
        """dv: the module to drive (default: the reference build); the product's
        drop-in shim diffvg_amd.diffvg is driven the same way with use_gpu=True."""
        if dv is None:
            dv = load_ref()
        if dv is None:
            raise RuntimeError('reference CPU build not available')
        self.dv = dv
        self.use_gpu, self.gpu_index = use_gpu, gpu_index
        self.flat = flat
        f = flat
        keep = []
        shapes = []
        self.shape_objs = []
        for i in range(f['num_shapes']):
            t = int(f['shape_type'][i])
            sw = float(f['stroke_width'][i])
            if t == 2:
                c0, nb = int(f['path_ctrl_offset'][i]), int(f['path_num_base_points'][i])
                p0, npt = int(f['path_point_offset'][i]), int(f['path_num_points'][i])
                ncp = np.ascontiguousarray(f['num_control_points'][c0:c0 + nb], np.int32)
                pts = np.ascontiguousarray(f['points'][2 * p0:2 * (p0 + npt)], np.float32)
                th = np.ascontiguousarray(f['thickness'][p0:p0 + npt], np.float32) if f['path_has_thickness'][i] \
                    else None
                keep += [ncp, pts, th]
                obj = dv.Path(dv.int_ptr(ncp.ctypes.data), dv.float_ptr(pts.ctypes.data),
                              dv.float_ptr(th.ctypes.data if th is not None else 0), nb, npt,
                              bool(f['path_is_closed'][i]), bool(f['path_use_distance_approx'][i]))
                st = dv.ShapeType.path
            else:
                o = int(f['shape_param_offset'][i])
                P = f['shape_params']
                if t == 0:
                    obj = dv.Circle(float(P[o]), dv.Vector2f(float(P[o + 1]), float(P[o + 2])))
                    st = dv.ShapeType.circle
                elif t == 1:
                    obj = dv.Ellipse(dv.Vector2f(float(P[o]), float(P[o + 1])),
                                     dv.Vector2f(float(P[o + 2]), float(P[o + 3])))
                    st = dv.ShapeType.ellipse
                else:
                    obj = dv.Rect(dv.Vector2f(float(P[o]), float(P[o + 1])),
                                  dv.Vector2f(float(P[o + 2]), float(P[o + 3])))
                    st = dv.ShapeType.rect
            self.shape_objs.append(obj)
            shapes.append(dv.Shape(st, obj.get_ptr(), sw))
        groups = []
        self.color_objs = []
        ctypes_map = {0: dv.ColorType.constant, 1: dv.ColorType.linear_gradient, 2: dv.ColorType.radial_gradient}
        C = f['color_params']

        def paint(t, off, ns_):
            if t < 0:
                return dv.ColorType.constant, dv.void_ptr(0)
            if t == 0:
                obj = dv.Constant(dv.Vector4f(*[float(x) for x in C[off:off + 4]]))
            else:
                offs = np.ascontiguousarray(C[off + 4:off + 4 + ns_], np.float32)
                cols = np.ascontiguousarray(C[off + 4 + ns_:off + 4 + 5 * ns_], np.float32)
                keep.extend([offs, cols])
                a = dv.Vector2f(float(C[off]), float(C[off + 1]))
                b = dv.Vector2f(float(C[off + 2]), float(C[off + 3]))
                cls = dv.LinearGradient if t == 1 else dv.RadialGradient
                obj = cls(a, b, ns_, dv.float_ptr(offs.ctypes.data), dv.float_ptr(cols.ctypes.data))
            self.color_objs.append(obj)
            return ctypes_map[t], obj.get_ptr()

        for g in range(f['num_groups']):
            o, n = int(f['group_shape_offset'][g]), int(f['group_num_shapes'][g])
            ids = np.ascontiguousarray(f['group_shape_ids'][o:o + n], np.int32)
            m = np.ascontiguousarray(f['shape_to_canvas'][9 * g:9 * g + 9], np.float32)
            keep += [ids, m]
            ft, fp = paint(int(f['fill_type'][g]), int(f['fill_offset'][g]), int(f['fill_num_stops'][g]))
            st_, sp_ = paint(int(f['stroke_type'][g]), int(f['stroke_offset'][g]), int(f['stroke_num_stops'][g]))
            groups.append(dv.ShapeGroup(dv.int_ptr(ids.ctypes.data), n, ft, fp, st_, sp_,
                                        bool(f['use_even_odd_rule'][g]), dv.float_ptr(m.ctypes.data)))
        ft_enum = [dv.FilterType.box, dv.FilterType.tent, dv.FilterType.parabolic, dv.FilterType.hann]
        if f['filter_type'] >= 4 and not hasattr(dv.FilterType, 'gaussian'):
            raise ValueError('the reference has no Gaussian filter')
        if f['filter_type'] >= 4:
            ft_enum.append(dv.FilterType.gaussian)
        self.filt = dv.Filter(ft_enum[f['filter_type']], float(f['filter_radius']))
        self.shapes = shapes
        self.groups = groups
        self._keep = keep
        self.scene = dv.Scene(f['canvas_width'], f['canvas_height'], shapes, groups, self.filt, use_gpu, gpu_index)

    def render(self, width, height, nsx, nsy, seed, *, output='color', prefilter=False, background=None,
               d_image=None, d_sdf=None, translation=False, eval_positions=None, d_background=False):
        dv = self.dv
        eval_n = 0 if eval_positions is None else int(eval_positions.shape[0])
        img = None
        sdf = None
        if d_image is None and d_sdf is None:
            if output == 'color':
                img = np.zeros((height, width, 4), np.float32)
            else:
                sdf = np.zeros((eval_n, 1) if eval_n else (height, width, 1), np.float32)
        d_tr = np.zeros((height, width, 2), np.float32) if translation else None
        d_bg = np.zeros((height, width, 4), np.float32) if (d_background and background is not None) else None
        bg = None if background is None else np.ascontiguousarray(background, np.float32)
        ev = None if eval_positions is None else np.ascontiguousarray(eval_positions, np.float32)
        di = None if d_image is None else np.ascontiguousarray(d_image, np.float32)
        ds = None if d_sdf is None else np.ascontiguousarray(d_sdf, np.float32)

        def P(a):
            return dv.float_ptr(0 if a is None else a.ctypes.data)

        # A fresh scene per call so that the gradient buffers start at zero.
        if d_image is not None or d_sdf is not None:
            self.scene = dv.Scene(self.flat['canvas_width'], self.flat['canvas_height'], self.shapes, self.groups,
                                  self.filt, self.use_gpu, self.gpu_index)
        dv.render(self.scene, P(bg), P(img), P(sdf), width, height, nsx, nsy, seed, P(d_bg), P(di), P(ds),
                  P(d_tr), bool(prefilter), P(ev), eval_n)
        out = dict(image=img, sdf=sdf, d_translation=d_tr, d_background=d_bg)
        if d_image is not None or d_sdf is not None:
            out['grads'] = self._grads()
        return out

    def _grads(self):
        dv = self.dv
        f = self.flat
        g = {k: v.copy() for k, v in trim_grads(f, grad_buffers(f)).items()}
        sc = self.scene
        for i in range(f['num_shapes']):
            d = sc.get_d_shape(i)
            t = int(f['shape_type'][i])
            o = int(f['shape_param_offset'][i])
            if t == 0:
                c = d.as_circle()
                g['shape_params'][o:o + 3] = [c.radius, c.center.x, c.center.y]
            elif t == 1:
                e = d.as_ellipse()
                g['shape_params'][o:o + 4] = [e.radius.x, e.radius.y, e.center.x, e.center.y]
            elif t == 3:
                r = d.as_rect()
                g['shape_params'][o:o + 4] = [r.p_min.x, r.p_min.y, r.p_max.x, r.p_max.y]
            else:
                p = d.as_path()
                p0, npt = int(f['path_point_offset'][i]), int(f['path_num_points'][i])
                pts = np.zeros(2 * npt, np.float32)
                th = np.zeros(npt, np.float32)
                p.copy_to(dv.float_ptr(pts.ctypes.data), dv.float_ptr(th.ctypes.data if p.has_thickness() else 0))
                g['points'][2 * p0:2 * (p0 + npt)] = pts
                if p.has_thickness():
                    g['thickness'][p0:p0 + npt] = th
            g['stroke_width'][i] = d.stroke_width
        C = f['color_params']
        for gi in range(f['num_groups']):
            d = sc.get_d_shape_group(gi)
            for kind in ('fill', 'stroke'):
                t = int(f[kind + '_type'][gi])
                off = int(f[kind + '_offset'][gi])
                ns_ = int(f[kind + '_num_stops'][gi])
                if t < 0:
                    continue
                if t == 0:
                    c = getattr(d, kind + '_color_as_constant')().color
                    g['color_params'][off:off + 4] = [c.x, c.y, c.z, c.w]
                    continue
                if kind == 'stroke':
                    # Gradient stroke adjoints are not retrievable from the reference: its d-buffer
                    # pointer for them is never set (scene.cpp:866-869, 885-888).
                    g['color_params'][off:off + 4 + 5 * ns_] = np.nan
                    continue
                lg = d.fill_color_as_linear_gradient() if t == 1 else d.fill_color_as_radial_gradient()
                offs = np.zeros(ns_, np.float32)
                cols = np.zeros(4 * ns_, np.float32)
                lg.copy_to(dv.float_ptr(offs.ctypes.data), dv.float_ptr(cols.ctypes.data))
                a, b = (lg.begin, lg.end) if t == 1 else (lg.center, lg.radius)
                g['color_params'][off:off + 4] = [a.x, a.y, b.x, b.y]
                g['color_params'][off + 4:off + 4 + ns_] = offs
                g['color_params'][off + 4 + ns_:off + 4 + 5 * ns_] = cols
            m = np.zeros(9, np.float32)
            d.copy_to(dv.float_ptr(m.ctypes.data))
            g['shape_to_canvas'][9 * gi:9 * gi + 9] = m
        g['filter_radius'] = np.array([sc.get_d_filter_radius()], np.float32)
        return g


if __name__ == '__main__':
    _child_main(sys.argv[1:])
