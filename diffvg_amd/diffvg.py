"""Drop-in counterpart of the reference's native ``diffvg`` module.

Same names and signatures as the pybind11 module (reference diffvg.cpp:1651-1792):
pointer wrappers, vectors, shapes, paints, filters, ``Scene`` and ``render`` -- backed
by libdiffvg_amd.so.  As in the reference, scene objects hold raw host addresses of
caller-owned arrays; ``Scene`` copies them to the GPU.  Image pointers passed to
``render`` are device addresses.  ``FilterType.gaussian`` is an extension.
"""
import ctypes
import weakref
from enum import IntEnum

import numpy as np

from . import _lib


class ShapeType(IntEnum):
    circle = 0
    ellipse = 1
    path = 2
    rect = 3


class ColorType(IntEnum):
    constant = 0
    linear_gradient = 1
    radial_gradient = 2


class FilterType(IntEnum):
    box = 0
    tent = 1
    parabolic = 2
    hann = 3
    gaussian = 4


class _Ptr:
    def __init__(self, address):
        self.address = int(address)

    def as_size_t(self):
        return self.address


class void_ptr(_Ptr):
    pass


class float_ptr(_Ptr):
    pass


class int_ptr(_Ptr):
    pass


class Vector2f:
    def __init__(self, x, y):
        self.x = float(x)
        self.y = float(y)


class Vector3f:
    def __init__(self, x, y, z):
        self.x, self.y, self.z = float(x), float(y), float(z)


class Vector4f:
    def __init__(self, x, y, z, w):
        self.x, self.y, self.z, self.w = float(x), float(y), float(z), float(w)


def _floats(address, n):
    if n == 0:
        return np.zeros(0, np.float32)
    return np.ctypeslib.as_array((ctypes.c_float * n).from_address(address)).copy()


def _ints(address, n):
    if n == 0:
        return np.zeros(0, np.int32)
    return np.ctypeslib.as_array((ctypes.c_int32 * n).from_address(address)).copy()


# get_ptr() hands out the object's id and remembers the object, so that Shape and
# ShapeGroup can resolve it as the reference resolves the raw pointer.  As in the
# reference, the caller keeps the object alive while it is in use.
_objects = weakref.WeakValueDictionary()


class _Value:
    def get_ptr(self):
        _objects[id(self)] = self
        return void_ptr(id(self))


class Circle(_Value):
    def __init__(self, radius, center):
        self.radius = float(radius)
        self.center = center


class Ellipse(_Value):
    def __init__(self, radius, center):
        self.radius = radius
        self.center = center


class Rect(_Value):
    def __init__(self, p_min, p_max):
        self.p_min = p_min
        self.p_max = p_max


class Path(_Value):
    def __init__(self, num_control_points, points, thickness, num_base_points, num_points, is_closed,
                 use_distance_approx):
        self.num_control_points = _ints(num_control_points.address, num_base_points)
        self.points = _floats(points.address, 2 * num_points)
        self.thickness = _floats(thickness.address, num_points) if thickness.address else None
        self.num_base_points = int(num_base_points)
        self.num_points = int(num_points)
        self.is_closed = bool(is_closed)
        self.use_distance_approx = bool(use_distance_approx)

    def has_thickness(self):
        return self.thickness is not None

    def copy_to(self, points, thickness):
        ctypes.memmove(points.address, self.points.ctypes.data, 4 * self.points.size)
        if self.thickness is not None and thickness.address:
            ctypes.memmove(thickness.address, self.thickness.ctypes.data, 4 * self.thickness.size)


class Constant(_Value):
    def __init__(self, color):
        self.color = color


class LinearGradient(_Value):
    def __init__(self, begin, end, num_stops, stop_offsets, stop_colors):
        self.begin, self.end, self.num_stops = begin, end, int(num_stops)
        self.stop_offsets = _floats(stop_offsets.address, num_stops)
        self.stop_colors = _floats(stop_colors.address, 4 * num_stops)

    def copy_to(self, stop_offsets, stop_colors):
        ctypes.memmove(stop_offsets.address, self.stop_offsets.ctypes.data, 4 * self.num_stops)
        ctypes.memmove(stop_colors.address, self.stop_colors.ctypes.data, 16 * self.num_stops)


class RadialGradient(LinearGradient):
    def __init__(self, center, radius, num_stops, stop_offsets, stop_colors):
        super().__init__(center, radius, num_stops, stop_offsets, stop_colors)
        self.center, self.radius = center, radius


class Filter:
    def __init__(self, type, radius):
        self.type = FilterType(int(type))
        self.radius = float(radius)


class Shape:
    def __init__(self, type, ptr, stroke_width):
        self.type = ShapeType(int(type))
        self.obj = _objects.get(ptr.address) if isinstance(ptr, _Ptr) else ptr
        self.stroke_width = float(stroke_width)

    def as_circle(self):
        return self.obj

    def as_ellipse(self):
        return self.obj

    def as_path(self):
        return self.obj

    def as_rect(self):
        return self.obj


class ShapeGroup:
    def __init__(self, shape_ids, num_shapes, fill_color_type, fill_color, stroke_color_type, stroke_color,
                 use_even_odd_rule, shape_to_canvas):
        self.shape_ids = _ints(shape_ids.address, num_shapes)
        self.num_shapes = int(num_shapes)
        self.fill_color_type = ColorType(int(fill_color_type))
        self.fill_color = _objects.get(fill_color.address) if fill_color.address else None
        self.stroke_color_type = ColorType(int(stroke_color_type))
        self.stroke_color = _objects.get(stroke_color.address) if stroke_color.address else None
        self.use_even_odd_rule = bool(use_even_odd_rule)
        self.shape_to_canvas = _floats(shape_to_canvas.address, 9)

    def has_fill_color(self):
        return self.fill_color is not None

    def has_stroke_color(self):
        return self.stroke_color is not None

    def fill_color_as_constant(self):
        return self.fill_color

    def fill_color_as_linear_gradient(self):
        return self.fill_color

    def fill_color_as_radial_gradient(self):
        return self.fill_color

    def stroke_color_as_constant(self):
        return self.stroke_color

    def stroke_color_as_linear_gradient(self):
        return self.stroke_color

    def stroke_color_as_radial_gradient(self):
        return self.stroke_color

    def copy_to(self, shape_to_canvas):
        ctypes.memmove(shape_to_canvas.address, self.shape_to_canvas.ctypes.data, 36)


def _flat_from_objects(canvas_width, canvas_height, shapes, groups, filt):
    ns = len(shapes)
    shape_type = np.zeros(ns, np.int32)
    sw = np.zeros(ns, np.float32)
    poff = np.full(ns, -1, np.int32)
    params, ncp, pts, thick = [], [], [], []
    ctrl_off = np.zeros(ns, np.int32)
    nbase = np.zeros(ns, np.int32)
    pt_off = np.zeros(ns, np.int32)
    npts = np.zeros(ns, np.int32)
    has_thick = np.zeros(ns, np.uint8)
    closed = np.zeros(ns, np.uint8)
    approx = np.zeros(ns, np.uint8)
    n_params = n_ctrl = n_pts = 0
    for i, sh in enumerate(shapes):
        shape_type[i] = int(sh.type)
        sw[i] = sh.stroke_width
        o = sh.obj
        if sh.type == ShapeType.path:
            ctrl_off[i], nbase[i], pt_off[i], npts[i] = n_ctrl, o.num_base_points, n_pts, o.num_points
            has_thick[i] = 1 if o.thickness is not None else 0
            closed[i] = 1 if o.is_closed else 0
            approx[i] = 1 if o.use_distance_approx else 0
            ncp.append(o.num_control_points)
            pts.append(o.points)
            thick.append(o.thickness if o.thickness is not None else np.zeros(o.num_points, np.float32))
            n_ctrl += o.num_base_points
            n_pts += o.num_points
            continue
        poff[i] = n_params
        if sh.type == ShapeType.circle:
            p = [o.radius, o.center.x, o.center.y]
        elif sh.type == ShapeType.ellipse:
            p = [o.radius.x, o.radius.y, o.center.x, o.center.y]
        else:
            p = [o.p_min.x, o.p_min.y, o.p_max.x, o.p_max.y]
        params.append(np.array(p, np.float32))
        n_params += len(p)
    ng = len(groups)
    g_off = np.zeros(ng, np.int32)
    g_n = np.zeros(ng, np.int32)
    ids = []
    paint = {k: np.zeros(ng, np.int32) for k in ('ft', 'fo', 'fn', 'st', 'so', 'sn')}
    colors = []
    even_odd = np.zeros(ng, np.uint8)
    xf = np.zeros((ng, 9), np.float32)
    n_ids = n_col = 0
    for g, grp in enumerate(groups):
        g_off[g], g_n[g] = n_ids, grp.num_shapes
        ids.append(grp.shape_ids)
        n_ids += grp.num_shapes
        for key, ct, c in (('f', grp.fill_color_type, grp.fill_color), ('s', grp.stroke_color_type, grp.stroke_color)):
            if c is None:
                paint[key + 't'][g] = -1
                continue
            paint[key + 't'][g] = int(ct)
            paint[key + 'o'][g] = n_col
            if ct == ColorType.constant:
                arr = np.array([c.color.x, c.color.y, c.color.z, c.color.w], np.float32)
            else:
                a, b = (c.begin, c.end)
                arr = np.concatenate([np.array([a.x, a.y, b.x, b.y], np.float32), c.stop_offsets, c.stop_colors])
                paint[key + 'n'][g] = c.num_stops
            colors.append(arr)
            n_col += arr.size
        even_odd[g] = 1 if grp.use_even_odd_rule else 0
        xf[g] = grp.shape_to_canvas

    def cat(lst, dt):
        return np.ascontiguousarray(np.concatenate(lst).astype(dt)) if lst else np.zeros(1, dt)

    return dict(canvas_width=int(canvas_width), canvas_height=int(canvas_height), num_shapes=ns,
                shape_type=shape_type, stroke_width=sw, shape_param_offset=poff, shape_params=cat(params, np.float32),
                num_shape_params=n_params, path_ctrl_offset=ctrl_off, path_num_base_points=nbase,
                path_point_offset=pt_off, path_num_points=npts, path_has_thickness=has_thick, path_is_closed=closed,
                path_use_distance_approx=approx, num_control_points=cat(ncp, np.int32), num_ctrl_total=n_ctrl,
                points=cat(pts, np.float32), thickness=cat(thick, np.float32), num_points_total=n_pts,
                num_groups=ng, group_shape_offset=g_off, group_num_shapes=g_n, group_shape_ids=cat(ids, np.int32),
                num_group_shape_ids=n_ids, fill_type=paint['ft'], fill_offset=paint['fo'], fill_num_stops=paint['fn'],
                stroke_type=paint['st'], stroke_offset=paint['so'], stroke_num_stops=paint['sn'],
                color_params=cat(colors, np.float32), num_color_params=n_col, use_even_odd_rule=even_odd,
                shape_to_canvas=np.ascontiguousarray(xf.reshape(-1)), filter_type=int(filt.type),
                filter_radius=float(filt.radius))


class Scene:
    """Scene(canvas_width, canvas_height, shapes, shape_groups, filter, use_gpu, gpu_index)."""

    def __init__(self, canvas_width, canvas_height, shapes, shape_groups, filter, use_gpu, gpu_index):
        if not use_gpu:
            raise RuntimeError('diffvg_amd renders on the GPU only (use_gpu=False is not supported)')
        self.num_shapes = len(shapes)
        self.num_shape_groups = len(shape_groups)
        self._shapes = shapes
        self._groups = shape_groups
        self.flat = _flat_from_objects(canvas_width, canvas_height, shapes, shape_groups, filter)
        self._scene = _lib.DeviceScene(self.flat, gpu_index)
        self._grads = None

    def _g(self):
        if self._grads is None:
            self._grads = self._scene.grads()
        return self._grads

    def get_d_shape(self, shape_id):
        g = self._g()
        sh = self._shapes[shape_id]
        f = self.flat
        o = int(f['shape_param_offset'][shape_id])
        if sh.type == ShapeType.circle:
            obj = Circle(g['shape_params'][o], Vector2f(g['shape_params'][o + 1], g['shape_params'][o + 2]))
        elif sh.type == ShapeType.ellipse:
            obj = Ellipse(Vector2f(*g['shape_params'][o:o + 2]), Vector2f(*g['shape_params'][o + 2:o + 4]))
        elif sh.type == ShapeType.rect:
            obj = Rect(Vector2f(*g['shape_params'][o:o + 2]), Vector2f(*g['shape_params'][o + 2:o + 4]))
        else:
            p0, n = int(f['path_point_offset'][shape_id]), int(f['path_num_points'][shape_id])
            pts = np.ascontiguousarray(g['points'][2 * p0:2 * (p0 + n)])
            th = np.ascontiguousarray(g['thickness'][p0:p0 + n]) if f['path_has_thickness'][shape_id] else None
            obj = Path.__new__(Path)
            obj.points, obj.thickness, obj.num_points = pts, th, n
            obj.num_base_points = int(f['path_num_base_points'][shape_id])
            obj.num_control_points = None
            obj.is_closed = bool(f['path_is_closed'][shape_id])
            obj.use_distance_approx = bool(f['path_use_distance_approx'][shape_id])
        d = Shape.__new__(Shape)
        d.type, d.obj = sh.type, obj
        d.stroke_width = float(g['stroke_width'][shape_id])
        return d

    def get_d_shape_group(self, group_id):
        g = self._g()
        f = self.flat
        grp = self._groups[group_id]
        d = ShapeGroup.__new__(ShapeGroup)
        d.num_shapes = grp.num_shapes
        d.shape_ids = grp.shape_ids
        d.use_even_odd_rule = grp.use_even_odd_rule
        d.fill_color_type, d.stroke_color_type = grp.fill_color_type, grp.stroke_color_type
        for key, attr in (('fill', 'fill_color'), ('stroke', 'stroke_color')):
            t = int(f[key + '_type'][group_id])
            if t < 0:
                setattr(d, attr, None)
                continue
            off, ns = int(f[key + '_offset'][group_id]), int(f[key + '_num_stops'][group_id])
            c = g['color_params']
            if t == 0:
                setattr(d, attr, Constant(Vector4f(*c[off:off + 4])))
            else:
                obj = LinearGradient.__new__(LinearGradient if t == 1 else RadialGradient)
                obj.begin, obj.end = Vector2f(*c[off:off + 2]), Vector2f(*c[off + 2:off + 4])
                obj.center, obj.radius = obj.begin, obj.end
                obj.num_stops = ns
                obj.stop_offsets = np.ascontiguousarray(c[off + 4:off + 4 + ns])
                obj.stop_colors = np.ascontiguousarray(c[off + 4 + ns:off + 4 + 5 * ns])
                setattr(d, attr, obj)
        d.shape_to_canvas = np.ascontiguousarray(g['shape_to_canvas'][9 * group_id:9 * group_id + 9])
        return d

    def get_d_filter_radius(self):
        return float(self._g()['filter_radius'][0])


class _Dev:
    """Minimal stand-in exposing data_ptr() for raw device addresses."""

    def __init__(self, address):
        self.address = address

    def data_ptr(self):
        return self.address


def render(scene, background_image, render_image, render_sdf, width, height, num_samples_x, num_samples_y, seed,
           d_background_image, d_render_image, d_render_sdf, d_translation, use_prefiltering, eval_positions,
           num_eval_positions):
    """render(...) with the reference's argument list (diffvg.cpp:1477-1492); image
    pointers are device addresses on the scene's GPU."""
    def P(p):
        return _Dev(p.address) if p is not None and p.address else None
    scene._grads = None
    scene._scene.render(width, height, num_samples_x, num_samples_y, seed, background=P(background_image),
                        render_image=P(render_image), render_sdf=P(render_sdf), d_background=P(d_background_image),
                        d_render_image=P(d_render_image), d_render_sdf=P(d_render_sdf),
                        d_translation=P(d_translation), use_prefiltering=use_prefiltering,
                        eval_positions=P(eval_positions), num_eval_positions=num_eval_positions)
    # the reference's render returns when the device is done (diffvg.cpp:1639-1641)
    scene._scene.status()
