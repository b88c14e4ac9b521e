"""PyTorch front end with the interface of pydiffvg/render_pytorch.py.

``RenderFunction.serialize_scene`` produces the reference's positional argument list
(render_pytorch.py:22-172) and ``RenderFunction.apply(width, height, num_samples_x,
num_samples_y, seed, background_image, *args)`` renders it (render_pytorch.py:174-428);
``backward`` returns gradients in the reference's order (render_pytorch.py:668-868) and
``render_grad`` returns the screen-space translation gradient (render_pytorch.py:430-666).

Differences from the reference, all deliberate:
* scene tensors are not copied to the CPU one by one; the forward gathers them with one
  concatenation per kind and one host copy, and gradients come back on the device of
  the corresponding input;
* rendering runs on the GPU only (``set_use_gpu(False)`` makes ``apply`` raise);
* the backward clears the scene's gradient buffers before it runs, so a second backward
  through the same graph does not add to the first;
* ``render_grad`` with a background image works (the reference reads an undefined
  variable there, render_pytorch.py:638).
"""
import itertools
import os
import operator
import time
import weakref
import warnings
from enum import IntEnum

import numpy as np
import torch

from .device import get_device, get_use_gpu
from . import diffvg
from . import _lib
from .color import LinearGradient, RadialGradient
from .pixel_filter import PixelFilter
from .shape import Circle, Ellipse, Path, Polygon, Rect

print_timing = False

_render_mode = _lib.MODE_TILED


def set_print_timing(val):
    global print_timing
    print_timing = val


def set_render_mode(mode):
    """Selects the TILED (default) or PLAIN render schedule for RenderFunction."""
    global _render_mode
    if isinstance(mode, str):
        mode = {'tiled': _lib.MODE_TILED, 'plain': _lib.MODE_PLAIN}[mode]
    _render_mode = int(mode)


def get_render_mode():
    return _render_mode


_direct_leaf_grads = False


def set_direct_leaf_gradients(enabled):
    """Opt-in delivery of scene gradients straight into the ``.grad`` of the argument
    tensors, bypassing the autograd engine's per-tensor accumulation.

    With it on, a render whose float argument tensors that require a gradient are all
    leaves (no grad_fn), each used once, float32, on the render device and without
    gradient hooks is recorded as one autograd node without edges to those tensors; its
    backward writes each tensor's gradient (a view of the scene's gradient block) into
    ``.grad``, adding to a gradient already there, as ``loss.backward()`` would.  The
    saving is the engine's per-tensor cost, which dominates the host time of scenes of
    thousands of shapes.  What changes: ``torch.autograd.grad`` and
    ``backward(inputs=...)`` cannot name those tensors (they are not in the graph and
    the engine says so), and the gradients are not taken when a render's argument
    tensors fail the conditions above -- such a render goes through the engine as
    usual.  Off by default."""
    global _direct_leaf_grads
    _direct_leaf_grads = bool(enabled)


class OutputType(IntEnum):
    color = 1
    sdf = 2


def _filter_int(t):
    return int(getattr(t, 'value', t))


class _Flat:
    """Flattened scene: the dvg_scene_desc arrays plus the map from argument positions
    to gradient slices."""


def _cat_kind(tensors):
    """Concatenates same-kind tensors into one flat float32 tensor with few torch calls
    (a per-tensor reshape of an autograd tensor costs microseconds, and scenes hold
    thousands of them)."""
    if not tensors:
        return None
    try:
        return _cat_same_device(tensors)
    except RuntimeError:
        # tensors of one kind on several devices: gather them on the host
        return _cat_same_device([t.detach().cpu() for t in tensors])


def _cat_same_device(tensors):
    ranks = {t.dim() for t in tensors}
    if ranks == {0}:
        out = torch.stack(tensors)
    elif ranks == {1}:
        out = torch.cat(tensors)
    elif len(ranks) == 1 and len({tuple(t.shape[1:]) for t in tensors}) == 1:
        out = torch.cat(tensors).reshape(-1)
    else:
        out = torch.cat([t.reshape(-1) for t in tensors])
    return out if out.dtype == torch.float32 else out.to(torch.float32)


def _gather_kinds(kinds):
    """List of tensor lists -> list of host float32 arrays (one device-to-host copy when
    all tensors share a device)."""
    with torch.no_grad():
        parts = [_cat_kind(k) for k in kinds]
        present = [p for p in parts if p is not None]
        if not present:
            return [np.zeros(0, np.float32) for _ in kinds]
        if len({p.device for p in present}) == 1:
            host = torch.cat(present).cpu().numpy()
        else:
            host = np.concatenate([p.cpu().numpy() for p in present])
    out = []
    k = 0
    for p in parts:
        n = 0 if p is None else int(p.numel())
        out.append(host[k:k + n])
        k += n
    return out


def _cat_int(tensors):
    if not tensors:
        return np.zeros(0, np.int32)
    with torch.no_grad():
        flat = [t.reshape(-1) for t in tensors] if any(t.dim() != 1 for t in tensors) else tensors
        if len({t.device for t in flat}) > 1:
            # index tensors on several devices (shape ids made on the host, control
            # point counts moved to the GPU): gather them on the host
            flat = [t.cpu() for t in flat]
        return torch.cat(flat).to(torch.int32).cpu().numpy()


def flatten_args(args):
    """serialize_scene argument list -> (_Flat, canvas info)."""
    canvas_width, canvas_height, num_shapes, num_shape_groups = args[0], args[1], args[2], args[3]
    output_type, use_prefiltering, eval_positions = args[4], args[5], args[6]
    f = _Flat()
    f.canvas = (int(canvas_width), int(canvas_height))
    f.output_type = output_type
    f.use_prefiltering = bool(use_prefiltering)
    f.eval_positions = eval_positions
    i = 7
    ns = num_shapes
    shape_type = np.zeros(ns, np.int32)
    param_off = np.full(ns, -1, np.int32)
    ctrl_off = np.zeros(ns, np.int32)
    nbase = np.zeros(ns, np.int32)
    pt_off = np.zeros(ns, np.int32)
    npts = np.zeros(ns, np.int32)
    has_thick = np.zeros(ns, np.uint8)
    closed = np.zeros(ns, np.uint8)
    approx = np.zeros(ns, np.uint8)
    # float sources in description order, per kind, and their argument positions
    src_params, src_points, src_thick, src_sw, src_color, src_xform = [], [], [], [], [], []
    pos = {k: [] for k in ('params', 'points', 'thick', 'sw', 'color', 'xform', 'ncp', 'ids')}
    thick_fill = []  # (offset, count) of zero thickness ranges
    ncp_list = []
    grad_map = []  # (arg index, kind, offset, count, shape)
    n_params = n_points = n_ctrl = 0
    for s in range(ns):
        t = args[i]
        i += 1
        if t == diffvg.ShapeType.circle:
            radius, center = args[i], args[i + 1]
            shape_type[s] = 0
            param_off[s] = n_params
            src_params += [radius, center]
            pos['params'] += [i, i + 1]
            grad_map.append((i, 'shape_params', n_params, 1, tuple(radius.shape)))
            grad_map.append((i + 1, 'shape_params', n_params + 1, 2, tuple(center.shape)))
            n_params += 3
            i += 2
        elif t == diffvg.ShapeType.ellipse:
            radius, center = args[i], args[i + 1]
            shape_type[s] = 1
            param_off[s] = n_params
            src_params += [radius, center]
            pos['params'] += [i, i + 1]
            grad_map.append((i, 'shape_params', n_params, 2, tuple(radius.shape)))
            grad_map.append((i + 1, 'shape_params', n_params + 2, 2, tuple(center.shape)))
            n_params += 4
            i += 2
        elif t == diffvg.ShapeType.path:
            ncp, points, thickness, is_closed, use_approx = args[i], args[i + 1], args[i + 2], args[i + 3], args[i + 4]
            n = int(points.shape[0])
            nb = int(ncp.shape[0])
            shape_type[s] = 2
            ctrl_off[s] = n_ctrl
            nbase[s] = nb
            pt_off[s] = n_points
            npts[s] = n
            closed[s] = 1 if is_closed else 0
            approx[s] = 1 if use_approx else 0
            ncp_list.append(ncp)
            src_points.append(points)
            pos['ncp'].append(i)
            pos['points'].append(i + 1)
            grad_map.append((i + 1, 'points', 2 * n_points, 2 * n, tuple(points.shape)))
            if thickness is not None:
                has_thick[s] = 1
                src_thick.append((n_points, thickness))
                pos['thick'].append(i + 2)
                grad_map.append((i + 2, 'thickness', n_points, n, tuple(thickness.shape)))
            n_points += n
            n_ctrl += nb
            i += 5
        elif t == diffvg.ShapeType.rect:
            p_min, p_max = args[i], args[i + 1]
            shape_type[s] = 3
            param_off[s] = n_params
            src_params += [p_min, p_max]
            pos['params'] += [i, i + 1]
            grad_map.append((i, 'shape_params', n_params, 2, tuple(p_min.shape)))
            grad_map.append((i + 1, 'shape_params', n_params + 2, 2, tuple(p_max.shape)))
            n_params += 4
            i += 2
        else:
            raise ValueError('unknown shape type %r' % (t,))
        sw = args[i]
        src_sw.append(sw)
        pos['sw'].append(i)
        if not (shape_type[s] == 2 and has_thick[s]):
            grad_map.append((i, 'stroke_width', s, 1, tuple(sw.shape)))
        i += 1
    ng = num_shape_groups
    g_off = np.zeros(ng, np.int32)
    g_n = np.zeros(ng, np.int32)
    ids_list = []
    paint = {k: np.zeros(ng, np.int32) for k in ('fill_type', 'fill_offset', 'fill_num_stops', 'stroke_type',
                                                  'stroke_offset', 'stroke_num_stops')}
    paint['fill_type'][:] = -1
    paint['stroke_type'][:] = -1
    even_odd = np.zeros(ng, np.uint8)
    n_color = 0
    n_ids = 0
    for g in range(ng):
        ids = args[i]
        pos['ids'].append(i)
        i += 1
        ids_list.append(ids)
        g_off[g] = n_ids
        g_n[g] = int(ids.shape[0])
        n_ids += int(ids.shape[0])
        for kind in ('fill', 'stroke'):
            ct = args[i]
            i += 1
            if ct is None:
                continue
            paint[kind + '_offset'][g] = n_color
            if ct == diffvg.ColorType.constant:
                c = args[i]
                paint[kind + '_type'][g] = 0
                src_color.append(c)
                pos['color'].append(i)
                grad_map.append((i, 'color_params', n_color, 4, tuple(c.shape)))
                n_color += 4
                i += 1
            else:
                a, b, offsets, colors = args[i], args[i + 1], args[i + 2], args[i + 3]
                nstops = int(offsets.shape[0])
                if int(colors.shape[0]) != nstops:
                    raise ValueError('gradient offsets and stop colours differ in count')
                paint[kind + '_type'][g] = 1 if ct == diffvg.ColorType.linear_gradient else 2
                paint[kind + '_num_stops'][g] = nstops
                src_color += [a, b, offsets, colors]
                pos['color'] += [i, i + 1, i + 2, i + 3]
                grad_map.append((i, 'color_params', n_color, 2, tuple(a.shape)))
                grad_map.append((i + 1, 'color_params', n_color + 2, 2, tuple(b.shape)))
                grad_map.append((i + 2, 'color_params', n_color + 4, nstops, tuple(offsets.shape)))
                grad_map.append((i + 3, 'color_params', n_color + 4 + nstops, 4 * nstops, tuple(colors.shape)))
                n_color += 4 + 5 * nstops
                i += 4
        even_odd[g] = 1 if args[i] else 0
        xf = args[i + 1]
        src_xform.append(xf)
        pos['xform'].append(i + 1)
        grad_map.append((i + 1, 'shape_to_canvas', 9 * g, 9, tuple(xf.shape)))
        i += 2
    filter_type = args[i]
    filter_radius = args[i + 1]
    grad_map.append((i + 1, 'filter_radius', 0, 1, tuple(filter_radius.shape)))
    f.num_args = i + 2

    # one concatenation per kind and one host copy for all float data
    params, points, thick_vals, sw, color, xform, frad = _gather_kinds(
        [src_params, src_points, [t for _, t in src_thick], src_sw, src_color, src_xform, [filter_radius]])
    # the reference asserts finite points and thicknesses (render_pytorch.py:63, 68)
    assert np.isfinite(points).all() and np.isfinite(thick_vals).all(), 'non-finite path points or thickness'
    thickness = np.zeros(n_points, np.float32)
    k = 0
    for off, t in src_thick:
        n = int(t.numel())
        thickness[off:off + n] = thick_vals[k:k + n]
        k += n
    # stroke widths are scalars in practice; the first element of each is used
    sw_sizes = np.fromiter((int(t.numel()) for t in src_sw), np.int64, count=ns)
    sw_first = np.concatenate([[0], np.cumsum(sw_sizes)[:-1]]).astype(np.int64)
    sw_arr = np.ascontiguousarray(sw[sw_first] if ns else np.zeros(0), np.float32)
    sw_arr[(shape_type == 2) & (has_thick != 0)] = 0.0
    ncp = _cat_int(ncp_list)
    gids = _cat_int(ids_list)
    f.desc = dict(
        canvas_width=int(canvas_width), canvas_height=int(canvas_height), num_shapes=ns,
        shape_type=shape_type, stroke_width=sw_arr, shape_param_offset=param_off,
        shape_params=np.ascontiguousarray(params, np.float32), num_shape_params=int(n_params),
        path_ctrl_offset=ctrl_off, path_num_base_points=nbase, path_point_offset=pt_off, path_num_points=npts,
        path_has_thickness=has_thick, path_is_closed=closed, path_use_distance_approx=approx,
        num_control_points=np.ascontiguousarray(ncp, np.int32), num_ctrl_total=int(n_ctrl),
        points=np.ascontiguousarray(points, np.float32), thickness=thickness, num_points_total=int(n_points),
        num_groups=ng, group_shape_offset=g_off, group_num_shapes=g_n,
        group_shape_ids=np.ascontiguousarray(gids, np.int32), num_group_shape_ids=int(n_ids),
        fill_type=paint['fill_type'], fill_offset=paint['fill_offset'], fill_num_stops=paint['fill_num_stops'],
        stroke_type=paint['stroke_type'], stroke_offset=paint['stroke_offset'],
        stroke_num_stops=paint['stroke_num_stops'],
        color_params=np.ascontiguousarray(color, np.float32), num_color_params=int(n_color),
        use_even_odd_rule=even_odd, shape_to_canvas=np.ascontiguousarray(xform, np.float32),
        filter_type=_filter_int(filter_type), filter_radius=float(frad[0]))
    f.grad_map = grad_map
    f.pos = pos
    f.thick_ranges = [(off, int(t.numel())) for off, t in src_thick]
    f.sw_first = sw_first
    f.sw_zero = (shape_type == 2) & (has_thick != 0)
    f.int_bytes = ncp.tobytes() + b'|' + gids.tobytes()
    return f


def _picker(pos):
    """args -> tuple of args[i] for i in pos, in one call (scenes hold thousands)."""
    if not pos:
        return lambda args: ()
    if len(pos) == 1:
        i = pos[0]
        return lambda args: (args[i],)
    return operator.itemgetter(*pos)


_size = torch.Tensor.size
_unflatten = torch._C._nn.unflatten_dense_tensors
_deref = weakref.ref.__call__


def _weak(tensors):
    return tuple(map(weakref.ref, tensors))


def _same(tensors, refs):
    """`tensors` are the objects `refs` (weak references) point to, in order."""
    return len(tensors) == len(refs) and all(map(operator.is_, map(_deref, refs), tensors))
_flatten = torch._C._nn.flatten_dense_tensors
_version = operator.attrgetter('_version')
_requires_grad = operator.attrgetter('requires_grad')
_grad_fn = operator.attrgetter('grad_fn')
_hooks = operator.attrgetter('_backward_hooks')
_device = operator.attrgetter('device')
_dtype = operator.attrgetter('dtype')


_NATIVE = []


def _native():
    """The host-side delivery helper (_dvg_deliver.so, built with the library), or None
    when it is not built: delivery then cuts the views in Python, with the same result."""
    if not _NATIVE:
        import importlib.machinery
        import importlib.util
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_dvg_deliver.so')
        mod = None
        if os.path.exists(path):
            loader = importlib.machinery.ExtensionFileLoader('_dvg_deliver', path)
            spec = importlib.util.spec_from_file_location('_dvg_deliver', path, loader=loader)
            mod = importlib.util.module_from_spec(spec)
            loader.exec_module(mod)
        _NATIVE.append(mod)
    return _NATIVE[0]


_I32 = {}


def _as_i32(t):
    """`t` as int32.  The conversion of another integer type is kept while `t` lives
    unmodified, so a scene serialized again passes the same tensor objects (and the
    template's identity check holds)."""
    if t.dtype == torch.int32:
        return t
    key = id(t)
    e = _I32.get(key)
    if e is not None and e[0]() is t and e[1] == t._version:
        return e[2]
    out = t.to(torch.int32)
    _I32[key] = (weakref.ref(t, lambda _, k=key: _I32.pop(k, None)), t._version, out)
    return out


class _Resized(Exception):
    """A template's argument tensor was resized in place since its last match."""


class _SceneTemplate:
    """Topology of one serialize_scene argument list: everything but the float values
    (shape kinds, segment counts, point counts, group members, paint kinds, flags,
    canvas and filter type), the argument positions of the float values per kind, and a
    pool of device scenes of that topology.  A later argument list with the same
    topology gets a pooled scene rebuilt on the device from its values
    (dvg_scene_update) instead of a host build and upload."""

    MAX_IDLE = 64

    def __init__(self, args, flat):
        self.nargs = len(args)
        self.flat = flat
        self.float_pos = sorted({i for (i, *_rest) in flat.grad_map} | set(flat.pos['sw']))
        tensor_pos = set(self.float_pos) | set(flat.pos['ncp']) | set(flat.pos['ids'])
        self.scalar_pos = [i for i in range(self.nargs) if i != 6 and i not in tensor_pos]
        self.scalar_vals = tuple(args[i] for i in self.scalar_pos)
        self._get_scalars = _picker(self.scalar_pos)
        self._get_floats = _picker(self.float_pos)
        self.float_shapes = list(map(_size, self._get_floats(args)))
        # how each kind's tensors concatenate (their ranks are part of the topology)
        self.cat_mode = {}
        for kind in ('params', 'points', 'thick', 'sw', 'color', 'xform'):
            ts = [args[i] for i in flat.pos[kind]]
            ranks = {t.dim() for t in ts}
            if not ts:
                mode = None
            elif ranks == {0}:
                mode = 'stack'
            elif ranks == {1}:
                mode = 'cat'
            elif len(ranks) == 1 and len({tuple(t.shape[1:]) for t in ts}) == 1:
                mode = 'cat_flat'
            else:
                mode = 'flatten_each'
            self.cat_mode[kind] = mode
        self.int_pos = flat.pos['ncp'] + flat.pos['ids']
        self._get_ints = _picker(self.int_pos)
        self._get_kind = {kind: _picker(flat.pos[kind]) for kind in ('params', 'points', 'thick', 'sw', 'color',
                                                                     'xform')}
        self.int_bytes = flat.int_bytes
        # the integer topology tensors last seen with these contents, and their versions:
        # the same tensors, unmodified, need no comparison of contents
        # identity caches hold weak references: a template keeps no argument tensor alive
        ints = self._get_ints(args)
        self._int_seen = _weak(ints)
        self._int_versions = list(map(_version, ints))
        self.idle = []
        self._dev_tables = {}
        self._runs = None
        self._direct_last = None
        # the float argument tensors of the last match and the element count of each
        # kind's gather: the same tensor objects need no shape comparison, and a gather
        # whose count differs (a tensor resized in place since) is caught in device_params
        self._float_seen = None
        self.kind_numel = {kind: sum(int(t.numel()) for t in self._get_kind[kind](args))
                           for kind in self._get_kind}

    def matches(self, args):
        if len(args) != self.nargs:
            return False
        try:
            if self._get_scalars(args) != self.scalar_vals:
                return False
            floats = self._get_floats(args)
            seen = self._float_seen
            if seen is None or not _same(floats, seen):
                if list(map(_size, floats)) != self.float_shapes:
                    return False
                self._float_seen = _weak(floats)
        except (RuntimeError, TypeError, AttributeError):
            return False
        cur = self._get_ints(args)
        if _same(cur, self._int_seen) and list(map(_version, cur)) == self._int_versions:
            return True
        ncp = _cat_int([args[i] for i in self.flat.pos['ncp']])
        gids = _cat_int([args[i] for i in self.flat.pos['ids']])
        if ncp.tobytes() + b'|' + gids.tobytes() != self.int_bytes:
            return False
        self._int_seen = _weak(cur)
        self._int_versions = list(map(_version, cur))
        return True

    def _tables(self, dev):
        t = self._dev_tables.get(dev)
        if t is None:
            f = self.flat
            npts = f.desc['num_points_total']
            thick_idx = None
            if f.thick_ranges:
                thick_idx = torch.from_numpy(np.concatenate([np.arange(o, o + n) for o, n in f.thick_ranges])).to(dev)
            t = dict(sw_first=torch.from_numpy(f.sw_first).to(dev),
                     sw_zero=torch.from_numpy(f.sw_zero.astype(np.bool_)).to(dev),
                     thick_idx=thick_idx, npts=npts)
            self._dev_tables[dev] = t
        return t

    def device_params(self, args, dev, grad=False):
        """Float values of `args` as device arrays in the description's layout.  With
        `grad` the arrays are built by differentiable ops, so that gradients of the
        arrays flow back to the argument tensors (through torch's own cat/index
        backward, one node per kind)."""
        f, T = self.flat, self._tables(dev)
        with torch.set_grad_enabled(grad):
            def cat(kind):
                mode = self.cat_mode[kind]
                if mode is None:
                    return None
                srcs = self._get_kind[kind](args)
                for attempt in range(2):
                    try:
                        if mode == 'stack':
                            out = torch.stack(srcs)
                        elif mode == 'cat':
                            out = torch.cat(srcs)
                        elif mode == 'cat_flat':
                            out = torch.cat(srcs).reshape(-1)
                        else:
                            out = _flatten(srcs)
                        break
                    except RuntimeError:
                        if attempt:
                            raise
                        # tensors of one kind on several devices: move each first
                        srcs = [t.to(dev) for t in srcs]
                if out.numel() != self.kind_numel[kind]:
                    raise _Resized()
                return out.to(device=dev, dtype=torch.float32).contiguous()
            params = cat('params')
            points = cat('points')
            thickness = None
            if T['thick_idx'] is not None:
                thickness = torch.zeros(T['npts'], dtype=torch.float32, device=dev).index_put(
                    (T['thick_idx'],), cat('thick'))
            sw = cat('sw')
            if sw is not None:
                sw = torch.where(T['sw_zero'], torch.zeros((), device=dev), sw[T['sw_first']]).contiguous()
            return dict(shape_params=params, points=points, thickness=thickness, stroke_width=sw,
                        color_params=cat('color'), shape_to_canvas=cat('xform'))

    def _direct_runs(self):
        """The gradient slots of the float arguments (the map from argument positions to
        gradient slices) grouped into runs that one view + unbind of the gradient block
        covers: consecutive slots of one kind, shape and size.  False when some argument's
        slot is not the whole tensor (a stroke width tensor of several elements)."""
        if self._runs is None:
            runs = []
            # the filter radius is an input of the render node itself
            slots = [e for e in self.flat.grad_map if e[1] != 'filter_radius']
            for pos, kind, off, cnt, shape in sorted(slots, key=lambda e: (e[1], e[2])):
                if int(np.prod(shape, dtype=np.int64)) != cnt:
                    self._runs = False
                    return False
                r = runs[-1] if runs else None
                if r and r[0] == kind and r[2] == cnt and r[3] == shape and r[1] + r[4] * cnt == off:
                    r[4] += 1
                    r[5].append(pos)
                else:
                    runs.append([kind, off, cnt, shape, 1, [pos]])
            run_pos = [p for r in runs for p in r[5]]
            self._run_get = _picker(run_pos)
            # per tensor in run order: its kind and its offset within the kind's slots
            self._kinds = sorted({r[0] for r in runs})
            kind_idx = {k: i for i, k in enumerate(self._kinds)}
            self._leaf_kind = torch.tensor([kind_idx[r[0]] for r in runs for _ in range(r[4])], dtype=torch.int64)
            self._leaf_off = torch.tensor([r[1] + j * r[2] for r in runs for j in range(r[4])], dtype=torch.int64)
            # float arguments without a gradient slot (the stroke width of a path with
            # per-point thickness) receive zeros, as the engine gives them
            self._zero_pos = sorted(set(self.float_pos) - set(run_pos))
            self._zero_get = _picker(self._zero_pos)
            self._runs = runs
            # how deliver cuts the views: a long run with one view + unbind; consecutive
            # short runs of one kind (paths of differing point counts) with one split
            # along the first dimension when their shapes differ only there (one unbind
            # for scalars), else with one unflatten_dense_tensors call over their slots.
            # (kind, block offset, count of floats, first and last-plus-one index into
            # the run order, run or None, (trailing shape, first-dimension sizes, None for
            # scalars) or None)
            ops, shapes, i = [], [], 0
            for r in runs:
                kind, off, cnt, shape, m = r[0], r[1], r[2], r[3], r[4]
                last = ops[-1] if ops else None
                if m >= 8:
                    ops.append([kind, off, m * cnt, i, i + m, r, None])
                    shapes.append(None)
                elif last and last[5] is None and last[0] == kind and last[1] + last[2] == off:
                    last[2] += m * cnt
                    last[4] = i + m
                    shapes[-1] += [shape] * m
                else:
                    ops.append([kind, off, m * cnt, i, i + m, None, None])
                    shapes.append([shape] * m)
                i += m
            for op, shp in zip(ops, shapes):
                if shp and all(len(t) == 0 for t in shp):
                    op[6] = ((), None)  # scalars: one unbind
                elif shp and all(len(t) >= 1 and t[1:] == shp[0][1:] for t in shp):
                    op[6] = (tuple(shp[0][1:]), [t[0] for t in shp])
            self._ops = ops
        return self._runs

    def direct_leaves(self, args, dev):
        """(tensors in run order, which of them require a gradient, zero-gradient tensors,
        whether a tensor appears more than once) when gradients of `args` can be delivered
        directly, else None."""
        if self._direct_runs() is False:
            return None
        ts = self._run_get(args)
        rg = list(map(_requires_grad, ts))
        last = self._direct_last
        if last is not None and last[1] == rg and _same(ts, last[0]):
            # the tensors of the last accepted call, with the same gradient flags: they are
            # still leaves on the same device (a tensor that requires no gradient and
            # becomes a non-leaf changes its flag); only hooks may have been added since
            sel = list(itertools.compress(ts, rg)) + [z for z in self._zero_get(args) if z.requires_grad]
            if any(map(_hooks, sel)):
                return None
            return ts, rg, sel[len(sel) - last[2]:] if last[2] else [], last[3]
        sel = list(itertools.compress(ts, rg))
        if not sel:
            return None
        zs = [z for z in self._zero_get(args) if z.requires_grad]
        sel += zs
        # only the tensors that receive a gradient matter: they must be leaves without
        # hooks, float32 on the render device, and each slot's tensor used once (or the
        # slots of a shared tensor are summed, below)
        if any(map(_grad_fn, sel)) or any(map(_hooks, sel)):
            return None
        if dev.type == 'cuda' and dev.index is None:
            dev = torch.device('cuda', torch.cuda.current_device())
        if set(map(_device, sel)) != {dev} or set(map(_dtype, sel)) != {torch.float32}:
            return None
        dup = len(set(map(id, sel))) != len(sel)
        self._direct_last = (_weak(ts), rg, len(zs), dup)
        return ts, rg, zs, dup

    def deliver(self, block, layout, plan):
        """Writes the gradients of a direct-delivery plan into the tensors' .grad."""
        ts, rg, zeros, dup = plan
        native = _native()
        if native is not None:
            native.deliver(block, ts, rg, self._leaf_kind, self._leaf_off, [layout[k] for k in self._kinds])
            ops = ()
        else:
            ops = self._ops
        acc_t, acc_g = [], []
        for kind, off, n, i0, i1, run, split in ops:
            sel = rg[i0:i1]
            if not any(sel):
                continue
            o = layout[kind] + off
            part = ts[i0:i1]
            if run is not None:
                views = block[o:o + n].view((run[4],) + run[3]).unbind(0)
            elif split is not None:
                views = (block[o:o + n].unbind(0) if split[1] is None else
                         block[o:o + n].view((-1,) + split[0]).split(split[1]))
            else:
                views = _unflatten(block[o:o + n], part)
            for t, g, r in zip(part, views, sel):
                if r:
                    if t.grad is None:
                        t.grad = g
                    else:
                        acc_t.append(t.grad)
                        acc_g.append(g)
        if dup:
            # a tensor used several times (a shared default transform) sums its slots:
            # one add at a time, since a fused add must not see one tensor twice
            for a, g in zip(acc_t, acc_g):
                a.add_(g)
        elif acc_t:
            torch._foreach_add_(acc_t, acc_g)
        for z in zeros:
            if z.grad is None:
                z.grad = torch.zeros_like(z)

    def acquire(self, dev_index):
        if self.idle:
            return self.idle.pop()
        return _lib.DeviceScene(self.flat.desc, dev_index, _lib.MODE_TILED)

    def release(self, scene):
        if len(self.idle) < self.MAX_IDLE:
            self.idle.append(scene)


class _Lease:
    """A pooled device scene held by one forward pass (and its backward); it goes back to
    its template's pool when the autograd context lets go of it."""

    def __init__(self, template, scene):
        self._template, self._scene = template, scene

    def __getattr__(self, name):
        return getattr(self._scene, name)

    def release(self):
        """Returns the scene to its pool now (once); the lease is unusable afterwards."""
        tpl, self._template = self._template, None
        if tpl is not None:
            tpl.release(self._scene)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


_TEMPLATES = {}  # number of arguments -> templates, most recently used first
_MAX_TEMPLATES = 1024


def _templates():
    return [t for ts in _TEMPLATES.values() for t in ts]


def _find_template(args, hint=None, skip=()):
    """The pooled topology `args` match (moved to the front of its bucket), or None.
    `hint`: a template to try first (the one the caller's previous call with these
    arguments used); `skip`: templates not to try."""
    bucket = _TEMPLATES.get(len(args))
    if not bucket:
        return None
    order = enumerate(bucket)
    if hint is not None and hint.nargs == len(args):
        k0 = next((k for k, t in enumerate(bucket) if t is hint), None)
        if k0 is not None:
            order = itertools.chain([(k0, hint)], ((k, t) for k, t in enumerate(bucket) if k != k0))
    for k, tpl in order:
        if tpl not in skip and tpl.matches(args):
            if k:
                bucket.insert(0, bucket.pop(k))
            return tpl
    return None


def _pooled_scene(tpl, args, dev, vals, plan, check):
    """(flat, lease, arrays) for a pooled scene of `tpl` rebuilt on the device from `vals`."""
    scene = tpl.acquire(dev.index if dev.index is not None else -1)
    scene.update(float(args[-1]), _stream(), **vals)
    if check:
        # what Scene construction rejects is reported now, as the reference reports it
        # (one wait for the rebuild kernels, before the render is enqueued)
        try:
            check_rebuild(scene)
        except Exception:
            tpl.release(scene)
            raise
    f = _Flat()
    f.__dict__.update(tpl.flat.__dict__)
    f.eval_positions = args[6]
    f.direct = None if plan is None else (tpl, plan)
    return f, _Lease(tpl, scene), vals


def _host_built_scene(args, dev, grad):
    """(flat, lease, arrays) for a topology seen for the first time: built on the host,
    starting a new pool."""
    flat = flatten_args(args)
    tpl = _SceneTemplate(args, flat)
    if sum(len(ts) for ts in _TEMPLATES.values()) >= _MAX_TEMPLATES:
        clear_scene_cache()
    _TEMPLATES.setdefault(len(args), []).insert(0, tpl)
    scene = _lib.DeviceScene(flat.desc, dev.index if dev.index is not None else -1, _lib.MODE_TILED)
    vals = tpl.device_params(args, dev, True) if grad else None
    return flat, _Lease(tpl, scene), vals


def _device_scene_for(args, dev, check=True, grad=False, direct=False, hint=None):
    """(flat, scene, arrays) for a TILED render of `args`: a pooled scene of the same
    topology rebuilt on the device when one exists, else a host build that starts a new
    pool.  `arrays` are the float arrays in the description's layout (differentiable
    with `grad`; None for a host build without `grad`).  With check=False the caller
    reports the rebuild's findings (scene.status()).  With `grad` and `direct`, a pooled
    topology whose arguments allow it sets flat.direct to its direct-delivery plan and
    returns arrays that are not differentiable.  `hint`: a template to try first."""
    skip = []
    while True:
        tpl = _find_template(args, hint, skip)
        if tpl is None:
            return _host_built_scene(args, dev, grad)
        plan = tpl.direct_leaves(args, dev) if grad and direct else None
        try:
            vals = tpl.device_params(args, dev, grad and plan is None)
        except _Resized:
            # compare the shapes after all; a template they no longer match is passed
            tpl._float_seen = None
            if not tpl.matches(args):
                skip.append(tpl)
                continue
            vals = tpl.device_params(args, dev, grad and plan is None)
        return _pooled_scene(tpl, args, dev, vals, plan, check)


_GATHER_KINDS = (('params', 'shape_params'), ('points', 'points'), ('color', 'color_params'),
                 ('xform', 'shape_to_canvas'))
_BATCH_TABLES = {}


def _batch_tables(tpls, dev):
    """Index tables of a batch of templates (cached per sequence of templates): element
    offsets of each scene in every batch-wide gather, and the stroke-width and thickness
    index maps of the whole batch."""
    key = (tuple(map(id, tpls)), dev)
    t = _BATCH_TABLES.get(key)
    if t is not None and all(a is b for a, b in zip(t['tpls'], tpls)):
        return t
    offs = {}
    for kind in ('params', 'points', 'thick', 'sw', 'color', 'xform'):
        offs[kind] = np.concatenate([[0], np.cumsum([tp.kind_numel[kind] for tp in tpls])]).astype(np.int64)
    # stroke widths: per scene, shape s takes its stroke-width argument's first element
    sw_first, sw_zero, ns_off = [], [], [0]
    pt_off = [0]
    for i, tp in enumerate(tpls):
        sw_first.append(tp.flat.sw_first + offs['sw'][i])
        sw_zero.append(tp.flat.sw_zero)
        ns_off.append(ns_off[-1] + len(tp.flat.sw_first))
        pt_off.append(pt_off[-1] + tp.flat.desc['num_points_total'])
    thick_idx = [np.arange(o, o + n) + pt_off[i] for i, tp in enumerate(tpls) for o, n in tp.flat.thick_ranges]
    t = dict(tpls=list(tpls), offs=offs, ns_off=ns_off, pt_off=pt_off,
             sw_first=torch.from_numpy(np.concatenate(sw_first) if sw_first else np.zeros(0, np.int64)).to(dev),
             sw_zero=torch.from_numpy(np.concatenate(sw_zero).astype(np.bool_) if sw_zero else np.zeros(0, bool)).to(dev),
             thick_idx=torch.from_numpy(np.concatenate(thick_idx)).to(dev) if thick_idx else None)
    if len(_BATCH_TABLES) > 16:
        _BATCH_TABLES.clear()
    _BATCH_TABLES[key] = t
    return t


def _batch_params(tpls, arg_lists, dev, grad):
    """Device arrays of many pooled scenes gathered with one op per kind over all of
    them, then cut per scene (views): the per-scene form of device_params, for batches.
    Raises _Resized when a tensor was resized in place since its template matched."""
    T = _batch_tables(tpls, dev)
    offs = T['offs']
    with torch.set_grad_enabled(grad):
        def gather(kind):
            srcs, modes = [], set()
            for tp, a in zip(tpls, arg_lists):
                mode = tp.cat_mode[kind]
                if mode is not None:
                    srcs.extend(tp._get_kind[kind](a))
                    modes.add(mode)
            if not srcs:
                return None
            try:
                if modes == {'cat'}:
                    out = torch.cat(srcs)
                elif modes == {'stack'}:
                    out = torch.stack(srcs)
                elif modes == {'cat_flat'}:
                    out = torch.cat(srcs).reshape(-1)
                else:
                    out = _flatten(srcs)
            except RuntimeError:
                out = _flatten([t.to(dev) for t in srcs])
            if out.numel() != int(offs[kind][-1]):
                raise _Resized()
            return out.to(device=dev, dtype=torch.float32).contiguous()
        big = {name: gather(kind) for kind, name in _GATHER_KINDS}
        sw = gather('sw')
        if sw is not None:
            sw = torch.where(T['sw_zero'], torch.zeros((), device=dev), sw[T['sw_first']])
        thick = None
        if T['thick_idx'] is not None:
            thick = torch.zeros(T['pt_off'][-1], dtype=torch.float32, device=dev).index_put(
                (T['thick_idx'],), gather('thick'))
    out = []
    for i, tp in enumerate(tpls):
        vals = {}
        for kind, name in _GATHER_KINDS:
            o0, o1 = int(offs[kind][i]), int(offs[kind][i + 1])
            vals[name] = big[name][o0:o1] if tp.cat_mode[kind] is not None else None
        vals['stroke_width'] = sw[T['ns_off'][i]:T['ns_off'][i + 1]] if tp.cat_mode['sw'] is not None else None
        vals['thickness'] = (thick[T['pt_off'][i]:T['pt_off'][i + 1]]
                             if tp.flat.thick_ranges else None)
        out.append(vals)
    return out


def check_rebuild(scene):
    """Raises what Scene construction raises for the values of the last rebuild (an
    AssertionError for non-finite points, as render_pytorch.py:63 asserts)."""
    try:
        scene.status()
    except RuntimeError as e:
        if 'non-finite path points' in str(e):
            raise AssertionError('non-finite path points or thickness') from None
        raise


def clear_scene_cache():
    """Drops the pooled device scenes (they are rebuilt on demand), and every table that
    refers to their templates."""
    _TEMPLATES.clear()
    _BATCH_TABLES.clear()
    from . import batch_render
    batch_render._HINTS.clear()


# Render passes whose device-side findings (fragment overflow, a non-finite grad_img, a
# prefiltered circle stroke) have not been read yet: the passes return while their
# kernels run, so the host goes on (the autograd engine hands out the gradients
# meanwhile); the findings are read at the next entry into the render path or at
# synchronize().
_PENDING = []


def synchronize():
    """Waits for the render passes still in flight and raises what they found (the
    reference's backward asserts a finite grad_img and aborts on fragment overflow; its
    prefiltered forward asserts on a stroked circle)."""
    while _PENDING:
        work, finite = _PENDING.pop(0)
        try:
            work.status()
        except RuntimeError as e:
            raise RuntimeError('%s (found by an earlier render pass)' % e) from None
        if finite is not None:
            assert bool(finite), 'non-finite grad_img in an earlier backward pass'


def _check_gpu(read_findings=True):
    if read_findings:
        synchronize()
    if not get_use_gpu() or get_device().type != 'cuda':
        raise RuntimeError('diffvg_amd renders on the GPU only; call diffvg_amd.set_use_gpu(True) '
                           'on a machine with a ROCm GPU')


def _stream():
    return ctypes_stream(torch.cuda.current_stream())


def ctypes_stream(s):
    return s.cuda_stream


class RenderFunction(torch.autograd.Function):
    """The PyTorch interface of the render path."""

    @staticmethod
    def serialize_scene(canvas_width,
                        canvas_height,
                        shapes,
                        shape_groups,
                        filter=PixelFilter(type=diffvg.FilterType.box, radius=torch.tensor(0.5)),
                        output_type=OutputType.color,
                        use_prefiltering=False,
                        eval_positions=torch.tensor([])):
        """Scene -> flat argument list in the reference's order (render_pytorch.py:22-172)."""
        num_shapes = len(shapes)
        num_shape_groups = len(shape_groups)
        args = [canvas_width, canvas_height, num_shapes, num_shape_groups, output_type, use_prefiltering,
                eval_positions.to(get_device())]
        # locals: the loops run once per shape and group of scenes of thousands
        extend, append = args.extend, args.append
        i32 = torch.int32
        t_path, t_circle, t_ellipse = diffvg.ShapeType.path, diffvg.ShapeType.circle, diffvg.ShapeType.ellipse
        t_rect = diffvg.ShapeType.rect
        c_const = diffvg.ColorType.constant
        Tensor = torch.Tensor
        for shape in shapes:
            kind = type(shape)
            if kind is Path or (kind is not Polygon and isinstance(shape, Path)):
                points = shape.points
                assert points.size(1) == 2
                ncp = shape.num_control_points
                sw = shape.stroke_width
                if sw.dim() > 0 and sw.size(0) > 1:
                    # per-point thickness; the stroke width slot then holds zero
                    extend((t_path, ncp if ncp.dtype == i32 else _as_i32(ncp), points, sw, shape.is_closed,
                            shape.use_distance_approx, torch.tensor(0.0)))
                else:
                    extend((t_path, ncp if ncp.dtype == i32 else _as_i32(ncp), points, None, shape.is_closed,
                            shape.use_distance_approx, sw))
            elif isinstance(shape, Circle):
                extend((t_circle, shape.radius, shape.center, shape.stroke_width))
            elif isinstance(shape, Ellipse):
                extend((t_ellipse, shape.radius, shape.center, shape.stroke_width))
            elif isinstance(shape, Polygon):
                assert shape.points.shape[1] == 2
                n = shape.points.shape[0] if shape.is_closed else shape.points.shape[0] - 1
                extend((t_path, torch.zeros(n, dtype=i32), shape.points, None, shape.is_closed, False,
                        shape.stroke_width))
            elif isinstance(shape, Rect):
                extend((t_rect, shape.p_min, shape.p_max, shape.stroke_width))
            else:
                raise TypeError('unsupported shape %r' % (type(shape),))
        for shape_group in shape_groups:
            ids = shape_group.shape_ids
            append(ids if ids.dtype == i32 else _as_i32(ids))
            fill = shape_group.fill_color
            for color in (fill, shape_group.stroke_color):
                if color is None:
                    append(None)
                elif isinstance(color, Tensor):
                    extend((c_const, color))
                elif isinstance(color, LinearGradient):
                    extend((diffvg.ColorType.linear_gradient, color.begin, color.end, color.offsets,
                            color.stop_colors))
                elif isinstance(color, RadialGradient):
                    extend((diffvg.ColorType.radial_gradient, color.center, color.radius, color.offsets,
                            color.stop_colors))
                else:
                    raise TypeError('unsupported paint %r' % (type(color),))
            if fill is not None:
                for shape_id in ids.tolist():
                    if isinstance(shapes[shape_id], Path) and not shapes[shape_id].is_closed:
                        warnings.warn('Detected non-closed paths with fill color. This might causes '
                                      'unexpected results.', Warning)
            extend((shape_group.use_even_odd_rule, shape_group.shape_to_canvas.contiguous()))
        args.append(filter.type)
        args.append(filter.radius)
        return args

    @staticmethod
    def _prepare(width, height, background_image, args, output_type=None, grad=False):
        """-> (flat, scene, arrays); `arrays` (the float arrays in the description's
        layout, differentiable) only with `grad`."""
        _check_gpu(read_findings=False)
        dev = get_device()
        start = time.time()
        if _render_mode == _lib.MODE_TILED:
            flat, scene, vals = _device_scene_for(args, dev, grad=grad, direct=_direct_leaf_grads)
        else:
            # the comparator: a host-built scene per call, as the reference builds it
            flat = flatten_args(args)
            scene = _lib.DeviceScene(flat.desc, dev.index if dev.index is not None else -1, _render_mode)
            vals = _SceneTemplate(args, flat).device_params(args, dev, True) if grad else None
        # Findings of earlier backward passes, read once this call's scene work is on the
        # stream: the host prepares this render while the previous backward's kernels
        # run, and the rebuild check above has already waited for them.
        try:
            synchronize()
        except BaseException:
            if isinstance(scene, _Lease):
                scene.release()
            raise
        if print_timing:
            print('Scene construction, time: %.5f s' % (time.time() - start))
        return flat, scene, vals

    @classmethod
    def apply(cls, width, height, num_samples_x, num_samples_y, seed, background_image, *args):
        """Renders the serialize_scene argument list `args` (render_pytorch.py:174-428).

        The autograd node is taken over the scene's float arrays -- one tensor per kind
        (shape parameters, points, thicknesses, stroke widths, paint parameters,
        transforms) gathered from the argument tensors by differentiable concatenation --
        and the filter radius, not over every argument tensor: each argument tensor that
        requires a gradient receives the same .grad as a node over all of them would give
        it, while the per-argument cost of the autograd machinery stays out of the call."""
        flat, scene, vals = RenderFunction._prepare(width, height, background_image, args,
                                                    grad=torch.is_grad_enabled())
        radius = args[flat.num_args - 1]
        if vals is None:
            vals = {k: None for k in _KINDS}
        meta = (flat, scene, width, height, num_samples_x, num_samples_y, seed)
        # with direct delivery the node's only edge to the scene tensors is a zero-size
        # leaf that makes the output require a gradient
        anchor = torch.empty(0, requires_grad=True) if getattr(flat, 'direct', None) else None
        return _RenderNode.apply(meta, background_image, radius if isinstance(radius, torch.Tensor) else None,
                                 anchor, *[vals[k] for k in _KINDS])

    @staticmethod
    def render_grad(grad_img, width, height, num_samples_x, num_samples_y, seed, background_image, *args):
        """Screen-space gradient of the image w.r.t. a translation of the scene
        (render_pytorch.py:430-666)."""
        if not grad_img.is_contiguous():
            grad_img = grad_img.contiguous()
        assert torch.isfinite(grad_img).all()
        flat, scene, _ = RenderFunction._prepare(width, height, background_image, args)
        dev = get_device()
        if flat.output_type == OutputType.color:
            assert grad_img.shape[2] == 4
        else:
            assert grad_img.shape[2] == 1
        if background_image is not None:
            background_image = background_image.to(dev)
            if background_image.shape[2] == 3:
                background_image = torch.cat((background_image, torch.ones(
                    background_image.shape[0], background_image.shape[1], 1, device=dev)), dim=2)
            background_image = background_image.contiguous().to(torch.float32)
            assert background_image.shape[0] == grad_img.shape[0]
            assert background_image.shape[1] == grad_img.shape[1]
        grad_img = grad_img.to(dev).to(torch.float32).contiguous()
        translation_grad_image = torch.zeros(height, width, 2, device=dev)
        ev = flat.eval_positions
        ev = ev.to(dev).to(torch.float32).contiguous() if ev.shape[0] > 0 else None
        start = time.time()
        scene.render(width, height, num_samples_x, num_samples_y, seed, background=background_image,
                     d_render_image=grad_img if flat.output_type == OutputType.color else None,
                     d_render_sdf=grad_img if flat.output_type == OutputType.sdf else None,
                     d_translation=translation_grad_image, use_prefiltering=flat.use_prefiltering,
                     eval_positions=ev, num_eval_positions=0 if ev is None else ev.shape[0], stream=_stream())
        scene.status()
        if print_timing:
            torch.cuda.synchronize()
            print('Gradient pass, time: %.5f s' % (time.time() - start))
        assert torch.isfinite(translation_grad_image).all()
        return translation_grad_image


_KINDS = ('shape_params', 'points', 'thickness', 'stroke_width', 'color_params', 'shape_to_canvas')


class _RenderNode(torch.autograd.Function):
    """The render as one autograd node: inputs are the background image, the filter
    radius, the anchor of a direct-delivery render (else None) and the scene's float
    arrays by kind (RenderFunction.apply builds them); the backward returns slices of
    the scene's gradient block for them, or, for a direct-delivery render, writes the
    argument tensors' gradients itself."""

    @staticmethod
    def forward(ctx, meta, background_image, filter_radius, anchor, *kinds):
        flat, scene, width, height, num_samples_x, num_samples_y, seed = meta
        dev = get_device()
        eval_positions = flat.eval_positions
        if flat.output_type == OutputType.color:
            assert eval_positions.shape[0] == 0
            rendered_image = torch.zeros(height, width, 4, device=dev)
        else:
            assert flat.output_type == OutputType.sdf
            if eval_positions.shape[0] == 0:
                rendered_image = torch.zeros(height, width, 1, device=dev)
            else:
                rendered_image = torch.zeros(eval_positions.shape[0], 1, device=dev)
        if background_image is not None:
            background_image = background_image.to(dev)
            if background_image.shape[2] == 3:
                raise NotImplementedError('Background image must have 4 channels, not 3. Add a fourth channel with '
                                          'all ones via torch.ones().')
            background_image = background_image.contiguous().to(torch.float32)
            assert background_image.shape[0] == rendered_image.shape[0]
            assert background_image.shape[1] == rendered_image.shape[1]
            assert background_image.shape[2] == 4
        ev = eval_positions.to(dev).to(torch.float32).contiguous() if eval_positions.shape[0] > 0 else None
        start = time.time()
        scene.render(width, height, num_samples_x, num_samples_y, seed,
                     background=background_image,
                     render_image=rendered_image if flat.output_type == OutputType.color else None,
                     render_sdf=rendered_image if flat.output_type == OutputType.sdf else None,
                     use_prefiltering=flat.use_prefiltering, eval_positions=ev,
                     num_eval_positions=0 if ev is None else ev.shape[0], stream=_stream())
        if flat.use_prefiltering:
            # a prefiltered stroke of a circle group, on which the reference asserts in the
            # forward: read with the other findings, at the next render call or
            # synchronize(), so the host does not wait for the forward's kernels here
            _PENDING.append((scene, None))
        if print_timing:
            torch.cuda.synchronize()
            print('Forward pass, time: %.5f s' % (time.time() - start))
        ctx.scene = scene
        ctx.flat = flat
        ctx.background_image = background_image
        ctx.width = width
        ctx.height = height
        ctx.num_samples_x = num_samples_x
        ctx.num_samples_y = num_samples_y
        ctx.seed = seed
        ctx.eval_positions = ev
        ctx.radius_shape = None if filter_radius is None else filter_radius.shape
        ctx.kind_shapes = [None if k is None else k.shape for k in kinds]
        return rendered_image

    @staticmethod
    def backward(ctx, grad_img):
        # earlier findings are read at the next render call or synchronize(), not here:
        # this call would wait for its own forward's kernels
        if not grad_img.is_contiguous():
            grad_img = grad_img.contiguous()
        # the reference asserts a finite grad_img before rendering; the check is taken on
        # the device and read after the backward is enqueued, so the host prepares the
        # gradient views while the device works
        finite = torch.isfinite(grad_img).all()
        scene = ctx.scene
        flat = ctx.flat
        background_image = ctx.background_image
        d_background_image = torch.zeros_like(background_image) if background_image is not None else None
        grad_img = grad_img.to(torch.float32).contiguous()
        start = time.time()
        scene.zero_grads(_stream())
        scene.render(ctx.width, ctx.height, ctx.num_samples_x, ctx.num_samples_y, ctx.seed,
                     background=background_image, d_background=d_background_image,
                     d_render_image=grad_img if flat.output_type == OutputType.color else None,
                     d_render_sdf=grad_img if flat.output_type == OutputType.sdf else None,
                     use_prefiltering=flat.use_prefiltering, eval_positions=ctx.eval_positions,
                     num_eval_positions=0 if ctx.eval_positions is None else ctx.eval_positions.shape[0],
                     stream=_stream())
        block, layout = scene.grads_device(grad_img.device, _stream())
        need = ctx.needs_input_grad
        d_radius = None
        if ctx.radius_shape is not None and need[2]:
            o = layout['filter_radius']
            d_radius = block[o:o + 1].view(ctx.radius_shape)
        direct = getattr(flat, 'direct', None)
        if direct is not None:
            direct[0].deliver(block, layout, direct[1])
        d_kinds = []
        for j, (name, shape) in enumerate(zip(_KINDS, ctx.kind_shapes)):
            if shape is None or not need[4 + j]:
                d_kinds.append(None)
                continue
            o = layout[name]
            n = 1
            for d in shape:
                n *= d
            d_kinds.append(block[o:o + n].view(shape))
        if print_timing:
            scene.status()
            print('Backward pass, time: %.5f s' % (time.time() - start))
        _PENDING.append((scene, finite))
        return (None, d_background_image, d_radius, None) + tuple(d_kinds)
