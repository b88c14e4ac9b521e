"""SVG documents as trainable scenes: ``SvgOptimizationSettings`` and ``OptimizableSvg``.

Interface kept from pydiffvg/optimize_svg.py (so that its callers, e.g.
apps/svg_brush.py:96-102 and apps/simple_transform_svg.py:90-104, switch over): the
settings schema and the settings methods, ``OptimizableSvg(filename, settings,
optimize_background, verbose, device)`` with ``build_scene``, ``render``, ``zero_grad``,
``step`` and ``write_xml``, and ``OptimizableSvg.TransformTools.decompose/recompose``
(apps/svg_parse_test.py:12-22).

Design.  A document becomes a tree of ``SvgNode`` records (one class; ``kind`` is the
SVG tag).  Every trainable tensor -- solid colours, opacities, path points, gradient
stops and end points, transform parameters -- is registered once in a document-wide
``ParameterBank`` together with its learning rate, optimizer family and the projection
that keeps it valid after a step (colours clamped to [1e-4, 1]; stop offsets clamped,
sorted and pinned to 0 and 1).  The bank owns one torch optimizer per family, with a
parameter group per tensor, so a step is one call per family plus the projections.
``build_scene`` walks the tree once per step, composing transforms and cascading the
appearance, and emits the (width, height, shapes, shape_groups) tuple that
``RenderFunction.serialize_scene`` takes.  Transforms train through a decomposition
M = R(theta) diag(sx, sy) shearX(k) translate(t) whose free parts depend on the
settings' transform mode.

Behaviours of the reference that change results are kept: presentation attributes beat
``style`` declarations, which beat class rules; style keys are not stripped (so a
declaration after "; " is not recognised); the root element's own appearance is not
inherited; rect, circle, ellipse and polygon geometry is not trained; stroke-width is
used as given; skewX/skewY take the shear coefficient itself, not an angle; lengths use
1 px = 0.25 mm.
"""
import copy
import json
import math
import re
import warnings
import xml.etree.ElementTree as ET
from collections import namedtuple
from xml.dom import minidom

import numpy as np
import torch

from .color import LinearGradient
from .render_pytorch import RenderFunction
from .shape import Circle, Ellipse, Path, Polygon, Rect, ShapeGroup, from_svg_path

__all__ = ['SvgOptimizationSettings', 'OptimizableSvg']


# ------------------------------------------------------------------------------ settings
class SvgOptimizationSettings:
    """Optimisation settings per SVG id, falling back to the "default" entry
    (optimize_svg.py:15-103).  Plain nested dicts, so they round-trip through JSON."""

    default_params = {
        "optimize_color": True,
        "color_lr": 2e-3,
        "optimize_alpha": False,
        "alpha_lr": 2e-3,
        "optimizer": "Adam",
        "transforms": {
            "optimize_transforms": True,
            "transform_mode": "rigid",
            "translation_mult": 1e-3,
            "transform_lr": 2e-3,
        },
        "circles": {
            "optimize_center": True,
            "optimize_radius": True,
            "shape_lr": 2e-1,
        },
        "paths": {
            "optimize_points": True,
            "shape_lr": 2e-1,
        },
        "gradients": {
            "optimize_stops": True,
            "stop_lr": 2e-3,
            "optimize_color": True,
            "color_lr": 2e-3,
            "optimize_alpha": False,
            "alpha_lr": 2e-3,
            "optimize_location": True,
            "location_lr": 2e-1,
        },
    }

    optims = {"Adam": torch.optim.Adam, "SGD": torch.optim.SGD, "ASGD": torch.optim.ASGD}

    def __init__(self, f=None):
        self.store = json.load(f) if f is not None else {"default": copy.deepcopy(self.default_params)}
        self._root_name = None

    def default_name(self, dname):
        """Marks `dname` (the document root's id) as the entry `save` writes as "default"."""
        self._root_name = dname
        self.store.setdefault(dname, self.store["default"])

    def retrieve(self, node_id):
        """-> (the entry for node_id, or the defaults; whether node_id has its own)."""
        own = node_id in self.store
        return self.store[node_id if own else "default"], own

    def reset_to_defaults(self, node_id):
        self.store.pop(node_id, None)
        return self.store["default"]

    def undefault(self, node_id):
        """Gives node_id an editable copy of the defaults (if it has no entry) and returns it."""
        if node_id not in self.store:
            self.store[node_id] = copy.deepcopy(self.store["default"])
        return self.store[node_id]

    def override_optimizer(self, optimizer):
        if optimizer is not None:
            self.global_override(["optimizer"], optimizer)

    def global_override(self, path, value):
        """Sets the entry at key path `path` to `value` in every stored entry."""
        for entry in self.store.values():
            for key in path[:-1]:
                entry = entry[key]
            entry[path[-1]] = value

    def save(self, file):
        if self._root_name is not None:
            self.store["default"] = self.store[self._root_name]
        json.dump(self.store, file, indent="\t")


# ------------------------------------------------------------------- affine 2-D algebra
TransformDecomposition = namedtuple("TransformDecomposition", "theta scale shear translate")


def _affine(a, b, c, d, e, f):
    """The SVG matrix(a b c d e f) as a 3x3 array."""
    return np.array([[a, c, e], [b, d, f], [0.0, 0.0, 1.0]], dtype=np.float64)


def _rotation(v):
    r = math.radians(v[0])
    m = _affine(math.cos(r), math.sin(r), -math.sin(r), math.cos(r), 0.0, 0.0)
    if len(v) == 3:  # about (cx, cy): move the centre to the origin, rotate, move back
        cx, cy = v[1], v[2]
        m = _affine(1, 0, 0, 1, cx, cy) @ m @ _affine(1, 0, 0, 1, -cx, -cy)
    return m


# name -> (accepted argument counts, matrix of the arguments)
_TRANSFORM_FUNCTIONS = {
    "matrix": ((6,), lambda v: _affine(*v)),
    "translate": ((1, 2), lambda v: _affine(1, 0, 0, 1, v[0], v[1] if len(v) == 2 else 0.0)),
    "scale": ((1, 2), lambda v: _affine(v[0], 0, 0, v[-1], 0, 0)),
    "rotate": ((1, 3), _rotation),
    "skewX": ((1,), lambda v: _affine(1, 0, v[0], 1, 0, 0)),
    "skewY": ((1,), lambda v: _affine(1, v[0], 0, 1, 0, 0)),
}
_TRANSFORM_ITEM = re.compile(r"\s*,?\s*([A-Za-z]+)\s*\(([^)]*)\)")
_NUMBER = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def parse_transform_list(text):
    """An SVG transform list -> 3x3 float64 matrix (the functions multiplied left to right)."""
    m, pos = np.eye(3), 0
    text = text.rstrip()
    while pos < len(text):
        item = _TRANSFORM_ITEM.match(text, pos)
        if item is None:
            raise ValueError("Cannot parse transform '{}'".format(text))
        name, args = item.group(1), [float(x) for x in re.findall(_NUMBER, item.group(2))]
        if name not in _TRANSFORM_FUNCTIONS:
            raise ValueError("Unknown transform type '{}'".format(name))
        counts, fn = _TRANSFORM_FUNCTIONS[name]
        if len(args) not in counts:
            raise ValueError("{} takes {} arguments, got {}".format(name, counts, len(args)))
        m = m @ fn(args)
        pos = item.end()
    return m


def _fmt(x):
    """A number as the shortest text that reads back to the same float32."""
    return "%.9g" % float(x)


class TransformTools:
    """The decomposition the transform parameters train in:
    M = [R(theta) diag(sx, sy) [[1, k], [0, 1]] | 0; 0 0 1] @ translate(tx, ty), sx > 0."""

    TransformDecomposition = TransformDecomposition

    @staticmethod
    def decompose(M):
        """3x3 affine matrix -> TransformDecomposition(theta, [sx, sy], k, [tx, ty]).

        The first column of the linear part A is R(theta) (sx, 0): sx is its length and
        theta its angle.  R(-theta) A is then upper triangular [[sx, sx k], [0, sy]], so
        sy = det(A) / sx and k = (R(-theta) A)[0, 1] / sx.  The translation is applied
        first, so t = A^-1 M[0:2, 2]."""
        M = np.asarray(M, dtype=np.float64)
        A, off = M[0:2, 0:2], M[0:2, 2]
        sx = math.hypot(A[0, 0], A[1, 0])
        theta = math.atan2(A[1, 0], A[0, 0])
        c, s = math.cos(theta), math.sin(theta)
        sy = float(np.linalg.det(A)) / sx
        k = (c * A[0, 1] + s * A[1, 1]) / sx
        return TransformDecomposition(theta, np.array([sx, sy]), k, np.linalg.solve(A, off))

    @staticmethod
    def _linear(theta, scale, shear, xp):
        c, s = xp.cos(theta), xp.sin(theta)
        sx, sy = (scale[0], scale[1]) if getattr(scale, "ndim", 0) and len(scale) == 2 else (scale, scale)
        # R(theta) @ diag(sx, sy) @ [[1, k], [0, 1]]
        return ((c * sx, c * sx * shear - s * sy), (s * sx, s * sx * shear + c * sy))

    @staticmethod
    def recompose(Theta, ScaleXY, ShearX, TXY):
        """Differentiable inverse of `decompose` on torch tensors (ScaleXY may be a
        scalar for a uniform scale)."""
        theta, shear = Theta.reshape(()), ShearX.reshape(())
        scale = ScaleXY.reshape(-1)
        scale = scale if scale.numel() == 2 else scale.reshape(())
        (a, c), (b, d) = TransformTools._linear(theta, scale, shear, torch)
        t = TXY.reshape(2)
        one, zero = torch.ones_like(theta), torch.zeros_like(theta)
        rows = [torch.stack([a, c, a * t[0] + c * t[1]]), torch.stack([b, d, b * t[0] + d * t[1]]),
                torch.stack([zero, zero, one])]
        return torch.stack(rows)

    @staticmethod
    def recompose_numpy(Theta, ScaleXY, ShearX, TXY):
        scale = np.asarray(ScaleXY, dtype=np.float64).reshape(-1)
        (a, c), (b, d) = TransformTools._linear(np.float64(Theta), scale if scale.size == 2 else scale[0],
                                                np.float64(ShearX), np)
        tx, ty = np.asarray(TXY, dtype=np.float64).reshape(2)
        return _affine(a, b, c, d, a * tx + c * ty, b * tx + d * ty)


def transform_features(d, eps=1e-3):
    """Which parts of a decomposition differ from the identity."""
    sx, sy = abs(float(d.scale[0])), abs(float(d.scale[1]))
    return dict(rotation=abs(d.theta) > eps, scale=max(abs(sx - 1), abs(sy - 1)) > eps,
                mirror=float(d.scale[0]) * float(d.scale[1]) < 0, uniform=abs(sx - sy) < eps,
                shear=abs(d.shear) > eps, translation=float(np.abs(d.translate).max()) > eps)


def _matrix_svg(M):
    M = np.asarray(M, dtype=np.float64)
    return "matrix({})".format(" ".join(_fmt(v) for v in (M[0, 0], M[1, 0], M[0, 1], M[1, 1], M[0, 2], M[1, 2])))


# ------------------------------------------------------------------- trainable tensors
def _project_colour(t):
    t.clamp_(1e-4, 1.0)


def _project_stops(t):
    t.copy_(torch.sort(t.clamp(0.0, 1.0)).values)
    t[0], t[-1] = 0.0, 1.0


class ParameterBank:
    """Every trainable tensor of a document with its learning rate, optimizer family and
    post-step projection; one torch optimizer per family, a parameter group per tensor."""

    def __init__(self):
        self._entries = []
        self._optimizers = None

    def add(self, tensor, lr, family, project=None):
        tensor.requires_grad_(True)
        self._entries.append((tensor, float(lr), family, project))
        self._optimizers = None
        return tensor

    def tensors(self):
        return [e[0] for e in self._entries]

    def learning_rates(self, tensor):
        return [lr for t, lr, _, _ in self._entries if t is tensor]

    def _build(self):
        groups = {}
        for t, lr, family, _ in self._entries:
            groups.setdefault(family, []).append({"params": [t], "lr": lr})
        self._optimizers = [SvgOptimizationSettings.optims[f](g) for f, g in groups.items()]

    def zero_grad(self):
        for t in self.tensors():
            t.grad = None

    def step(self):
        if self._optimizers is None:
            self._build()
        for opt in self._optimizers:
            opt.step()
        with torch.no_grad():
            for t, _, _, project in self._entries:
                if project is not None:
                    project(t)


class TrainableTransform:
    """One element's transform.  Mode "move" trains rotation and translation, "rigid"
    adds a uniform scale, "similarity" a uniform scale that keeps the parsed mirror, and
    "affine" both scales and the shear.  A parsed transform with parts the mode cannot
    represent is kept as a fixed `residual`, and the trained parameters start from the
    identity: matrix() = residual @ recompose(parameters).  Translations train at the
    transform learning rate times "translation_mult"."""

    MODES = ("move", "rigid", "similarity", "affine")

    def __init__(self, matrix, entry, bank, device):
        cfg = entry["transforms"]
        self.mode = cfg["transform_mode"]
        self.fixed = torch.tensor(matrix, dtype=torch.float32, device=device)
        self.trained = bool(cfg["optimize_transforms"])
        self.residual = None
        if not self.trained:
            return
        if self.mode not in self.MODES:
            raise ValueError("Unrecognized transform mode '{}'".format(self.mode))
        d = TransformTools.decompose(matrix)
        f = transform_features(d)
        cannot = {"move": f["scale"] or f["shear"] or f["mirror"],
                  "rigid": f["shear"] or f["mirror"] or not f["uniform"],
                  "similarity": f["shear"] or not f["uniform"],
                  "affine": False}[self.mode]
        if cannot:
            warnings.warn("transform mode '{}' cannot represent {}; it is kept fixed and trained on top".format(
                self.mode, _matrix_svg(matrix)))
            self.residual = self.fixed.clone()
            d = TransformDecomposition(0.0, np.array([1.0, 1.0]), 0.0, np.zeros(2))

        def leaf(v):
            return torch.tensor(v, dtype=torch.float32, device=device)
        lr, family = cfg["transform_lr"], entry["optimizer"]
        self.theta = bank.add(leaf(d.theta), lr, family)
        self.scale = self.shear = None
        self.mirror = 1.0
        if self.mode in ("rigid", "similarity"):
            self.scale = bank.add(leaf(d.scale[0]), lr, family)
            if self.mode == "similarity":
                self.mirror = float(np.sign(d.scale[0] * d.scale[1]))
        elif self.mode == "affine":
            self.scale = bank.add(leaf(d.scale), lr, family)
            self.shear = bank.add(leaf(d.shear), lr, family)
        self.translate = bank.add(leaf(d.translate), lr * cfg["translation_mult"], family)

    def _parts(self):
        dev = self.theta.device
        if self.scale is None:
            scale = torch.ones((), device=dev)
        elif self.mode == "similarity":
            scale = torch.stack([self.scale, self.scale * self.mirror])
        else:
            scale = self.scale
        shear = self.shear if self.shear is not None else torch.zeros((), device=dev)
        return self.theta, scale, shear, self.translate

    def matrix(self):
        if not self.trained:
            return self.fixed
        m = TransformTools.recompose(*self._parts())
        return m if self.residual is None else self.residual @ m

    def svg(self):
        """The transform as SVG text: the fixed part as a matrix, then the trained
        parts that differ from the identity, in composition order."""
        if not self.trained:
            return _matrix_svg(self.fixed.detach().cpu().numpy())
        theta, scale, shear, t = (p.detach().cpu().double().numpy() for p in self._parts())
        scale = np.broadcast_to(scale, (2,))
        f = transform_features(TransformDecomposition(float(theta), scale, float(shear), t))
        items = [] if self.residual is None else [_matrix_svg(self.residual.detach().cpu().numpy())]
        if f["rotation"]:
            items.append("rotate({})".format(_fmt(math.degrees(float(theta)))))
        if f["scale"] or f["mirror"]:
            items.append("scale({} {})".format(_fmt(scale[0]), _fmt(scale[1])))
        if f["shear"]:
            items.append("skewX({})".format(_fmt(shear)))
        if f["translation"]:
            items.append("translate({} {})".format(_fmt(t[0]), _fmt(t[1])))
        return " ".join(items)


# --------------------------------------------------------------------- document nodes
Paint = namedtuple("Paint", "kind rgb ref")       # kind: "none" | "solid" | "url"
NO_PAINT = Paint("none", None, None)
APPEARANCE_KEYS = ("fill", "fill-opacity", "fill-rule", "opacity", "stroke", "stroke-opacity", "stroke-width")
OPACITY_KEYS = ("opacity", "fill-opacity", "stroke-opacity")
SHAPE_TAGS = ("path", "circle", "rect", "ellipse", "polygon")
# CSS units in pixels at 1 px = 0.25 mm (the reference's unit table: document sizes and
# stroke widths depend on it)
UNITS = {"px": 1.0, "mm": 4.0, "cm": 40.0, "in": 101.6, "pt": 101.6 / 72, "pc": 101.6 / 6}


class SvgNode:
    """One element: `kind` is its tag ("svg", "g", a shape tag or "linearGradient"),
    `transform` a TrainableTransform or None, `style` its own appearance entries and
    `geometry` what its kind needs (paths, a parameter tensor, or gradient data)."""

    def __init__(self, kind, node_id, transform, style, geometry=None):
        self.kind, self.id = kind, node_id
        self.transform, self.style = transform, style
        self.geometry = geometry if geometry is not None else {}
        self.children = []

    def local(self, ctm):
        return ctm if self.transform is None else ctm @ self.transform.matrix()

    # gradient nodes ---------------------------------------------------------------
    def stop_data(self):
        g = self.geometry
        if g.get("href") is not None:
            return g["href"].stop_data()
        return g["offsets"], torch.cat([g["stop_rgb"], g["stop_alpha"].unsqueeze(1)], 1)

    def linear_gradient(self, opacity, ctm):
        """The gradient in canvas space for a shape drawn under `ctm`, with the stop
        alphas scaled by the shape's combined opacity."""
        g = self.geometry
        if g.get("begin") is None or g.get("end") is None:
            raise ValueError("Gradient '{}' has no x1/y1/x2/y2 of its own".format(self.id))
        offsets, stops = self.stop_data()
        stops = torch.cat([stops[:, 0:3], stops[:, 3:4] * opacity.reshape(1, 1)], 1)
        m = self.local(ctm)
        ends = torch.stack([g["begin"], g["end"]]) @ m[0:2, 0:2].t() + m[0:2, 2]
        return LinearGradient(ends[0], ends[1], offsets, stops)


def _default_style(device):
    one = torch.ones(1, device=device)
    return {"fill": Paint("solid", torch.zeros(3, device=device), None), "fill-opacity": one,
            "fill-rule": "nonzero", "opacity": one, "stroke": NO_PAINT, "stroke-opacity": one,
            "stroke-width": torch.zeros(1, device=device)}


def _cascade(inherited, own):
    """Paints, fill rule and stroke width replace the inherited values; opacities multiply."""
    out = dict(inherited)
    for key, value in own.items():
        out[key] = out[key] * value if key in OPACITY_KEYS else value
    return out


def _paint_colour(paint, opacity, ctm, device):
    if paint.kind == "none":
        return None
    if paint.kind == "solid":
        return torch.cat([paint.rgb, opacity.reshape(1)]).to(device)
    return paint.ref.linear_gradient(opacity, ctm)


def _emit_shapes(node, width):
    """The diffvg shapes of one shape element at stroke width `width`."""
    g = node.geometry
    if node.kind == "path":
        return [Path(p.num_control_points, p.points, p.is_closed, width, p.id) for p in g["paths"]]
    v = g["values"]
    if node.kind == "rect":        # x y width height
        return [Rect(v[0:2], v[0:2] + v[2:4], width, node.id)]
    if node.kind == "circle":      # cx cy r
        return [Circle(v[2], v[0:2], width, node.id)]
    if node.kind == "ellipse":     # cx cy rx ry
        return [Ellipse(v[2:4], v[0:2], width, node.id)]
    return [Polygon(v, True, width, node.id)]


# ---------------------------------------------------------------------------- parsing
def parse_length(text):
    """'12.5mm' -> pixels; a bare number is in pixels."""
    m = re.fullmatch(r"\s*({})\s*([a-z]*)\s*".format(_NUMBER), text)
    if m is None:
        raise ValueError("Cannot parse length '{}'".format(text))
    unit = m.group(2) or "px"
    if unit not in UNITS:
        raise ValueError("Unknown or unsupported unit '{}'".format(unit))
    return float(m.group(1)) * UNITS[unit]


def parse_opacity(text):
    """'0.5' or '50%' -> value clamped to [0, 1]."""
    text = text.strip()
    v = float(text[:-1]) / 100 if text.endswith("%") else float(text)
    return min(max(v, 0.0), 1.0)


def parse_colour(text):
    """'#rrggbb' or '#rgb' -> RGB tensor in [0, 1] (no other colour syntax)."""
    h = text.strip()
    if not re.fullmatch(r"#(?:[0-9a-fA-F]{3}){1,2}", h):
        raise ValueError("Color argument `{}` not supported".format(text))
    h = h[1:]
    if len(h) == 3:
        return torch.tensor([int(c, 16) / 15.0 for c in h])
    return torch.tensor([int(h[i:i + 2], 16) / 255.0 for i in (0, 2, 4)])


def colour_hex(rgb):
    """RGB in [0, 1] -> '#rrggbb', truncating each channel to a byte."""
    b = (rgb.detach().cpu() * 255).to(torch.int).clamp(0, 255).tolist()
    return "#" + "".join("{:02x}".format(v) for v in b)


def _declarations(text, keys=None):
    """'k:v;k2:v2' -> {k: v}.  Keys are taken as written, not stripped; the value is the
    text between the first and second colon."""
    out = {}
    for item in text.split(";"):
        if item:
            parts = item.split(":")
            if keys is None or parts[0] in keys:
                out[parts[0]] = parts[1]
    return out


def _stylesheet(text):
    """Class rules of a text/css sheet -> {'.name': 'key: value;key: value'}."""
    text = re.sub(r"/\*.*?\*/", "", text or "", flags=re.S)
    out = {}
    for selector, body in re.findall(r"([^{}]+)\{([^}]*)\}", text):
        selector = " ".join(selector.split())
        if len(selector) < 2 or selector[0] != ".":
            raise ValueError("Unrecognized CSS selector {}".format(selector))
        decls = [d.split(":", 1) for d in body.split(";") if ":" in d]
        out[selector] = ";".join("{}: {}".format(k.strip().lower(), v.strip()) for k, v in decls)
    return out


def _tag(name):
    return name.rsplit("}", 1)[-1]


class _DocumentReader:
    """Builds the node tree of one document, registering trainables in `bank`."""

    def __init__(self, settings, bank, device, verbose):
        self.settings, self.bank, self.device, self.verbose = settings, bank, device, verbose
        self.gradients = {}    # id -> linearGradient node
        self.classes = {}      # '.name' -> declarations
        self.depth = 0

    def log(self, text):
        if self.verbose:
            print("\t" * self.depth + text)

    def entry(self, node_id):
        return self.settings.retrieve(node_id)[0]

    def transform(self, el, node_id, attr="transform"):
        text = el.attrib.get(attr)
        if text is None:
            return None
        return TrainableTransform(parse_transform_list(text), self.entry(node_id), self.bank, self.device)

    def paint(self, text):
        text = text.strip()
        if text == "none":
            return NO_PAINT
        if text.startswith("#"):
            return Paint("solid", parse_colour(text).to(self.device), None)
        m = re.fullmatch(r"url\(\s*['\"]?#?([^'\")]+)['\"]?\s*\)", text)
        if m is None:
            raise ValueError("Unrecognized paint string: '{}'".format(text))
        if m.group(1) not in self.gradients:
            raise ValueError("Paint-type attribute referencing an unknown object with ID '#{}'".format(m.group(1)))
        return Paint("url", None, self.gradients[m.group(1)])

    def appearance(self, el, node_id, trainable=True):
        raw = {}
        cls = el.attrib.get("class")
        if cls is not None and "." + cls in self.classes:
            raw.update(_declarations(self.classes["." + cls], APPEARANCE_KEYS))
        raw.update(_declarations(el.attrib.get("style", ""), APPEARANCE_KEYS))
        raw.update({k: v for k, v in el.attrib.items() if k in APPEARANCE_KEYS})
        entry = self.entry(node_id)
        style = {}
        for key, text in raw.items():
            if key in ("fill", "stroke"):
                p = self.paint(text)
                if trainable and p.kind == "solid" and entry["optimize_color"]:
                    self.bank.add(p.rgb, entry["color_lr"], entry["optimizer"], _project_colour)
                style[key] = p
            elif key in OPACITY_KEYS:
                v = torch.tensor([parse_opacity(text)], dtype=torch.float32, device=self.device)
                if trainable and entry["optimize_alpha"]:
                    self.bank.add(v, entry["alpha_lr"], entry["optimizer"], _project_colour)
                style[key] = v
            elif key == "stroke-width":
                style[key] = torch.tensor([parse_length(text)], dtype=torch.float32, device=self.device)
            else:
                style[key] = text.strip()
        return style

    def skipped(self, el, known, children=True):
        if not self.verbose:
            return
        attrs = sorted(k for k in el.attrib if not k.startswith("{") and k not in known)
        if attrs:
            self.log("Warning: skipping attributes of '{}': {}".format(_tag(el.tag), ", ".join(attrs)))
        if children and len(el):
            self.log("Warning: skipping children of '{}': {}".format(
                _tag(el.tag), ", ".join(_tag(c.tag) for c in el)))

    # elements ------------------------------------------------------------------------
    def document(self, el):
        canvas = self.canvas(el)
        self.settings.global_override(["transforms", "translation_mult"], int(max(canvas)))
        if el.attrib.get("version", "<unknown version>") != "2.0":
            self.log("Warning: Version {} is not 2.0, strange things may happen".format(
                el.attrib.get("version", "<unknown version>")))
        rid = el.attrib.get("id")
        root = SvgNode("svg", rid, self.transform(el, rid), self.appearance(el, rid, trainable=False))
        self.skipped(el, ("width", "height", "id", "transform", "version", "style", "viewBox") + APPEARANCE_KEYS,
                     children=False)
        self.children(el, root)
        return root, canvas

    @staticmethod
    def canvas(el):
        def px(v):
            return int(math.ceil(parse_length(v)))
        if "width" in el.attrib and "height" in el.attrib:
            return np.array([px(el.attrib["width"]), px(el.attrib["height"])])
        if "viewBox" in el.attrib:
            vb = el.attrib["viewBox"].replace(",", " ").split()
            return np.array([px(vb[2]), px(vb[3])])
        raise ValueError("Size information is missing from document definition")

    def children(self, el, parent):
        self.depth += 1
        for child in el:
            tag = _tag(child.tag)
            if tag in SHAPE_TAGS:
                parent.children.append(self.shape(child, tag))
            elif tag == "g":
                parent.children.append(self.group(child))
            elif tag == "defs":
                self.defs(child)
            elif tag == "style":
                self.stylesheet(child)
            else:
                self.log("Warning: skipping element '{}'".format(tag))
        self.depth -= 1

    def group(self, el):
        gid = el.attrib.get("id", "<No ID>")
        self.log("Parsing g#{}".format(gid))
        node = SvgNode("g", gid, self.transform(el, gid), self.appearance(el, gid))
        self.skipped(el, ("id", "transform", "style") + APPEARANCE_KEYS, children=False)
        self.children(el, node)
        return node

    def shape(self, el, tag):
        sid = el.attrib.get("id", "")
        self.log("Parsing {}#{}".format(tag, sid or "<No ID>"))
        a = el.attrib
        if tag == "path":
            paths = from_svg_path(a["d"])
            entry = self.entry(sid)
            for k, p in enumerate(paths):
                p.points = p.points.to(self.device)
                p.num_control_points = p.num_control_points.to(self.device)
                p.id = "{}-{}".format(sid, k) if len(paths) > 1 else sid
                if entry["paths"]["optimize_points"]:
                    self.bank.add(p.points, entry["paths"]["shape_lr"], entry["optimizer"])
            geometry, used = {"paths": paths}, ("d",)
        elif tag == "polygon":
            pts = [[float(x) for x in item.split(",")] for item in a["points"].split(" ") if item]
            if any(len(p) != 2 for p in pts):
                raise ValueError("Malformed polygon points '{}'".format(a["points"]))
            geometry, used = {"values": torch.tensor(pts, dtype=torch.float32, device=self.device)}, ("points",)
        else:
            used = {"rect": ("x", "y", "width", "height"), "circle": ("cx", "cy", "r"),
                    "ellipse": ("cx", "cy", "rx", "ry")}[tag]
            optional = used[:2]    # the position defaults to 0; sizes are required
            vals = [float(a[n]) if (n in a or n not in optional) else 0.0 for n in used]
            geometry = {"values": torch.tensor(vals, dtype=torch.float32, device=self.device)}
        node = SvgNode(tag, sid, self.transform(el, sid), self.appearance(el, sid), geometry)
        self.skipped(el, ("id", "transform", "style", "class") + used + APPEARANCE_KEYS)
        return node

    def stylesheet(self, el):
        if el.attrib.get("type") != "text/css":
            raise ValueError("Only text/css style recognized, got {}".format(el.attrib.get("type")))
        if len(el):
            raise ValueError("Style node should not have children (has {})".format(len(el)))
        self.classes.update(_stylesheet(el.text))

    def defs(self, el):
        self.depth += 1
        for child in el:
            tag = _tag(child.tag)
            if tag == "linearGradient":
                node = self.gradient(child)
                self.gradients[node.id] = node
            elif tag in SHAPE_TAGS:
                raise NotImplementedError("Definition/instantiation of shapes not supported")
            elif tag in ("defs", "g"):
                raise NotImplementedError("'{}' within definitions not supported".format(tag))
            else:
                self.log("Warning: skipping definition '{}'".format(tag))
        self.depth -= 1

    def gradient(self, el):
        gid = el.attrib.get("id", "<No ID>")
        self.log("Parsing linearGradient#{}".format(gid))
        a = el.attrib
        href = next((v for k, v in a.items() if _tag(k) == "href"), None)
        stops = []
        for s in el:
            if _tag(s.tag) != "stop":
                continue
            attrs = {k: v for k, v in s.attrib.items() if k in ("offset", "stop-color", "stop-opacity")}
            attrs.update(_declarations(s.attrib.get("style", "")))
            stops.append((parse_opacity(attrs["offset"]), parse_colour(attrs["stop-color"]).tolist(),
                          parse_opacity(attrs["stop-opacity"]) if "stop-opacity" in attrs else 1.0))
        if not stops and href is None:
            raise ValueError("Gradient {} has neither stops nor a href link to them".format(gid))
        entry = self.entry(gid)
        cfg, family = entry["gradients"], entry["optimizer"]

        def tensor(v):
            return torch.tensor(v, dtype=torch.float32, device=self.device)

        def point(x, y):
            return tensor([float(a.get(x, 0.0)), float(a.get(y, 0.0))]) if (x in a or y in a) else None
        g = {"begin": point("x1", "y1"), "end": point("x2", "y2"),
             "href": self.gradients[href.lstrip("#")] if href is not None else None}
        if stops:
            stops.sort(key=lambda s: s[0])
            g["offsets"] = tensor([s[0] for s in stops])
            g["stop_rgb"] = tensor([s[1] for s in stops])
            g["stop_alpha"] = tensor([s[2] for s in stops])
            if cfg["optimize_stops"]:
                self.bank.add(g["offsets"], cfg["stop_lr"], family, _project_stops)
            if cfg["optimize_color"]:
                self.bank.add(g["stop_rgb"], cfg["color_lr"], family, _project_colour)
            if cfg["optimize_alpha"]:
                self.bank.add(g["stop_alpha"], cfg["alpha_lr"], family, _project_colour)
        if cfg["optimize_location"] and g["begin"] is not None and g["end"] is not None:
            self.bank.add(g["begin"], cfg["location_lr"], family)
            self.bank.add(g["end"], cfg["location_lr"], family)
        return SvgNode("linearGradient", gid, self.transform(el, gid, "gradientTransform"), {}, g)


# ---------------------------------------------------------------------------- writing
def _style_text(style):
    items = []
    for key, value in style.items():
        if isinstance(value, Paint):
            text = {"none": lambda p: "none", "solid": lambda p: colour_hex(p.rgb),
                    "url": lambda p: "url(#{})".format(p.ref.id)}[value.kind](value)
        elif torch.is_tensor(value):
            text = _fmt(value.reshape(-1)[0])
        else:
            text = value
        items.append("{}:{}".format(key, text))
    return ";".join(items) + (";" if items else "")


def _path_data(path):
    """One Path as SVG path data: every segment written out (a closed path's last
    segment ends on its first point), then Z for a closed path."""
    pts = path.points.detach().cpu().double().numpy()
    n = len(pts)
    words = ["M", _fmt(pts[0, 0]), _fmt(pts[0, 1])]
    i = 1
    for k in path.num_control_points.tolist():
        words.append("LQC"[k])
        for j in range(i, i + k + 1):
            words += [_fmt(pts[j % n, 0]), _fmt(pts[j % n, 1])]
        i += k + 1
    if path.is_closed:
        words.append("Z")
    return " ".join(words)


_GEOMETRY_ATTRS = {"rect": ("x", "y", "width", "height"), "circle": ("cx", "cy", "r"),
                   "ellipse": ("cx", "cy", "rx", "ry")}


def _write_node(parent, node):
    el = ET.SubElement(parent, node.kind)
    if node.transform is not None:
        el.set("gradientTransform" if node.kind == "linearGradient" else "transform", node.transform.svg())
    if node.style:
        el.set("style", _style_text(node.style))
    if node.id is not None:
        el.set("id", node.id)
    g = node.geometry
    if node.kind == "path":
        el.set("d", " ".join(_path_data(p) for p in g["paths"]))
    elif node.kind == "polygon":
        el.set("points", " ".join("{},{}".format(_fmt(x), _fmt(y)) for x, y in g["values"].tolist()))
    elif node.kind in _GEOMETRY_ATTRS:
        for name, v in zip(_GEOMETRY_ATTRS[node.kind], g["values"].tolist()):
            el.set(name, _fmt(v))
    elif node.kind == "linearGradient":
        if g["href"] is not None:
            el.set("xlink:href", "#" + g["href"].id)
        else:
            for off, rgb, alpha in zip(g["offsets"].tolist(), g["stop_rgb"].detach(), g["stop_alpha"].tolist()):
                ET.SubElement(el, "stop", {"offset": _fmt(off), "stop-color": colour_hex(rgb),
                                           "stop-opacity": _fmt(alpha)})
        if g["begin"] is not None and g["end"] is not None:
            for name, v in zip(("x1", "y1", "x2", "y2"), torch.cat([g["begin"], g["end"]]).tolist()):
                el.set(name, _fmt(v))
            el.set("gradientUnits", "userSpaceOnUse")
    for c in node.children:
        _write_node(el, c)


# --------------------------------------------------------------------------- document
class OptimizableSvg:
    """An SVG document whose parameters train through the differentiable renderer
    (optimize_svg.py:974-1046 for the interface)."""

    TransformTools = TransformTools

    def __init__(self, filename, settings=None, optimize_background=False, verbose=False,
                 device=torch.device("cpu")):
        self.settings = settings if settings is not None else SvgOptimizationSettings()
        self.device = device
        self.bank = ParameterBank()
        # kept for the interface: `render` composites over nothing
        self.background = torch.tensor([1.0, 1.0, 1.0], device=device)
        if optimize_background:
            p = self.settings.retrieve("default")[0]
            self.bank.add(self.background, p["color_lr"], p["optimizer"], _project_colour)
        reader = _DocumentReader(self.settings, self.bank, device, verbose)
        self.root, self.canvas = reader.document(ET.parse(filename).getroot())
        self.gradients, self.classes = reader.gradients, reader.classes
        self._scene = None

    def nodes(self, kind=None):
        """The elements of the tree in document order (optionally of one kind)."""
        out, todo = [], [self.root]
        while todo:
            n = todo.pop()
            if kind is None or n.kind == kind:
                out.append(n)
            todo.extend(reversed(n.children))
        return out

    def node(self, node_id):
        for n in self.nodes() + list(self.gradients.values()):
            if n.id == node_id:
                return n
        raise KeyError(node_id)

    def build_scene(self):
        """-> (width, height, shapes, shape_groups), rebuilt after each `step`."""
        if self._scene is None:
            shapes, groups = [], []
            ctm = self.root.local(torch.eye(3, device=self.device))
            style = _default_style(self.device)   # the root's own appearance is not inherited
            for c in self.root.children:
                self._emit(c, ctm, style, shapes, groups)
            self._scene = (self.canvas[0], self.canvas[1], shapes, groups)
        return self._scene

    def _emit(self, node, ctm, style, shapes, groups):
        ctm, style = node.local(ctm), _cascade(style, node.style)
        if node.kind == "g":
            for c in node.children:
                self._emit(c, ctm, style, shapes, groups)
            return
        built = _emit_shapes(node, style["stroke-width"])
        fill = _paint_colour(style["fill"], style["opacity"] * style["fill-opacity"], ctm, self.device)
        stroke = _paint_colour(style["stroke"], style["opacity"] * style["stroke-opacity"], ctm, self.device)
        groups.append(ShapeGroup(shape_ids=torch.arange(len(shapes), len(shapes) + len(built)),
                                 fill_color=fill, use_even_odd_rule=style["fill-rule"] == "evenodd",
                                 stroke_color=stroke, shape_to_canvas=ctm, id=node.id))
        shapes.extend(built)

    def render(self, scale=None, seed=0):
        """The scene at 2x2 samples; `scale` scales the output size only (the scene keeps
        its canvas, so the image shows its top-left part at 1:1 when scale > 1)."""
        w, h, shapes, groups = self.build_scene()
        args = RenderFunction.serialize_scene(w, h, shapes, groups)
        if scale is not None:
            w, h = int(w * scale), int(h * scale)
        return RenderFunction.apply(w, h, 2, 2, seed, None, *args)

    def zero_grad(self):
        self.bank.zero_grad()

    def step(self):
        self.bank.step()
        self._scene = None

    def write_xml(self):
        """The document with the current parameter values, as pretty-printed text."""
        el = ET.Element("svg")
        if self.root.transform is not None:
            el.set("transform", self.root.transform.svg())
        if self.root.style:
            el.set("style", _style_text(self.root.style))
        if self.root.id is not None:
            el.set("id", self.root.id)
        for k, v in (("version", "2.0"), ("width", str(self.canvas[0])), ("height", str(self.canvas[1])),
                     ("xmlns", "http://www.w3.org/2000/svg"), ("xmlns:xlink", "http://www.w3.org/1999/xlink")):
            el.set(k, v)
        if self.gradients or self.classes:
            defs = ET.SubElement(el, "defs")
            written = set()

            def put(node):     # a gradient after the one it links to (parsing needs that order)
                if node.id in written:
                    return
                if node.geometry["href"] is not None:
                    put(node.geometry["href"])
                _write_node(defs, node)
                written.add(node.id)
            for node in self.gradients.values():
                put(node)
            sheet = ET.SubElement(el, "style", {"type": "text/css"})
            sheet.text = "".join("{} {{{}}}\n".format(k, v) for k, v in self.classes.items())
        for c in self.root.children:
            _write_node(el, c)
        return minidom.parseString(ET.tostring(el, "utf-8")).toprettyxml(indent="  ")
