"""SVG documents as trainable scenes: ``SvgOptimizationSettings`` and ``OptimizableSvg``
with the names, arguments and results of pydiffvg/optimize_svg.py.

An ``OptimizableSvg`` parses a document into a tree of nodes (root, groups, shapes,
gradients kept in ``defs``).  Every node owns small torch optimizers over its own
parameters -- transform (through a rotation/scale/shear/translation decomposition),
solid colours, path points, gradient stops and end points -- configured per node id by
the settings.  ``build_scene`` walks the tree, composing transforms and inherited
appearance, into the (width, height, shapes, shape_groups) tuple that
``RenderFunction.serialize_scene`` takes; ``render`` draws it with 2x2 samples through
the GPU path; ``step``/``zero_grad`` drive all optimizers; ``write_xml`` writes the
document back.

Where the reference raises on its own inputs, this module computes what the code
plainly means instead (each case is marked "reference:" below).  Other reference
behaviours are kept as they are because they change results: local presentation
attributes override ``style`` declarations, style keys are not stripped, the root's own
appearance is not inherited by its children, rects/circles/ellipses/polygons get an
optimizer but never ``requires_grad`` (so they do not move), and ``stroke-width`` is used
as given, not halved as ``svg_to_scene`` does.
"""
import copy
import json
import math
import re
import xml.etree.ElementTree as etree
from collections import namedtuple
from xml.dom import minidom

import numpy as np
import torch

from .color import LinearGradient
from .render_pytorch import RenderFunction
from .shape import Circle, Ellipse, Path, Polygon, Rect, ShapeGroup, from_svg_path

__all__ = ['SvgOptimizationSettings', 'OptimizableSvg']


class SvgOptimizationSettings:
    """Per-node optimisation settings keyed by SVG id, with a "default" entry
    (optimize_svg.py:15-103).  Stored as plain nested dicts so they save to JSON."""

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

    optims = {
        "Adam": torch.optim.Adam,
        "SGD": torch.optim.SGD,
        "ASGD": torch.optim.ASGD,
    }

    def __init__(self, f=None):
        if f is None:
            self.store = {"default": copy.deepcopy(SvgOptimizationSettings.default_params)}
        else:
            self.store = json.load(f)

    def default_name(self, dname):
        """Names the entry `save` writes back as "default" (the root's id)."""
        self.dname = dname
        if dname not in self.store:
            self.store[dname] = self.store["default"]

    def retrieve(self, node_id):
        """-> (settings for node_id or the defaults, whether node_id has its own)."""
        if node_id in self.store:
            return self.store[node_id], True
        return self.store["default"], False

    def reset_to_defaults(self, node_id):
        self.store.pop(node_id, None)
        return self.store["default"]

    def undefault(self, node_id):
        """Gives node_id its own copy of the defaults (to be edited) and returns it."""
        if node_id not in self.store:
            self.store[node_id] = copy.deepcopy(self.store["default"])
        return self.store[node_id]

    def override_optimizer(self, optimizer):
        if optimizer is not None:
            for entry in self.store.values():
                entry["optimizer"] = optimizer

    def global_override(self, path, value):
        """Sets entry[path[0]]...[path[-1]] = value in every entry."""
        for entry in self.store.values():
            d = entry
            for key in path[:-1]:
                d = d[key]
            d[path[-1]] = value

    def save(self, file):
        # reference: `save` needs `default_name` to have been called first; without it
        # the "default" entry is written as it is
        if hasattr(self, 'dname'):
            self.store["default"] = self.store[self.dname]
        json.dump(self.store, file, indent="\t")


# ---------------------------------------------------------------------------------------
# 2-D affine transforms as 3x3 matrices (optimize_svg.py:108-326)
# ---------------------------------------------------------------------------------------
TransformDecomposition = namedtuple("TransformDecomposition", "theta scale shear translate")
TransformProperties = namedtuple("TransformProperties",
                                 "has_rotation has_scale has_mirror scale_uniform has_shear has_translation")


def _eye3(device=None):
    return torch.eye(3, device=device)


class TransformTools:
    """SVG transform functions as matrices, and the decomposition
    M = [R(theta) diag(sx, sy) [[1, k], [0, 1]] | 0; 0 0 1] @ translate(tx, ty)
    that the transform optimizers work in."""

    TransformDecomposition = TransformDecomposition
    TransformProperties = TransformProperties

    @staticmethod
    def parse_matrix(vals):
        assert len(vals) == 6
        a, b, c, d, e, f = vals
        return np.array([[a, c, e], [b, d, f], [0., 0., 1.]])

    @staticmethod
    def parse_translate(vals):
        assert 1 <= len(vals) <= 2
        m = np.eye(3)
        m[0, 2] = vals[0]
        if len(vals) == 2:
            m[1, 2] = vals[1]
        return m

    @staticmethod
    def parse_rotate(vals):
        """rotate(deg) or rotate(deg cx cy), the latter about (cx, cy)."""
        assert len(vals) in (1, 3)
        rad = math.radians(vals[0])
        c, s = math.cos(rad), math.sin(rad)
        m = np.eye(3)
        m[0:2, 0:2] = [[c, -s], [s, c]]
        if len(vals) == 3:
            # reference: calls an unqualified parse_translate here and raises NameError
            m = TransformTools.parse_translate(vals[1:3]) @ m @ TransformTools.parse_translate([-vals[1], -vals[2]])
        return m

    @staticmethod
    def parse_scale(vals):
        assert 1 <= len(vals) <= 2
        return np.diag([vals[0], vals[-1], 1.])

    @staticmethod
    def parse_skewx(vals):
        """The coefficient itself goes into the matrix (not tan of an angle), which is
        also how the decomposition's shear is written back."""
        assert len(vals) == 1
        m = np.eye(3)
        m[0, 1] = vals[0]
        return m

    @staticmethod
    def parse_skewy(vals):
        assert len(vals) == 1
        m = np.eye(3)
        m[1, 0] = vals[0]
        return m

    @staticmethod
    def transformPoints(pointsTensor, transform):
        """[n, 2] points through a 3x3 matrix (affine part; no projective divide)."""
        assert transform is not None
        out = pointsTensor @ transform[0:2, 0:2].t() + transform[0:2, 2]
        assert out.shape == pointsTensor.shape
        return out.contiguous()

    @staticmethod
    def promote_numpy(M):
        ret = np.eye(3)
        ret[0:2, 0:2] = M
        return ret

    @staticmethod
    def recompose_numpy(Theta, ScaleXY, ShearX, TXY):
        c, s = math.cos(Theta), math.sin(Theta)
        lin = np.array([[c, -s], [s, c]]) @ np.diag(ScaleXY) @ np.array([[1., ShearX], [0., 1.]])
        T = np.eye(3)
        T[0:2, 2] = TXY
        return TransformTools.promote_numpy(lin) @ T

    @staticmethod
    def promote(m):
        M = torch.eye(3, device=m.device, dtype=m.dtype)
        M[0:2, 0:2] = m
        return M

    @staticmethod
    def make_rot(Theta):
        c, s = Theta.cos().squeeze(), Theta.sin().squeeze()
        return torch.stack((torch.stack((c, -s)), torch.stack((s, c))))

    @staticmethod
    def make_scale(ScaleXY):
        sxy = ScaleXY.squeeze()
        if sxy.dim() == 0:
            return torch.diag(torch.stack([sxy, sxy]))
        return torch.diag(sxy)

    @staticmethod
    def make_shear(ShearX):
        m = torch.eye(2, device=ShearX.device)
        m[0, 1] = ShearX
        return m

    @staticmethod
    def make_translate(TXY):
        m = torch.eye(3, device=TXY.device)
        m[0:2, 2] = TXY
        return m

    @staticmethod
    def recompose(Theta, ScaleXY, ShearX, TXY):
        """Differentiable inverse of `decompose`."""
        lin = TransformTools.make_rot(Theta).mm(TransformTools.make_scale(ScaleXY)).mm(TransformTools.make_shear(ShearX))
        return TransformTools.promote(lin).mm(TransformTools.make_translate(TXY))

    @staticmethod
    def make_named(decomp):
        if not isinstance(decomp, TransformDecomposition):
            decomp = TransformDecomposition(theta=decomp[0], scale=decomp[1], shear=decomp[2], translate=decomp[3])
        return decomp

    @staticmethod
    def analyze_transform(decomp):
        """Which parts of a decomposition differ from identity by more than 1e-3."""
        d = TransformTools.make_named(decomp)
        eps = 1e-3
        scale = d.scale
        two = (scale.squeeze().ndim > 0) if isinstance(scale, np.ndarray) else (
            scale.squeeze().dim() > 0 if torch.is_tensor(scale) else len(scale) > 1)
        absd = (lambda v: abs(v))
        if two:
            s0, s1 = scale[0], scale[1]
            has_scale = max(abs(absd(s0) - 1), abs(absd(s1) - 1)) > eps
            has_mirror = bool(s0 * s1 < 0)
            uniform = bool(abs(absd(s0) - absd(s1)) < eps)
        else:
            has_scale = abs(absd(scale) - 1) > eps
            has_mirror = False
            uniform = True
        return TransformProperties(
            has_rotation=bool(abs(d.theta) > eps),
            has_scale=bool(has_scale),
            has_mirror=has_mirror,
            scale_uniform=uniform,
            has_shear=bool(abs(d.shear) > eps),
            has_translation=bool(max(abs(d.translate[0]), abs(d.translate[1])) > eps))

    @staticmethod
    def check_and_decomp(M):
        if M is None:
            decomp = TransformDecomposition(theta=0., scale=np.array([1., 1.]), shear=0., translate=np.array([0., 0.]))
        else:
            decomp = TransformTools.decompose(M)
        return decomp, TransformTools.analyze_transform(decomp)

    @staticmethod
    def tf_to_string(M):
        return "matrix({} {} {} {} {} {})".format(M[0, 0], M[1, 0], M[0, 1], M[1, 1], M[0, 2], M[1, 2])

    @staticmethod
    def decomp_to_string(decomp):
        """SVG transform list of the non-identity parts, in composition order."""
        d = TransformTools.make_named(decomp)
        props = TransformTools.analyze_transform(d)
        out = ""
        if props.has_rotation:
            out += "rotate({}) ".format(math.degrees(float(d.theta)))
        if props.has_scale:
            if d.scale.dim() == 0:
                out += "scale({}) ".format(d.scale.item())
            else:
                out += "scale({} {}) ".format(float(d.scale[0]), float(d.scale[1]))
        if props.has_shear:
            out += "skewX({}) ".format(float(d.shear))
        if props.has_translation:
            out += "translate({} {}) ".format(float(d.translate[0]), float(d.translate[1]))
        return out

    @staticmethod
    def decompose(M):
        """M (3x3, affine) -> (theta, [sx, sy], shear, [tx, ty]) with
        M = rot(theta) diag(sx, sy) shearX(shear) translate(tx, ty) and sx > 0.

        With A = M[0:2, 0:2]: the first column of A is rot(theta) (sx, 0), so
        sx = |A[:, 0]| and theta = atan2(A10, A00); rot(-theta) A is upper triangular
        [[sx, sx shear], [0, sy]], so sy = det(A) / sx and shear = (rot(-theta) A)01 / sx;
        the translation is A^-1 M[0:2, 2] (it is applied first).  This is the QR
        factorisation with signs normalised so that sx > 0, which is what
        optimize_svg.py:287-326 computes through numpy's QR."""
        A = np.asarray(M, dtype=np.float64)[0:2, 0:2]
        t = np.asarray(M, dtype=np.float64)[0:2, 2]
        sx = math.hypot(A[0, 0], A[1, 0])
        theta = math.atan2(A[1, 0], A[0, 0])
        c, s = math.cos(theta), math.sin(theta)
        sy = (A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0]) / sx
        shear = (c * A[0, 1] + s * A[1, 1]) / sx
        return TransformDecomposition(theta=theta, scale=np.array([sx, sy]), shear=shear,
                                      translate=np.linalg.solve(A, t))


# ---------------------------------------------------------------------------------------
# per-parameter optimizers (optimize_svg.py:328-505)
# ---------------------------------------------------------------------------------------
def _optim_class(params):
    return SvgOptimizationSettings.optims[params["optimizer"]]


class ColorOptimizer:
    """An optimizer over a tensor kept within [1e-4, 1] per entry."""

    def __init__(self, tensor, optim_type, lr):
        self.tensor = tensor
        self.optim = optim_type([tensor], lr=lr)

    def zero_grad(self):
        self.optim.zero_grad()

    def step(self):
        self.optim.step()
        self.tensor.data.clamp_(min=1e-4, max=1.)


class StopOptimizer:
    """Gradient stop offsets: clamped to [0, 1], kept sorted, ends pinned to 0 and 1."""

    def __init__(self, stops, optim_type, lr):
        self.stops = stops
        self.optim = optim_type([stops], lr=lr)

    def zero_grad(self):
        self.optim.zero_grad()

    def step(self):
        self.optim.step()
        self.stops.data.clamp_(min=0., max=1.)
        self.stops.data, _ = self.stops.data.sort()
        self.stops.data[0] = 0.
        self.stops.data[-1] = 1.


class GradientOptimizer:
    """A linear gradient's end points, stop offsets, stop colours and stop alphas, each
    optimised when its setting says so."""

    def __init__(self, begin, end, offsets, stops, optim_params):
        def own(t):
            return t.clone().detach() if t is not None else None
        self.begin, self.end, self.offsets = own(begin), own(end), own(offsets)
        self.stop_colors = own(stops[:, 0:3]) if stops is not None else None
        self.stop_alphas = own(stops[:, 3]) if stops is not None else None
        self.optimizers = []
        g = optim_params["gradients"]
        opt = _optim_class(optim_params)
        if g["optimize_stops"] and self.offsets is not None:
            self.offsets.requires_grad_(True)
            self.optimizers.append(StopOptimizer(self.offsets, opt, g["stop_lr"]))
        if g["optimize_color"] and self.stop_colors is not None:
            self.stop_colors.requires_grad_(True)
            self.optimizers.append(ColorOptimizer(self.stop_colors, opt, g["color_lr"]))
        if g["optimize_alpha"] and self.stop_alphas is not None:
            self.stop_alphas.requires_grad_(True)
            self.optimizers.append(ColorOptimizer(self.stop_alphas, opt, g["alpha_lr"]))
        if g["optimize_location"] and self.begin is not None and self.end is not None:
            self.begin.requires_grad_(True)
            self.end.requires_grad_(True)
            self.optimizers.append(opt([self.begin, self.end], lr=g["location_lr"]))

    def get_vals(self):
        stops = None
        if self.stop_colors is not None and self.stop_alphas is not None:
            stops = torch.cat((self.stop_colors, self.stop_alphas.unsqueeze(1)), 1)
        return self.begin, self.end, self.offsets, stops

    def zero_grad(self):
        for o in self.optimizers:
            o.zero_grad()

    def step(self):
        for o in self.optimizers:
            o.step()


class TransformOptimizer:
    """A node's transform as trainable decomposition parameters.

    transform_mode "move" trains rotation and translation, "rigid" adds a uniform scale,
    "similarity" a uniform scale that may mirror, "affine" everything (2 scales and a
    shear).  When the parsed transform has a part the mode cannot represent, it is kept
    as a fixed `residual` and the trained parameters start from identity:
    transform = residual @ recompose(parameters).  Translations train with the learning
    rate times "translation_mult" (set to the larger canvas side by the document)."""

    MODES = ("move", "rigid", "similarity", "affine")

    def __init__(self, transform, optim_params):
        self.transform = transform
        self.optimizes = optim_params["transforms"]["optimize_transforms"] and transform is not None
        self.params = copy.deepcopy(optim_params)
        self.transform_mode = optim_params["transforms"]["transform_mode"]
        if not self.optimizes:
            return
        if self.transform_mode not in self.MODES:
            raise ValueError("Unrecognized transform mode '{}'".format(self.transform_mode))
        lr = optim_params["transforms"]["transform_lr"]
        tmult = optim_params["transforms"]["translation_mult"]
        dev = transform.device
        decomp, props = TransformTools.check_and_decomp(transform.cpu().numpy())

        def param(v, grad=True):
            return torch.tensor(v, dtype=torch.float32, requires_grad=grad, device=dev)

        mode = self.transform_mode
        unsupported = {
            "move": props.has_scale or props.has_shear or props.has_mirror,
            "rigid": props.has_shear or props.has_mirror or not props.scale_uniform,
            "similarity": props.has_shear or not props.scale_uniform,
            "affine": False,
        }[mode]
        self.residual = None
        if unsupported:
            print("Warning: set to optimize {} transforms only, but the input transform has parts "
                  "that mode cannot represent; they are kept fixed".format(mode))
            self.residual = transform.clone().detach().requires_grad_(False)
            self.Theta, self.translation = param(0.), param([0., 0.])
            if mode in ("rigid", "similarity"):
                self.scale = param(1.)
            if mode == "similarity":
                self.scale_sign = param(1., grad=False)
        else:
            self.Theta, self.translation = param(decomp.theta), param(decomp.translate)
            if mode in ("rigid", "similarity"):
                self.scale = param(decomp.scale[0])
            if mode == "similarity":
                self.scale_sign = param(float(np.sign(decomp.scale[0] * decomp.scale[1])), grad=False)
            if mode == "affine":
                self.scale, self.shear = param(decomp.scale), param(decomp.shear)
        groups = [{'params': self.Theta, 'lr': lr}]
        if mode in ("rigid", "similarity"):
            groups.append({'params': self.scale, 'lr': lr})
        if mode == "affine":
            groups += [{'params': self.scale, 'lr': lr}, {'params': self.shear, 'lr': lr}]
        groups.append({'params': self.translation, 'lr': lr * tmult})
        self.optimizer = _optim_class(optim_params)(groups)

    def _parts(self):
        dev = self.Theta.device
        mode = self.transform_mode
        if mode == "move":
            return self.Theta, torch.tensor([1.], device=dev), torch.tensor(0., device=dev), self.translation
        if mode == "rigid":
            return self.Theta, self.scale, torch.tensor(0., device=dev), self.translation
        if mode == "similarity":
            # reference: torch.cat of two 0-d tensors, which raises
            return self.Theta, torch.stack((self.scale, self.scale * self.scale_sign)), \
                torch.tensor(0., device=dev), self.translation
        return self.Theta, self.scale, self.shear, self.translation

    def get_transform(self):
        if not self.optimizes:
            return self.transform
        composed = TransformTools.recompose(*self._parts())
        return self.residual.mm(composed) if self.residual is not None else composed

    def tfToString(self):
        if self.transform is None:
            return None
        if not self.optimizes:
            return TransformTools.tf_to_string(self.transform)
        parts = tuple(p.detach().cpu() for p in self._parts())
        text = TransformTools.decomp_to_string(parts)
        if self.transform_mode == "affine":
            # reference: returns an undefined name here
            return text
        return (TransformTools.tf_to_string(self.residual) if self.residual is not None else "") + " " + text

    def zero_grad(self):
        if self.optimizes:
            self.optimizer.zero_grad()

    def step(self):
        if self.optimizes:
            self.optimizer.step()


# ---------------------------------------------------------------------------------------
# document tree (optimize_svg.py:507-972)
# ---------------------------------------------------------------------------------------
def _num(v):
    """A tensor or number as the text of a number."""
    return str(v.item() if torch.is_tensor(v) else v)


class SvgNode:
    """A node with an optional transform, its own appearance entries and optimizers."""

    def __init__(self, id, transform, appearance, settings):
        self.id = id
        self.children = []
        self.optimizers = []
        self.device = settings.device
        self.transform = torch.tensor(transform, dtype=torch.float32, device=self.device) \
            if transform is not None else None
        params = settings.retrieve(self.id)[0]
        self.transform_optim = TransformOptimizer(self.transform, params)
        self.optimizers.append(self.transform_optim)
        self.proc_appearance(appearance, params)

    def tftostring(self):
        return self.transform_optim.tfToString()

    def appearanceToString(self):
        out = ""
        for key, value in self.appearance.items():
            if key in ("fill", "stroke"):
                kind = value[0]
                if kind == "none":
                    out += "{}:none;".format(key)
                elif kind == "solid":
                    out += "{}:{};".format(key, OptimizableSvg.rgb_to_string(value[1]))
                elif kind == "url":
                    out += "{}:url(#{});".format(key, value[1].id)
            elif key in ("opacity", "fill-opacity", "stroke-opacity", "stroke-width"):
                # reference: writes the tensor's repr, which does not parse back
                out += "{}:{};".format(key, _num(value))
            elif key == "fill-rule":
                out += "{}:{};".format(key, value)
            else:
                raise ValueError("Don't know how to write appearance parameter '{}'".format(key))
        return out

    def write_xml_common_attrib(self, node, tfname="transform"):
        if self.transform is not None:
            node.set(tfname, self.tftostring())
        if len(self.appearance) > 0:
            node.set('style', self.appearanceToString())
        if self.id is not None:
            node.set('id', self.id)

    def proc_appearance(self, appearance, optim_params):
        """Colours (and opacities when "optimize_alpha" is set) become trainable."""
        self.appearance = appearance
        opt = _optim_class(optim_params)
        for key, value in appearance.items():
            if key in ("fill", "stroke"):
                if optim_params["optimize_color"] and value[0] == "solid":
                    value[1].requires_grad_(True)
                    self.optimizers.append(ColorOptimizer(value[1], opt, optim_params["color_lr"]))
            elif key in ("fill-opacity", "stroke-opacity", "opacity"):
                if optim_params["optimize_alpha"]:
                    # reference: indexes the 0-d opacity and passes the optimizer's name,
                    # which raises
                    value.requires_grad_(True)
                    self.optimizers.append(ColorOptimizer(value, opt, optim_params["alpha_lr"]))
            elif key not in ("fill-rule", "stroke-width"):
                raise RuntimeError("Unrecognized appearance key '{}'".format(key))

    def prop_transform(self, intform):
        return intform.matmul(self.transform_optim.get_transform()) if self.transform is not None else intform

    def prop_appearance(self, inappearance):
        """Paints, fill rule and stroke width replace the inherited ones; opacities
        multiply them."""
        out = copy.copy(inappearance)
        for key, value in self.appearance.items():
            if key in ("fill-opacity", "opacity", "stroke-opacity"):
                out[key] = out[key] * value
            elif key in ("fill", "fill-rule", "stroke", "stroke-width"):
                out[key] = value
            else:
                raise RuntimeError("Unrecognized appearance key '{}'".format(key))
        return out

    def zero_grad(self):
        for o in self.optimizers:
            o.zero_grad()
        for c in self.children:
            c.zero_grad()

    def step(self):
        for o in self.optimizers:
            o.step()
        for c in self.children:
            c.step()

    def get_type(self):
        return "Generic node"

    def is_shape(self):
        return False

    def build_scene(self, shapes, shape_groups, transform, appearance):
        raise NotImplementedError("Abstract SvgNode cannot recurse")


class GroupNode(SvgNode):
    def get_type(self):
        return "Group node"

    def build_scene(self, shapes, shape_groups, transform, appearance):
        tf = self.prop_transform(transform)
        app = self.prop_appearance(appearance)
        for c in self.children:
            c.build_scene(shapes, shape_groups, tf, app)

    def write_xml(self, parent):
        elm = etree.SubElement(parent, "g")
        self.write_xml_common_attrib(elm)
        for c in self.children:
            c.write_xml(elm)


class RootNode(SvgNode):
    def write_xml(self, document):
        elm = etree.Element('svg')
        self.write_xml_common_attrib(elm)
        elm.set("version", "2.0")
        elm.set("width", str(document.canvas[0]))
        elm.set("height", str(document.canvas[1]))
        elm.set("xmlns", "http://www.w3.org/2000/svg")
        elm.set("xmlns:xlink", "http://www.w3.org/1999/xlink")
        document.write_defs(elm)
        for c in self.children:
            c.write_xml(elm)
        return elm

    def get_type(self):
        return "Root node"

    def build_scene(self, shapes, shape_groups, transform, appearance):
        # the root's own appearance is not passed on (as the reference does)
        tf = self.prop_transform(transform).to(self.device)
        for c in self.children:
            c.build_scene(shapes, shape_groups, tf, appearance)

    @staticmethod
    def get_default_appearance(device):
        return {"fill": ("solid", torch.tensor([0., 0., 0.], device=device)),
                "fill-opacity": torch.tensor([1.], device=device),
                "fill-rule": "nonzero",
                "opacity": torch.tensor([1.], device=device),
                "stroke": ("none", None),
                "stroke-opacity": torch.tensor([1.], device=device),
                "stroke-width": torch.tensor([0.], device=device)}

    @staticmethod
    def get_default_transform():
        return torch.eye(3)


class ShapeNode(SvgNode):
    def get_type(self):
        return "Generic shape node"

    def is_shape(self):
        return True

    def construct_paint(self, value, combined_opacity, transform):
        kind = value[0]
        if kind == "none":
            return None
        if kind == "solid":
            return torch.cat([value[1], combined_opacity]).to(self.device)
        if kind == "url":
            return value[1].getGrad(combined_opacity, transform)
        raise ValueError("Unknown paint value type '{}'".format(kind))

    def make_shape_group(self, appearance, transform, num_shapes, num_subobjects):
        fill = self.construct_paint(appearance["fill"], appearance["opacity"] * appearance["fill-opacity"], transform)
        stroke = self.construct_paint(appearance["stroke"], appearance["opacity"] * appearance["stroke-opacity"],
                                      transform)
        return ShapeGroup(shape_ids=torch.tensor(range(num_shapes, num_shapes + num_subobjects)),
                          fill_color=fill,
                          use_even_odd_rule=appearance["fill-rule"] == "evenodd",
                          stroke_color=stroke,
                          shape_to_canvas=transform,
                          id=self.id)

    def _emit(self, shapes, shape_groups, transform, appearance, make):
        tf = self.prop_transform(transform)
        app = self.prop_appearance(appearance)
        built = make(app["stroke-width"])
        shape_groups.append(self.make_shape_group(app, tf, len(shapes), len(built)))
        shapes.extend(built)

    def _shape_optimizer(self, settings, tensor):
        """Shapes other than paths get an optimizer under the "paths" settings, but their
        tensor never requires grad (as in the reference), so they stay where they are."""
        params = settings.retrieve(self.id)[0]
        if params["paths"]["optimize_points"]:
            self.optimizers.append(_optim_class(params)([tensor], lr=params["paths"]["shape_lr"]))


class PathNode(ShapeNode):
    def __init__(self, id, transform, appearance, settings, paths):
        super().__init__(id, transform, appearance, settings)
        self.proc_paths(paths, settings.retrieve(self.id)[0])

    def proc_paths(self, paths, optim_params):
        self.paths = paths
        if optim_params["paths"]["optimize_points"]:
            pts = [p.points.requires_grad_(True) for p in paths]
            self.optimizers.append(_optim_class(optim_params)(pts, lr=optim_params["paths"]["shape_lr"]))

    def get_type(self):
        return "Path node"

    def build_scene(self, shapes, shape_groups, transform, appearance):
        self._emit(shapes, shape_groups, transform, appearance,
                   lambda width: [Path(p.num_control_points, p.points, p.is_closed, width, p.id) for p in self.paths])

    def path_to_string(self, path):
        pts = path.points.detach()
        n = pts.shape[0]

        def xy(i):
            return "{},{} ".format(pts[i % n][0].item(), pts[i % n][1].item())
        out = "M " + xy(0)
        idx = 1
        for k in path.num_control_points.tolist():
            out += ("L ", "Q ", "C ")[k]
            for _ in range(k + 1):
                out += xy(idx)
                idx += 1
        if path.is_closed:
            out += "Z "
        return out

    def paths_string(self):
        return "".join(self.path_to_string(p) for p in self.paths)

    def write_xml(self, parent):
        elm = etree.SubElement(parent, "path")
        self.write_xml_common_attrib(elm)
        elm.set("d", self.paths_string())
        for c in self.children:
            c.write_xml(elm)


class RectNode(ShapeNode):
    def __init__(self, id, transform, appearance, settings, rect):
        super().__init__(id, transform, appearance, settings)
        self.rect = torch.tensor(rect, dtype=torch.float, device=settings.device)
        self._shape_optimizer(settings, self.rect)

    def get_type(self):
        return "Rect node"

    def build_scene(self, shapes, shape_groups, transform, appearance):
        r = self.rect
        self._emit(shapes, shape_groups, transform, appearance,
                   lambda width: [Rect(r[0:2], r[0:2] + r[2:4], width, self.id)])

    def write_xml(self, parent):
        # reference: writes the tensors' reprs for the four numbers
        elm = etree.SubElement(parent, "rect")
        self.write_xml_common_attrib(elm)
        for name, v in zip(("x", "y", "width", "height"), self.rect):
            elm.set(name, _num(v))
        for c in self.children:
            c.write_xml(elm)


class CircleNode(ShapeNode):
    def __init__(self, id, transform, appearance, settings, rect):
        super().__init__(id, transform, appearance, settings)
        self.circle = torch.tensor(rect, dtype=torch.float, device=settings.device)
        self._shape_optimizer(settings, self.circle)

    def get_type(self):
        return "Circle node"

    def build_scene(self, shapes, shape_groups, transform, appearance):
        c = self.circle
        self._emit(shapes, shape_groups, transform, appearance,
                   lambda width: [Circle(c[2], c[0:2], width, self.id)])

    def write_xml(self, parent):
        elm = etree.SubElement(parent, "circle")
        self.write_xml_common_attrib(elm)
        for name, v in zip(("cx", "cy", "r"), self.circle):
            elm.set(name, _num(v))
        for c in self.children:
            c.write_xml(elm)


class EllipseNode(ShapeNode):
    def __init__(self, id, transform, appearance, settings, ellipse):
        super().__init__(id, transform, appearance, settings)
        self.ellipse = torch.tensor(ellipse, dtype=torch.float, device=settings.device)
        self._shape_optimizer(settings, self.ellipse)

    def get_type(self):
        return "Ellipse node"

    def build_scene(self, shapes, shape_groups, transform, appearance):
        e = self.ellipse
        self._emit(shapes, shape_groups, transform, appearance,
                   lambda width: [Ellipse(e[2:4], e[0:2], width, self.id)])

    def write_xml(self, parent):
        elm = etree.SubElement(parent, "ellipse")
        self.write_xml_common_attrib(elm)
        for name, v in zip(("cx", "cy", "rx", "ry"), self.ellipse):
            elm.set(name, _num(v))
        for c in self.children:
            c.write_xml(elm)


class PolygonNode(ShapeNode):
    def __init__(self, id, transform, appearance, settings, points):
        super().__init__(id, transform, appearance, settings)
        self.points = points
        self._shape_optimizer(settings, self.points)

    def get_type(self):
        return "Polygon node"

    def build_scene(self, shapes, shape_groups, transform, appearance):
        self._emit(shapes, shape_groups, transform, appearance,
                   lambda width: [Polygon(self.points, True, width, self.id)])

    def point_string(self):
        return "".join("{},{} ".format(_num(p[0]), _num(p[1])) for p in self.points.detach())

    def write_xml(self, parent):
        elm = etree.SubElement(parent, "polygon")
        self.write_xml_common_attrib(elm)
        elm.set("points", self.point_string())
        for c in self.children:
            c.write_xml(elm)


class GradientNode(SvgNode):
    """A linearGradient from <defs>; `href` is the gradient whose stops it uses."""

    def __init__(self, id, transform, settings, begin, end, offsets, stops, href):
        super().__init__(id, transform, {}, settings)
        self.optim = GradientOptimizer(begin, end, offsets, stops, settings.retrieve(id)[0])
        self.optimizers.append(self.optim)
        self.href = href

    def is_ref(self):
        return self.href is not None

    def get_type(self):
        return "Gradient node"

    def get_stops(self):
        _, _, offsets, stops = self.optim.get_vals()
        return offsets, stops

    def get_points(self):
        begin, end, _, _ = self.optim.get_vals()
        return begin, end

    def write_xml(self, parent):
        elm = etree.SubElement(parent, "linearGradient")
        self.write_xml_common_attrib(elm, tfname="gradientTransform")
        begin, end, offsets, stops = self.optim.get_vals()
        if self.href is None:
            for k, off in enumerate(offsets):
                stop = etree.SubElement(elm, "stop")
                stop.set("offset", _num(off))
                stop.set("stop-color", OptimizableSvg.rgb_to_string(stops[k, 0:3]))
                stop.set("stop-opacity", _num(stops[k, 3]))
        else:
            elm.set('xlink:href', "#{}".format(self.href.id))
        if begin is not None and end is not None:
            elm.set('x1', _num(begin[0]))
            elm.set('y1', _num(begin[1]))
            elm.set('x2', _num(end[0]))
            elm.set('y2', _num(end[1]))
            elm.set("gradientUnits", "userSpaceOnUse")
        for c in self.children:
            c.write_xml(elm)

    def getGrad(self, combined_opacity, transform):
        """-> LinearGradient in canvas space: end points through the shape's transform
        then this gradient's own; stop alphas times the shape's combined opacity."""
        offsets, stops = self.href.get_stops() if self.is_ref() else self.get_stops()
        stops = stops.clone()
        stops[:, 3] *= combined_opacity
        begin, end = self.get_points()
        if begin is None or end is None:
            raise ValueError("Gradient '{}' has no x1/y1/x2/y2 of its own".format(self.id))
        tf = self.prop_transform(transform)
        begin = TransformTools.transformPoints(begin.unsqueeze(0), tf).squeeze()
        end = TransformTools.transformPoints(end.unsqueeze(0), tf).squeeze()
        return LinearGradient(begin, end, offsets, stops)


def _css_rules(text):
    """'.name { a: b; c: d }' rules -> [(selector, "a: b;c: d")] (comments removed;
    declarations joined the way cssutils serialises them once newlines are dropped)."""
    text = re.sub(r'/\*.*?\*/', '', text or '', flags=re.S)
    out = []
    for m in re.finditer(r'([^{}]+)\{([^}]*)\}', text):
        decls = []
        for d in m.group(2).split(';'):
            if ':' in d:
                k, v = d.split(':', 1)
                decls.append("{}: {}".format(k.strip().lower(), v.strip()))
        out.append((' '.join(m.group(1).split()), ';'.join(decls)))
    return out


def _split_decls(text, keys=None):
    """'k:v;k2:v2' -> {k: v}; keys are taken as written (not stripped), the value is
    what lies between the first and second colon."""
    out = {}
    for item in text.split(";"):
        if len(item) == 0:
            continue
        parts = item.split(":")
        if keys is None or parts[0] in keys:
            out[parts[0]] = parts[1]
    return out


class OptimizableSvg:
    """An SVG document whose parameters train through the differentiable renderer
    (optimize_svg.py:106-1607)."""

    TransformTools = TransformTools
    ColorOptimizer = ColorOptimizer
    StopOptimizer = StopOptimizer
    GradientOptimizer = GradientOptimizer
    TransformOptimizer = TransformOptimizer
    SvgNode = SvgNode
    GroupNode = GroupNode
    RootNode = RootNode
    ShapeNode = ShapeNode
    PathNode = PathNode
    RectNode = RectNode
    CircleNode = CircleNode
    EllipseNode = EllipseNode
    PolygonNode = PolygonNode
    GradientNode = GradientNode

    recognised_shapes = ["path", "circle", "rect", "ellipse", "polygon"]
    appearance_keys = ["fill", "fill-opacity", "fill-rule", "opacity", "stroke", "stroke-opacity", "stroke-width"]
    # CSS units in pixels, with 1 px = 0.25 mm
    unit_dict = {"px": 1, "mm": 4, "cm": 40, "in": 25.4 * 4, "pt": 25.4 * 4 / 72, "pc": 25.4 * 4 / 6}

    def __init__(self, filename, settings=None, optimize_background=False, verbose=False,
                 device=torch.device("cpu")):
        self.settings = settings if settings is not None else SvgOptimizationSettings()
        self.verbose = verbose
        self.device = device
        self.settings.device = device
        root = etree.parse(filename).getroot()
        self.optimizers = []
        # kept for interface parity: `render` composites over nothing, like the reference
        self.background = torch.tensor([1., 1., 1.], dtype=torch.float32, requires_grad=optimize_background,
                                       device=self.device)
        if optimize_background:
            p = self.settings.retrieve("default")[0]
            self.optimizers.append(ColorOptimizer(self.background, _optim_class(p), p["color_lr"]))
        self.defs = {}
        self.depth = 0
        self.dirty = True
        self.scene = None
        self.parseRoot(root)

    # ---- core ---------------------------------------------------------------------------
    def build_scene(self):
        """-> (width, height, shapes, shape_groups); rebuilt after each `step`."""
        if self.dirty:
            shapes, groups = [], []
            self.root.build_scene(shapes, groups, RootNode.get_default_transform().to(self.device),
                                  RootNode.get_default_appearance(self.device))
            self.scene = (self.canvas[0], self.canvas[1], shapes, groups)
            self.dirty = False
        return self.scene

    def _def_nodes(self):
        return [v for v in self.defs.values() if isinstance(v, SvgNode)]

    def zero_grad(self):
        self.root.zero_grad()
        for o in self.optimizers:
            o.zero_grad()
        for n in self._def_nodes():
            n.zero_grad()

    def render(self, scale=None, seed=0):
        """The scene at 2x2 samples; `scale` scales the output size only (the scene is
        not scaled, so the image shows the top-left part of the canvas at 1:1)."""
        scene = self.build_scene()
        args = RenderFunction.serialize_scene(*scene)
        w, h = (scene[0], scene[1]) if scale is None else (int(scene[0] * scale), int(scene[1] * scale))
        return RenderFunction.apply(w, h, 2, 2, seed, None, *args)

    def step(self):
        self.dirty = True
        self.root.step()
        for o in self.optimizers:
            o.step()
        for n in self._def_nodes():
            n.step()

    # ---- reporting ------------------------------------------------------------------------
    def offset_str(self, s):
        return ("\t" * self.depth) + s

    def _label(self, node):
        tag = OptimizableSvg.remove_namespace(node.tag)
        return tag if "id" not in node.attrib else "{}#{}".format(tag, node.attrib["id"])

    def reportSkippedAttribs(self, node, non_skipped=[]):
        skipped = set(k for k in node.attrib if not OptimizableSvg.is_namespace(k)) - set(non_skipped)
        if skipped:
            print(self.offset_str("Warning: Skipping the following attributes of node '{}': {}".format(
                self._label(node), ", ".join("'{}'".format(a) for a in skipped))))

    def reportSkippedChildren(self, node, skipped):
        if skipped:
            names = ["{}#{}".format(e.tag, e.attrib["id"]) if "id" in e.attrib else e.tag for e in skipped]
            print(self.offset_str("Warning: Skipping the following children of node '{}': {}".format(
                self._label(node), ", ".join("'{}'".format(n) for n in names))))

    # ---- parsing ---------------------------------------------------------------------------
    @staticmethod
    def remove_namespace(s):
        """'{uri}tag' -> 'tag'."""
        return re.sub('{.*}', '', s)

    @staticmethod
    def is_namespace(s):
        return re.match('{.*}', s) is not None

    @staticmethod
    def parseTransform(node):
        """transform (or gradientTransform) attribute -> 3x3 numpy matrix, or None."""
        text = node.attrib.get("transform", node.attrib.get("gradientTransform"))
        if text is None:
            return None
        parsers = {"matrix": TransformTools.parse_matrix, "translate": TransformTools.parse_translate,
                   "rotate": TransformTools.parse_rotate, "scale": TransformTools.parse_scale,
                   "skewX": TransformTools.parse_skewx, "skewY": TransformTools.parse_skewy}
        mat = np.eye(3)
        for item in text.split(")")[:-1]:
            # reference: a space before the function name or inside the parentheses
            # makes it raise
            kind, _, arg_text = item.partition("(")
            kind = kind.strip().lstrip(",").strip()
            if kind not in parsers:
                raise ValueError("Unknown transform type '{}'".format(kind))
            args = [float(v) for v in re.split("[, ]+", arg_text.strip())]
            mat = mat @ parsers[kind](args)
        return mat

    @staticmethod
    def parseLength(s):
        """'12.5mm' -> pixels (a bare number is in pixels)."""
        s = s.strip()
        m = re.fullmatch(r'([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)\s*([a-z]*)', s)
        if m is None:
            raise ValueError("Cannot parse length '{}'".format(s))
        val, unit = float(m.group(1)), m.group(2)
        if unit == "":
            return val
        if unit not in OptimizableSvg.unit_dict:
            raise ValueError("Unknown or unsupported unit '{}' encountered while parsing".format(unit))
        return val * OptimizableSvg.unit_dict[unit]

    @staticmethod
    def parseOpacity(s):
        s = s.strip()
        pct = s.endswith("%")
        val = float(s.rstrip("%"))
        return np.clip(val / 100 if pct else val, 0., 1.)

    @staticmethod
    def parse_color(s):
        """'#rrggbb' or '#rgb' -> RGB tensor in [0, 1]; nothing else is accepted."""
        s = s.strip()
        if not s.startswith('#'):
            raise ValueError("Color argument `{}` not supported".format(s))
        h = s[1:]
        if len(h) == 6:
            return torch.tensor([int(h[i:i + 2], 16) / 255.0 for i in (0, 2, 4)])
        if len(h) == 3:
            return torch.tensor([int(h[i], 16) / 15.0 for i in (0, 1, 2)])
        raise ValueError("Color argument `{}` not supported".format(s))

    @staticmethod
    def rgb_to_string(val):
        """RGB in [0, 1] -> '#rrggbb', truncating to bytes."""
        b = (val.clone().detach() * 255).type(torch.int).clamp_(min=0, max=255).tolist()
        return "#{:02x}{:02x}{:02x}".format(*b)

    @staticmethod
    def parsePaint(paintStr, defs, device):
        """-> ("none", None) | ("solid", rgb) | ("url", gradient node)."""
        p = paintStr.strip()
        if p == "none":
            return ("none", None)
        if p.startswith("#"):
            return ("solid", OptimizableSvg.parse_color(p).to(device))
        if p.startswith("url"):
            ref = re.sub(r'^url\(|\)$', '', p).strip().strip("'\"").lstrip("#")
            if ref not in defs:
                raise ValueError("Paint-type attribute referencing an unknown object with ID '#{}'".format(ref))
            return ("url", defs[ref])
        raise ValueError("Unrecognized paint string: '{}'".format(paintStr))

    @staticmethod
    def parseAppearance(node, defs, device):
        """Appearance entries of a node from its CSS class, then its `style`, then its
        presentation attributes (each overriding the one before)."""
        keys = OptimizableSvg.appearance_keys
        merged = {}
        cls = node.attrib.get("class")
        if cls is not None and "." + cls in defs:
            merged.update(_split_decls(defs["." + cls], keys))
        if "style" in node.attrib:
            merged.update(_split_decls(node.attrib["style"], keys))
        merged.update({k: v for k, v in node.attrib.items() if k in keys})
        out = {}
        for key, value in merged.items():
            if key in ("fill", "stroke"):
                out[key] = OptimizableSvg.parsePaint(value, defs, device)
            elif key in ("fill-opacity", "opacity", "stroke-opacity"):
                out[key] = torch.tensor(OptimizableSvg.parseOpacity(value), device=device)
            elif key == "fill-rule":
                out[key] = value
            elif key == "stroke-width":
                out[key] = torch.tensor(OptimizableSvg.parseLength(value), device=device)
        return out

    def _parse_children(self, node, parent):
        skipped = []
        for child in node:
            tag = OptimizableSvg.remove_namespace(child.tag)
            if tag in OptimizableSvg.recognised_shapes:
                self.parseShape(child, parent)
            elif tag == "defs":
                self.parseDefs(child)
            elif tag == "style":
                self.parseStyle(child)
            elif tag == "g":
                self.parseGroup(child, parent)
            else:
                skipped.append(child)
        if self.verbose:
            self.reportSkippedChildren(node, skipped)

    def parseRoot(self, root):
        if self.verbose:
            print(self.offset_str("Parsing root"))
        self.depth += 1
        self.parseViewport(root)
        self.settings.global_override(["transforms", "translation_mult"], int(np.max(self.canvas)))
        rid = root.attrib.get("id")
        transform = OptimizableSvg.parseTransform(root)
        appearance = OptimizableSvg.parseAppearance(root, self.defs, self.device)
        version = root.attrib.get("version", "<unknown version>")
        if version != "2.0":
            print(self.offset_str("Warning: Version {} is not 2.0, strange things may happen".format(version)))
        self.root = RootNode(rid, transform, appearance, self.settings)
        if self.verbose:
            self.reportSkippedAttribs(root, ["width", "height", "id", "transform", "version", "style"]
                                      + OptimizableSvg.appearance_keys)
        self._parse_children(root, self.root)
        self.depth -= 1

    def parseShape(self, shape, parent):
        tag = OptimizableSvg.remove_namespace(shape.tag)
        if self.verbose:
            print(self.offset_str("Parsing {}#{}".format(tag, shape.attrib.get("id", "<No ID>"))))
        self.depth += 1
        parse = {"path": self.parsePath, "circle": self.parseCircle, "rect": self.parseRect,
                 "ellipse": self.parseEllipse, "polygon": self.parsePolygon}.get(tag)
        if parse is None:
            raise ValueError("Encountered unknown shape type '{}'".format(tag))
        parse(shape, parent)
        self.depth -= 1

    def _attach(self, shape, parent, node_class, payload, known):
        name = shape.attrib.get('id', '')
        transform = OptimizableSvg.parseTransform(shape)
        appearance = OptimizableSvg.parseAppearance(shape, self.defs, self.device)
        parent.children.append(node_class(name, transform, appearance, self.settings, payload))
        if self.verbose:
            self.reportSkippedAttribs(shape, ["id", "transform", "style"] + known + OptimizableSvg.appearance_keys)
            self.reportSkippedChildren(shape, list(shape))

    def parsePath(self, shape, parent):
        name = shape.attrib.get('id', '')
        paths = from_svg_path(shape.attrib['d'])
        for k, p in enumerate(paths):
            p.stroke_width = torch.tensor([0.], device=self.device)
            p.num_control_points = p.num_control_points.to(self.device)
            p.points = p.points.to(self.device)
            p.source_id = name
            p.id = "{}-{}".format(name, k) if len(paths) > 1 else name
        self._attach(shape, parent, PathNode, paths, ["d"])

    @staticmethod
    def _floats(shape, names, optional=()):
        return tuple(float(shape.attrib[n]) if n in shape.attrib or n not in optional else 0.
                     for n in names)

    def parseEllipse(self, shape, parent):
        self._attach(shape, parent, EllipseNode, self._floats(shape, ("cx", "cy", "rx", "ry"), ("cx", "cy")),
                     ["cx", "cy", "rx", "ry"])

    def parseCircle(self, shape, parent):
        self._attach(shape, parent, CircleNode, self._floats(shape, ("cx", "cy", "r"), ("cx", "cy")),
                     ["cx", "cy", "r"])

    def parseRect(self, shape, parent):
        self._attach(shape, parent, RectNode, self._floats(shape, ("x", "y", "width", "height"), ("x", "y")),
                     ["x", "y", "width", "height"])

    def parsePolygon(self, shape, parent):
        pts = []
        for item in shape.attrib['points'].split(" "):
            if len(item) == 0:
                continue
            xy = item.split(",")
            assert len(xy) == 2
            pts.append([float(xy[0]), float(xy[1])])
        self._attach(shape, parent, PolygonNode, torch.tensor(pts, dtype=torch.float, device=self.device),
                     ["points"])

    def parseGroup(self, group, parent):
        gid = group.attrib.get("id", "<No ID>")
        if self.verbose:
            print(self.offset_str("Parsing {}#{}".format(OptimizableSvg.remove_namespace(group.tag), gid)))
        self.depth += 1
        node = GroupNode(gid, self.parseTransform(group), OptimizableSvg.parseAppearance(group, self.defs, self.device),
                         self.settings)
        parent.children.append(node)
        if self.verbose:
            self.reportSkippedAttribs(group, ["id", "transform", "style"] + OptimizableSvg.appearance_keys)
        self._parse_children(group, node)
        self.depth -= 1

    def parseStyle(self, style_node):
        """A text/css stylesheet of class rules -> defs['.name'] = declarations."""
        if self.verbose:
            print(self.offset_str("Parsing {}#{}".format(OptimizableSvg.remove_namespace(style_node.tag),
                                                         style_node.attrib.get("id", "<No ID>"))))
        if style_node.attrib.get("type") != "text/css":
            raise ValueError("Only text/css style recognized, got {}".format(style_node.attrib.get("type")))
        if len(style_node) > 0:
            raise ValueError("Style node should not have children (has {})".format(len(style_node)))
        for selector, decls in _css_rules(style_node.text):
            if len(selector) >= 2 and selector[0] == '.':
                self.defs[selector] = decls
            else:
                raise ValueError("Unrecognized CSS selector {}".format(selector))

    def parseDefs(self, def_node):
        """Linear gradients go into defs by id; radial gradients and anything else are
        skipped; shapes, groups and nested defs are not supported."""
        if self.verbose:
            print(self.offset_str("Parsing {}#{}".format(OptimizableSvg.remove_namespace(def_node.tag),
                                                         def_node.attrib.get("id", "<No ID>"))))
        self.depth += 1
        holder = SvgNode(def_node.attrib.get("id", "<No ID>"), None, {}, self.settings)
        if self.verbose:
            self.reportSkippedAttribs(def_node, ["id"])
        skipped = []
        for child in def_node:
            tag = OptimizableSvg.remove_namespace(child.tag)
            if tag == "linearGradient":
                self.parseGradient(child, holder)
            elif tag in OptimizableSvg.recognised_shapes:
                raise NotImplementedError("Definition/instantiation of shapes not supported")
            elif tag == "defs":
                raise NotImplementedError("Definition within definition not supported")
            elif tag == "g":
                raise NotImplementedError("Groups within definition not supported")
            else:
                skipped.append(child)
            if holder.children:
                node = holder.children.pop()
                self.defs[node.id] = node
        if self.verbose:
            self.reportSkippedChildren(def_node, skipped)
        self.depth -= 1

    def parseGradientStop(self, stop):
        """-> (offset, rgb, opacity); `style` entries override attributes."""
        params = {k: v for k, v in stop.attrib.items() if k in ("id", "offset", "stop-color", "stop-opacity")}
        if "style" in stop.attrib:
            params.update(_split_decls(stop.attrib["style"]))
        offset = OptimizableSvg.parseOpacity(params["offset"])
        color = OptimizableSvg.parse_color(params["stop-color"])
        opacity = OptimizableSvg.parseOpacity(params["stop-opacity"]) if "stop-opacity" in params else 1.
        return offset, color, opacity

    def parseGradient(self, gradient_node, parent):
        gid = gradient_node.attrib.get("id", "<No ID>")
        if self.verbose:
            print(self.offset_str("Parsing {}#{}".format(OptimizableSvg.remove_namespace(gradient_node.tag), gid)))
        self.depth += 1
        children = list(gradient_node)
        has_stops = any(OptimizableSvg.remove_namespace(c.tag) == "stop" for c in children)
        href_key = next((v for k, v in gradient_node.attrib.items()
                         if OptimizableSvg.remove_namespace(k) == "href"), None)
        if not has_stops and href_key is None:
            raise ValueError("Gradient {} has neither stops nor a href link to them".format(gid))
        a = gradient_node.attrib

        def point(xk, yk):
            if xk not in a and yk not in a:
                return None
            return torch.tensor([float(a.get(xk, 0.)), float(a.get(yk, 0.))], dtype=torch.float32,
                                device=self.device)
        begin, end = point("x1", "y1"), point("x2", "y2")
        offsets, stops = None, None
        stop_nodes = [c for c in children if OptimizableSvg.remove_namespace(c.tag) == "stop"]
        if stop_nodes:
            parsed = sorted((self.parseGradientStop(s) for s in stop_nodes), key=lambda t: t[0])
            offsets = torch.tensor([float(o) for o, _, _ in parsed], dtype=torch.float32, device=self.device)
            stops = torch.tensor([[*c.tolist(), float(op)] for _, c, op in parsed], dtype=torch.float32,
                                 device=self.device)
        href = self.defs[href_key.lstrip("#")] if href_key is not None else None
        parent.children.append(GradientNode(gid, self.parseTransform(gradient_node), self.settings,
                                            begin, end, offsets, stops, href))
        self.depth -= 1

    def parseViewport(self, root):
        """Canvas size = ceil(width), ceil(height), else the viewBox's size."""
        def size(v):
            try:
                return int(math.ceil(float(v)))
            except ValueError:
                return int(math.ceil(OptimizableSvg.parseLength(v)))
        if "width" in root.attrib and "height" in root.attrib:
            self.canvas = np.array([size(root.attrib["width"]), size(root.attrib["height"])])
        elif "viewBox" in root.attrib:
            vb = root.attrib["viewBox"].replace(",", " ").split()
            self.canvas = np.array([size(vb[2]), size(vb[3])])
        else:
            raise ValueError("Size information is missing from document definition")

    # ---- writing --------------------------------------------------------------------------
    def write_xml(self):
        """The document with the current parameter values, as pretty-printed text."""
        tree = self.root.write_xml(self)
        return minidom.parseString(etree.tostring(tree, 'utf-8')).toprettyxml(indent="  ")

    def write_defs(self, root):
        if len(self.defs) == 0:
            return
        defnode = etree.SubElement(root, 'defs')
        stylenode = etree.SubElement(root, 'style')
        stylenode.set('type', 'text/css')
        stylenode.text = ""
        # gradients that others link to go first
        pending = dict(self.defs)
        while pending:
            done = []
            for key, value in pending.items():
                if isinstance(value, SvgNode):
                    if value.href is not None and any(value.href is v for v in pending.values()):
                        continue
                    value.write_xml(defnode)
                else:
                    stylenode.text += key + " {" + value + "}\n"
                done.append(key)
            for key in done:
                del pending[key]
