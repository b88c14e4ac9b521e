"""SVG file -> (canvas_width, canvas_height, shapes, shape_groups).

Same entry point and result as pydiffvg.svg_to_scene (pydiffvg/parse_svg.py:575-586):
element and attribute handling follows parse_svg.py:16-573, path data goes through
diffvg_amd.svg_path (the svgpathtools behaviour it relies on), and CSS class rules are
read by a small parser of `.name { declarations }` blocks (the reference uses cssutils,
parse_svg.py:263-271, absent here).  Named colours are resolved by matplotlib, as in the
reference (parse_svg.py:79-82).

The reference's observable behaviour is kept, including what looks accidental, so that a
scene loads into the same tensors:
  * a fill colour inherited from the document default or from a <g fill=...> is ONE
    tensor shared by every shape that inherits it, and each shape writes its opacity
    into it (parse_svg.py:313-314, 494-495, 525);
  * gradient fills referenced with url(#id) are shared objects; a shape that has a
    `style` attribute maps the shared linear gradient's begin/end through its transform
    in place (parse_svg.py:364-377);
  * <radialGradient> yields RadialGradient(center=(0,0), radius=(0,0)) with its stops:
    the parsed centre and focus are not passed on (parse_svg.py:229-262);
  * <rect> ignores x and y and spans (0,0)-(width,height) (parse_svg.py:518-532);
  * <ellipse> elements are skipped (is_shape, parse_svg.py:384-385);
  * stroke widths are half the SVG stroke-width, default 0.5 (parse_svg.py:300, 331-335).
"""
import re
import warnings
import xml.etree.ElementTree as etree

import numpy as np
import torch

from .color import LinearGradient, RadialGradient
from .shape import Circle, Polygon, Rect, ShapeGroup, from_svg_path

__all__ = ['svg_to_scene', 'parse_scene', 'parse_transform', 'parse_color', 'parse_style']

_SHAPE_TAGS = ('path', 'polygon', 'line', 'circle', 'rect')


def _local(tag):
    """Element or attribute name without its '{namespace}' prefix."""
    return re.sub('{.*}', '', tag)


def _hex_rgb(s):
    s = s.lstrip('#')
    if len(s) == 3:
        s = ''.join(ch * 2 for ch in s)
    return torch.tensor([int(s[k:k + 2], 16) / 255.0 for k in (0, 2, 4)])


def parse_color(s, defs):
    """Colour value -> RGBA tensor, a gradient object from `defs`, or None
    (parse_svg.py:55-81).  Values already parsed (tensors, and gradients named in a
    style declaration, which the reference fails on) pass through unchanged."""
    if s is None or not isinstance(s, str):
        return s
    s = s.lstrip(' ')
    if s[0] == '#':
        c = torch.tensor([0.0, 0.0, 0.0, 1.0])
        c[:3] = _hex_rgb(s)
        return c
    if s[:3] == 'url':
        return defs[s[4:-1].lstrip('#')]
    if s == 'none':
        return None
    if s[:4] == 'rgb(':
        r, g, b = (int(v) / 255.0 for v in s[4:-1].split(',')[:3])
        return torch.tensor([r, g, b, 1.0])
    import matplotlib.colors
    try:
        return torch.tensor(matplotlib.colors.to_rgba(s))
    except ValueError:
        warnings.warn('Unknown color command ' + s)
        return torch.tensor([0.0, 0.0, 0.0, 1.0])


def parse_style(s, defs):
    """'key: value; ...' -> dict; fill and stroke values become colours
    (parse_svg.py:22-34)."""
    out = {}
    for decl in s.split(';'):
        kv = decl.split(':')
        if len(kv) != 2:
            continue
        key, value = kv[0].strip(), kv[1].strip()
        out[key] = parse_color(value, defs) if key in ('fill', 'stroke') else value
    return out


def _svg_int(s):
    """Integer part of a length, dropping unit letters (parse_svg.py:49-53)."""
    return int(float(''.join(ch for ch in s if not ch.isalpha())))


def _one_transform(kind, args):
    m = np.identity(3)
    if 'matrix' in kind:
        m[0:2, 0:3] = np.array([args[0:6:2], args[1:6:2]])
    elif 'translate' in kind:
        m[0, 2] = args[0]
        if len(args) > 1:
            m[1, 2] = args[1]
    elif 'scale' in kind:
        m[0, 0] = args[0]
        m[1, 1] = args[1] if len(args) > 1 else args[0]
    elif 'rotate' in kind:
        a = args[0] * np.pi / 180.0
        ox, oy = (args[1], args[2]) if len(args) == 3 else (0, 0)
        to = np.identity(3)
        to[0:2, 2] = (ox, oy)
        rot = np.identity(3)
        rot[0:2, 0:2] = [[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]]
        back = np.identity(3)
        back[0:2, 2] = (-ox, -oy)
        m = to.dot(rot).dot(back)
    elif 'skewX' in kind:
        m[0, 1] = np.tan(args[0] * np.pi / 180.0)
    elif 'skewY' in kind:
        m[1, 0] = np.tan(args[0] * np.pi / 180.0)
    else:
        warnings.warn('Unknown SVG transform type: {0}'.format(kind))
    return m


def parse_transform(s):
    """SVG transform list -> 3x3 float32 tensor, composed left to right
    (parse_svg.py:84-146); an empty value gives the identity."""
    if not s:
        return np.identity(3)
    if not isinstance(s, str):
        raise TypeError('Must provide a string to parse')
    total = np.identity(3)
    for part in s.split(')')[:-1]:
        kind, vals = part.split('(')
        args = [float(v) for v in vals.replace(',', ' ').split(' ') if v]
        total = total.dot(_one_transform(kind, args))
    return torch.from_numpy(total).type(torch.float32)


def _stops(node, defs, offsets, colors):
    for child in node:
        if _local(child.tag) != 'stop':
            continue
        off = float(child.attrib['offset'])
        rgba = [0.0, 0.0, 0.0, 1.0]
        for src in (child.attrib, parse_style(child.attrib['style'], defs) if 'style' in child.attrib else {}):
            if 'stop-color' in src:
                c = parse_color(src['stop-color'], defs)
                rgba[:3] = [c[0], c[1], c[2]]
            if 'stop-opacity' in src:
                rgba[3] = float(src['stop-opacity'])
        offsets.append(off)
        colors.append(rgba)
    offsets = torch.tensor(offsets) if isinstance(offsets, list) else offsets
    colors = torch.tensor(colors) if isinstance(colors, list) else colors
    return offsets, colors


def _map_point(m, p):
    h = m @ torch.cat((p, torch.ones([1])))
    return (h / h[2])[:2]


class _SceneParser:
    def __init__(self):
        self.defs = {}
        self.shapes = []
        self.groups = []

    # ---- definitions -------------------------------------------------------------
    def _inherit(self, node):
        for key in node.attrib:
            if _local(key) == 'href':
                return self.defs[node.attrib[key].lstrip('#')]
        return None

    def linear_gradient(self, node, transform):
        """parse_svg.py:148-198"""
        begin, end = torch.tensor([0.0, 0.0]), torch.tensor([0.0, 0.0])
        offsets, colors = [], []
        parent = self._inherit(node)
        if parent is not None:
            begin, end, offsets, colors = parent.begin, parent.end, parent.offsets, parent.stop_colors
        for key in node.attrib:
            name = _local(key)
            if name in ('x1', 'y1', 'x2', 'y2'):
                point = begin if name[1] == '1' else end
                point[0 if name[0] == 'x' else 1] = float(node.attrib[name])
            elif name == 'gradientTransform':
                transform = transform @ parse_transform(node.attrib['gradientTransform'])
        begin, end = _map_point(transform, begin), _map_point(transform, end)
        offsets, colors = _stops(node, self.defs, offsets, colors)
        return LinearGradient(begin, end, offsets, colors)

    def radial_gradient(self, node, transform):
        """parse_svg.py:200-261: the stops are read; centre and radius stay (0, 0)."""
        begin, end = torch.tensor([0.0, 0.0]), torch.tensor([0.0, 0.0])
        offsets, colors = [], []
        parent = self._inherit(node)
        if parent is not None:
            begin, end, offsets, colors = parent.begin, parent.end, parent.offsets, parent.stop_colors
        for key in node.attrib:
            if _local(key) == 'gradientTransform':
                transform = transform @ parse_transform(node.attrib['gradientTransform'])
        offsets, colors = _stops(node, self.defs, offsets, colors)
        return RadialGradient(begin, end, offsets, colors)

    def stylesheet(self, node):
        """CSS class rules -> defs[class] = declarations (parse_svg.py:263-271)."""
        text = re.sub(r'/\*.*?\*/', '', node.text or '', flags=re.S)
        for m in re.finditer(r'([^{}]+)\{([^}]*)\}', text):
            selector = ' '.join(m.group(1).split())
            if len(selector) >= 2 and selector[0] == '.':
                decls = ';\n'.join(d.strip() for d in m.group(2).split(';') if d.strip())
                self.defs[selector[1:]] = parse_style(decls, self.defs)

    def definition(self, child, tag, transform):
        if tag == 'linearGradient':
            if 'id' in child.attrib:
                self.defs[child.attrib['id']] = self.linear_gradient(child, transform)
        elif tag == 'radialGradient':
            if 'id' in child.attrib:
                self.defs[child.attrib['id']] = self.radial_gradient(child, transform)
        elif tag == 'style':
            self.stylesheet(child)

    # ---- presentation attributes ---------------------------------------------------
    def attributes(self, node, transform, fill):
        """parse_svg.py:273-382 -> (transform, fill, stroke, stroke_width, even_odd)."""
        attrs = {}
        if 'class' in node.attrib:
            attrs.update(self.defs[node.attrib['class']])
        attrs.update(node.attrib)
        name = node.attrib.get('id', '')
        stroke, width, even_odd = None, torch.tensor(0.5), False

        def fill_rule(v, cur):
            if v == 'evenodd':
                return True
            if v == 'nonzero':
                return False
            warnings.warn('Unknown fill-rule: {}'.format(v))
            return cur

        def half_width(v):
            return torch.tensor(float(v[:-2] if v[-2:] == 'px' else v) / 2.0)

        if 'transform' in attrs:
            transform = transform @ parse_transform(attrs['transform'])
        if 'fill' in attrs:
            fill = parse_color(attrs['fill'], self.defs)
        opacity = 1.0
        for key in ('fill-opacity', 'opacity'):
            if key in attrs:
                opacity *= float(attrs[key])
        if isinstance(fill, torch.Tensor):
            fill[3] = opacity
        if 'fill-rule' in attrs:
            even_odd = fill_rule(attrs['fill-rule'], even_odd)
        if 'stroke' in attrs:
            stroke = parse_color(attrs['stroke'], self.defs)
        if 'stroke-width' in attrs:
            width = half_width(attrs['stroke-width'])
        if 'stroke-opacity' in attrs:
            stroke[3] = torch.tensor(float(attrs['stroke-opacity']))
        if 'style' in attrs:
            st = parse_style(attrs['style'], self.defs)
            if 'fill' in st:
                fill = parse_color(st['fill'], self.defs)
            opacity = 1.0
            for key in ('fill-opacity', 'opacity'):
                if key in st:
                    opacity *= float(st[key])
            if 'fill-rule' in st:
                even_odd = fill_rule(st['fill-rule'], even_odd)
            if isinstance(fill, torch.Tensor):
                fill[3] = opacity
            if 'stroke' in st and st['stroke'] != 'none':
                stroke = parse_color(st['stroke'], self.defs)
                if isinstance(stroke, torch.Tensor):
                    if 'stroke-opacity' in st:
                        stroke[3] = float(st['stroke-opacity'])
                    if 'opacity' in st:
                        stroke[3] *= float(st['opacity'])
                if 'stroke-width' in st:
                    width = half_width(st['stroke-width'])
            for paint in (fill, stroke):
                if isinstance(paint, LinearGradient):
                    paint.begin = _map_point(transform, paint.begin)
                    paint.end = _map_point(transform, paint.end)
            if 'filter' in st:
                print('*** WARNING ***: Ignoring filter for path with id "{}"'.format(name))
        return transform, fill, stroke, width, even_odd

    # ---- shapes ------------------------------------------------------------------
    def _group(self, ids, fill, stroke, even_odd, transform=None, name=None):
        kw = dict(shape_ids=ids, fill_color=fill, stroke_color=stroke, use_even_odd_rule=even_odd)
        if transform is not None:
            kw['shape_to_canvas'] = transform
        if name is not None:
            kw['id'] = name
        self.groups.append(ShapeGroup(**kw))

    def _single(self, shape, width, fill, stroke, even_odd, transform, name=None):
        shape.stroke_width = width
        ids = torch.tensor([len(self.shapes)])
        self.shapes.append(shape)
        self._group(ids, fill, stroke, even_odd, transform, name)

    def shape(self, node, tag, transform, fill):
        """parse_svg.py:387-532"""
        transform, fill, stroke, width, even_odd = self.attributes(node, transform, fill)
        a = node.attrib
        name = a.get('id', '')
        if tag == 'path':
            paths = from_svg_path(a['d'], transform, fill is not None)
            for k, p in enumerate(paths):
                p.stroke_width = width
                p.source_id = name
                p.id = '{}-{}'.format(name, k) if len(paths) > 1 else name
            first = len(self.shapes)
            self.shapes = self.shapes + paths
            self._group(torch.tensor(list(range(first, len(self.shapes)))), fill, stroke, even_odd, name=name)
        elif tag == 'polygon':
            coords = [[float(v) for v in re.split(',| ', tok)] for tok in a['points'].strip().split(' ') if tok]
            pts = torch.tensor(coords, dtype=torch.float32).view(-1, 2)
            self._single(Polygon(pts, fill is not None), width, fill, stroke, even_odd, transform, name)
        elif tag == 'line':
            pts = torch.stack((torch.tensor([float(a['x1']), float(a['y1'])]),
                               torch.tensor([float(a['x2']), float(a['y2'])])))
            self._single(Polygon(pts, False), width, fill, stroke, even_odd, transform)
        elif tag == 'circle':
            c = Circle(radius=torch.tensor(float(a['r'])), center=torch.tensor([float(a['cx']), float(a['cy'])]))
            self._single(c, width, fill, stroke, even_odd, transform)
        elif tag == 'rect':
            w, h = float(a['width']), float(a['height'])
            r = Rect(p_min=torch.tensor([0.0, 0.0]), p_max=torch.tensor([0.0 + w, 0.0 + h]))
            self._single(r, width, fill, stroke, even_odd, transform)

    def children(self, node, transform, fill, top=False):
        for child in node:
            tag = _local(child.tag)
            if top and tag == 'defs':
                for d in child:
                    self.definition(d, _local(d.tag), transform)
            elif top and tag in ('style', 'linearGradient', 'radialGradient'):
                self.definition(child, tag, transform)
            elif tag in _SHAPE_TAGS:
                self.shape(child, tag, transform, fill)
            elif tag == 'g':
                self.group(child, transform, fill)

    def group(self, node, transform, fill):
        """parse_svg.py:534-546"""
        if 'transform' in node.attrib:
            transform = transform @ parse_transform(node.attrib['transform'])
        if 'fill' in node.attrib:
            fill = parse_color(node.attrib['fill'], self.defs)
        self.children(node, transform, fill)


def parse_scene(node):
    """<svg> element -> (canvas_width, canvas_height, shapes, shape_groups)
    (parse_svg.py:548-573)."""
    width = height = -1
    if 'viewBox' in node.attrib:
        box = node.attrib['viewBox'].split()
        width, height = _svg_int(box[2]), _svg_int(box[3])
    else:
        if 'width' in node.attrib:
            width = _svg_int(node.attrib['width'])
        else:
            print('Warning: Can\'t find canvas width.')
        if 'height' in node.attrib:
            height = _svg_int(node.attrib['height'])
        else:
            print('Warning: Can\'t find canvas height.')
    p = _SceneParser()
    p.children(node, torch.eye(3), torch.tensor([0.0, 0.0, 0.0, 1.0]), top=True)
    return width, height, p.shapes, p.groups


def svg_to_scene(filename):
    """Load an SVG file as (canvas_width, canvas_height, shapes, shape_groups)
    (parse_svg.py:575-586)."""
    return parse_scene(etree.parse(filename).getroot())
