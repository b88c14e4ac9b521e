"""save_svg(filename, width, height, shapes, shape_groups, use_gamma=False): write a
scene as SVG text, element for element and attribute for attribute as
pydiffvg.save_svg does (pydiffvg/save_svg.py:13-156), so that files written by either
are identical:
  * <defs> holds one <linearGradient id="shape_{i}_fill|stroke"> per linear-gradient
    paint (radial gradients are not written, and a radial paint cannot be saved), and
    the gamma filter when `use_gamma`;
  * each group is written as ONE element built from its first shape; fills and strokes
    are "rgb(R, G, B)" with R = int(255 * r) plus an opacity attribute; stroke widths are
    doubled back to SVG widths; strokes get round caps and joins;
  * numbers are formatted with str() / format() of the same float, numpy and tensor
    values the reference formats.
"""
import xml.etree.ElementTree as etree
from xml.dom import minidom

from .color import LinearGradient
from .shape import Circle, Ellipse, Path, Polygon, Rect

__all__ = ['save_svg', 'scene_to_svg_string']


def _rgb(c):
    return 'rgb({}, {}, {})'.format(int(255 * c[0]), int(255 * c[1]), int(255 * c[2]))


def _gamma_filter(defs, g):
    f = etree.SubElement(defs, 'filter')
    for k, v in (('id', 'gamma'), ('x', '0'), ('y', '0'), ('width', '100%'), ('height', '100%')):
        f.set(k, v)
    ct = etree.SubElement(f, 'feComponentTransfer')
    ct.set('color-interpolation-filters', 'sRGB')
    for ch in 'RGBA':
        fn = etree.SubElement(ct, 'feFunc' + ch)
        fn.set('type', 'gamma')
        fn.set('amplitude', str(1))
        fn.set('exponent', str(1 / 2.2))
    g.set('style', 'filter:url(#gamma)')


def _linear_def(defs, lg, name):
    el = etree.SubElement(defs, 'linearGradient')
    el.set('id', name)
    for key, v in (('x1', lg.begin[0]), ('y1', lg.begin[1]), ('x2', lg.end[0]), ('y2', lg.end[1])):
        el.set(key, str(v.item()))
    offsets = lg.offsets.data.cpu().numpy()
    for j in range(offsets.shape[0]):
        stop = etree.SubElement(el, 'stop')
        stop.set('offset', str(offsets[j]))
        c = lg.stop_colors[j, :]
        stop.set('stop-color', _rgb(c))
        stop.set('stop-opacity', '{}'.format(c[3]))


def _path_data(shape):
    ncp = shape.num_control_points.data.cpu().numpy()
    pts = shape.points.data.cpu().numpy()
    n = pts.shape[0]
    out = ['M {} {}'.format(pts[0, 0], pts[0, 1])]
    k = 1
    for c in ncp:
        if c == 0:
            q = k % n
            out.append('L {} {}'.format(pts[q, 0], pts[q, 1]))
            k += 1
        elif c == 1:
            q = (k + 1) % n
            out.append('Q {} {} {} {}'.format(pts[k, 0], pts[k, 1], pts[q, 0], pts[q, 1]))
            k += 2
        elif c == 2:
            q = (k + 2) % n
            out.append('C {} {} {} {} {} {}'.format(pts[k, 0], pts[k, 1], pts[k + 1, 0], pts[k + 1, 1],
                                                    pts[q, 0], pts[q, 1]))
            k += 3
    return ' '.join(out)


def _shape_element(g, shape):
    if isinstance(shape, Circle):
        el = etree.SubElement(g, 'circle')
        el.set('r', str(shape.radius.item()))
        el.set('cx', str(shape.center[0].item()))
        el.set('cy', str(shape.center[1].item()))
    elif isinstance(shape, Polygon):
        el = etree.SubElement(g, 'polygon')
        pts = shape.points.data.cpu().numpy()
        el.set('points', ' '.join('{} {}'.format(pts[j, 0], pts[j, 1]) for j in range(pts.shape[0])))
    elif isinstance(shape, Path):
        el = etree.SubElement(g, 'path')
        el.set('d', _path_data(shape))
    elif isinstance(shape, Rect):
        el = etree.SubElement(g, 'rect')
        x0, y0 = shape.p_min[0].item(), shape.p_min[1].item()
        el.set('x', str(x0))
        el.set('y', str(y0))
        el.set('width', str(shape.p_max[0].item() - x0))
        el.set('height', str(shape.p_max[1].item() - y0))
    elif isinstance(shape, Ellipse):
        el = etree.SubElement(g, 'ellipse')
        el.set('cx', str(shape.center[0].item()))
        el.set('cy', str(shape.center[1].item()))
        el.set('rx', str(shape.radius[0].item()))
        el.set('ry', str(shape.radius[1].item()))
    else:
        raise TypeError('save_svg: unsupported shape %r' % type(shape).__name__)
    return el


def _paint(el, paint, attr, opacity_attr, ref):
    if isinstance(paint, LinearGradient):
        el.set(attr, 'url(#{})'.format(ref))
    else:
        c = paint.data.cpu().numpy()
        el.set(attr, _rgb(c))
        el.set(opacity_attr, str(c[3]))


def scene_to_svg_string(width, height, shapes, shape_groups, use_gamma=False):
    root = etree.Element('svg')
    for k, v in (('version', '1.1'), ('xmlns', 'http://www.w3.org/2000/svg'), ('width', str(width)),
                 ('height', str(height))):
        root.set(k, v)
    defs = etree.SubElement(root, 'defs')
    g = etree.SubElement(root, 'g')
    if use_gamma:
        _gamma_filter(defs, g)
    for i, grp in enumerate(shape_groups):
        for paint, kind in ((grp.fill_color, 'fill'), (grp.stroke_color, 'stroke')):
            if isinstance(paint, LinearGradient):
                _linear_def(defs, paint, 'shape_{}_{}'.format(i, kind))
    for i, grp in enumerate(shape_groups):
        shape = shapes[grp.shape_ids[0]]
        el = _shape_element(g, shape)
        el.set('stroke-width', str(2 * shape.stroke_width.data.cpu().item()))
        if grp.fill_color is not None:
            _paint(el, grp.fill_color, 'fill', 'opacity', 'shape_{}_fill'.format(i))
        else:
            el.set('fill', 'none')
        if grp.stroke_color is not None:
            _paint(el, grp.stroke_color, 'stroke', 'stroke-opacity', 'shape_{}_stroke'.format(i))
            el.set('stroke-linecap', 'round')
            el.set('stroke-linejoin', 'round')
    return minidom.parseString(etree.tostring(root, 'utf-8')).toprettyxml(indent='  ')


def save_svg(filename, width, height, shapes, shape_groups, use_gamma=False):
    """Write the scene to `filename` (pydiffvg/save_svg.py:13-156)."""
    text = scene_to_svg_string(width, height, shapes, shape_groups, use_gamma)
    with open(filename, 'w') as f:
        f.write(text)
