"""SVG path data (the `d` attribute) to segments, and segments to diffvg paths.

pydiffvg delegates path parsing to the third-party package svgpathtools (unpinned in
setup.py:94; pydiffvg/shape.py:63 calls ``svgpathtools.parse_path``), which is not
installed here.  This module implements the same published behaviour from the SVG 1.1
path grammar (SVG 1.1 section 8.3) and the endpoint-to-centre arc conversion of SVG 1.1
appendix F.6.5, with points as Python complex numbers like svgpathtools:

  * commands M L H V C S Q T A Z in absolute and relative form; numbers after M are
    implicit line-tos; S and T reflect the previous segment's control point only when
    the previous command was C/S (Q/T); Z adds a closing line only when the current
    point differs from the subpath start;
  * zero-radius arcs become lines; radii too small to reach the end point are scaled
    up (F.6.6);
  * a path splits into continuous subpaths wherever a segment does not start exactly
    where the previous one ended.

`segments_to_paths` then follows pydiffvg/shape.py:63-172: near-closed subpaths snap
shut, filled open subpaths get a closing line, and each arc becomes cubics of at most a
quarter turn using the construction of pydiffvg/shape.py:107-157 (including its
placement of rotated-ellipse points, which scales after rotating).
"""
import math
import re

import torch

_COMMANDS = set('MmZzLlHhVvCcSsQqTtAa')
_TOKEN = re.compile(r'([MmZzLlHhVvCcSsQqTtAa])|([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)')


class Line:
    __slots__ = ('start', 'end')

    def __init__(self, start, end):
        self.start, self.end = start, end

    def length(self):
        return abs(self.end - self.start)


class QuadraticBezier:
    __slots__ = ('start', 'control', 'end')

    def __init__(self, start, control, end):
        self.start, self.control, self.end = start, control, end


class CubicBezier:
    __slots__ = ('start', 'control1', 'control2', 'end')

    def __init__(self, start, control1, control2, end):
        self.start, self.control1, self.control2, self.end = start, control1, control2, end


class Arc:
    """Elliptical arc in endpoint form, with its centre form derived at construction:
    `center` (complex), `radius` (complex rx + i ry, possibly scaled up), `rotation`
    (degrees), `theta` (start angle, degrees) and `delta` (signed sweep, degrees)."""

    def __init__(self, start, radius, rotation, large_arc, sweep, end):
        self.start, self.end = start, end
        self.rotation = rotation
        self.large_arc, self.sweep = bool(large_arc), bool(sweep)
        rx, ry = abs(radius.real), abs(radius.imag)
        phi = math.radians(rotation)
        cphi, sphi = math.cos(phi), math.sin(phi)
        # F.6.5.1: the half chord in the ellipse's axis frame
        hx, hy = (start.real - end.real) / 2.0, (start.imag - end.imag) / 2.0
        x1 = cphi * hx + sphi * hy
        y1 = -sphi * hx + cphi * hy
        # F.6.6.2-3: scale radii that cannot span the chord
        lam = (x1 * x1) / (rx * rx) + (y1 * y1) / (ry * ry)
        if lam > 1.0:
            s = math.sqrt(lam)
            rx, ry = rx * s, ry * s
        self.radius = complex(rx, ry)
        # F.6.5.2: the centre in the axis frame
        num = rx * rx * ry * ry - rx * rx * y1 * y1 - ry * ry * x1 * x1
        den = rx * rx * y1 * y1 + ry * ry * x1 * x1
        k = math.sqrt(max(num / den, 0.0)) if den > 0 else 0.0
        if self.large_arc == self.sweep:
            k = -k
        cx1, cy1 = k * rx * y1 / ry, -k * ry * x1 / rx
        # F.6.5.3: back to the user frame
        mx, my = (start.real + end.real) / 2.0, (start.imag + end.imag) / 2.0
        self.center = complex(cphi * cx1 - sphi * cy1 + mx, sphi * cx1 + cphi * cy1 + my)
        # F.6.5.5-6: start angle and sweep on the unit circle
        ux, uy = (x1 - cx1) / rx, (y1 - cy1) / ry
        vx, vy = (-x1 - cx1) / rx, (-y1 - cy1) / ry
        theta = math.degrees(math.atan2(uy, ux))
        delta = math.degrees(math.atan2(ux * vy - uy * vx, ux * vx + uy * vy))
        if not self.sweep and delta > 0:
            delta -= 360.0
        elif self.sweep and delta < 0:
            delta += 360.0
        self.theta, self.delta = theta, delta


class _Tokens:
    def __init__(self, d):
        self.toks = [m.group(0) for m in _TOKEN.finditer(d)]
        self.i = 0

    def more(self):
        return self.i < len(self.toks)

    def peek_command(self):
        return self.more() and self.toks[self.i] in _COMMANDS

    def next(self):
        if not self.more():
            raise ValueError('SVG path data ends inside a command')
        t = self.toks[self.i]
        self.i += 1
        return t

    def number(self):
        t = self.next()
        if t in _COMMANDS:
            raise ValueError('SVG path data: expected a number, found %r' % t)
        return float(t)

    def flag(self):
        # arc flags may be written without separators ("a1 1 0 011 1")
        t = self.next()
        if t in _COMMANDS or t[0] not in '01':
            raise ValueError('SVG path data: bad arc flag %r' % t)
        if len(t) > 1:
            self.toks[self.i - 1] = t[1:]
            self.i -= 1
        return t[0] == '1'


class SegmentList(list):
    """The parsed segments; `closed` is set once a Z command has been seen."""
    closed = False


def parse_path(d):
    """Path data string -> SegmentList of Line / QuadraticBezier / CubicBezier / Arc."""
    tok = _Tokens(d)
    segs = SegmentList()
    pos = 0j
    start = None
    cmd = None
    while tok.more():
        if tok.peek_command():
            prev = cmd
            raw = tok.next()
            absolute = raw.isupper()
            cmd = raw.upper()
        else:
            if cmd is None:
                raise ValueError('SVG path data must begin with a command: %r' % d)
            prev = cmd

        def point():
            p = complex(tok.number(), tok.number())
            return p if absolute else p + pos

        if cmd == 'M':
            pos = point()
            start = pos
            cmd = 'L'  # further pairs are line-tos
        elif cmd == 'Z':
            if pos != start:
                segs.append(Line(pos, start))
            segs.closed = True
            pos = start
            cmd = None
        elif cmd == 'L':
            end = point()
            segs.append(Line(pos, end))
            pos = end
        elif cmd == 'H':
            x = tok.number()
            end = complex(x if absolute else x + pos.real, pos.imag)
            segs.append(Line(pos, end))
            pos = end
        elif cmd == 'V':
            y = tok.number()
            end = complex(pos.real, y if absolute else y + pos.imag)
            segs.append(Line(pos, end))
            pos = end
        elif cmd == 'C':
            c1, c2, end = point(), point(), point()
            segs.append(CubicBezier(pos, c1, c2, end))
            pos = end
        elif cmd == 'S':
            c1 = pos + pos - segs[-1].control2 if prev in ('C', 'S') else pos
            c2, end = point(), point()
            segs.append(CubicBezier(pos, c1, c2, end))
            pos = end
        elif cmd == 'Q':
            c, end = point(), point()
            segs.append(QuadraticBezier(pos, c, end))
            pos = end
        elif cmd == 'T':
            c = pos + pos - segs[-1].control if prev in ('Q', 'T') else pos
            end = point()
            segs.append(QuadraticBezier(pos, c, end))
            pos = end
        elif cmd == 'A':
            radius = complex(tok.number(), tok.number())
            rotation = tok.number()
            large_arc, sweep = tok.flag(), tok.flag()
            end = point()
            if end == pos:
                pass  # F.6.2: an arc that ends where it starts is omitted
            elif radius.real == 0 or radius.imag == 0:
                segs.append(Line(pos, end))
            else:
                segs.append(Arc(pos, radius, rotation, large_arc, sweep, end))
            pos = end
    return segs


def continuous_subpaths(segs):
    """Split where a segment does not start exactly at the previous segment's end."""
    out, cur = [], []
    for s in segs:
        if cur and cur[-1].end != s.start:
            out.append(cur)
            cur = []
        cur.append(s)
    if cur:
        out.append(cur)
    return out


def _arc_to_cubics(e, ncp, points):
    """Cubic pieces of at most a quarter turn for arc `e` (pydiffvg/shape.py:107-157):
    appends their control counts and points, leaving out the arc's final end point."""
    start = math.radians(e.theta)
    stop = math.radians(e.theta + e.delta)
    sign = -1.0 if stop < start else 1.0
    eps = 1e-5
    rx, ry = e.radius.real, e.radius.imag
    cx, cy = e.center.real, e.center.imag
    rot = math.radians(e.rotation)
    cr, sr = math.cos(rot), math.sin(rot)
    while sign * (stop - start) > eps:
        step = stop - start
        step = min(step, 0.5 * math.pi) if step > 0 else max(step, -0.5 * math.pi)
        half = step / 2.0
        ca, sa = math.cos(half), math.sin(half)
        mid = start + half
        cm, sm = math.cos(mid), math.sin(mid)
        # control points of the unit-circle piece from start to start + step
        lam = (4.0 - ca) / 3.0
        mu = sa + (ca - lam) * (1.0 / math.tan(half))
        ncp.append(2)
        for ux, uy in ((lam * cm + mu * sm, lam * sm - mu * cm), (lam * cm - mu * sm, lam * sm + mu * cm)):
            points.append((cx + rx * (ux * cr - uy * sr), cy + ry * (ux * sr + uy * cr)))
        if sign * (stop - (start + step)) > eps:
            a = rot + start + step
            points.append((cx + rx * math.cos(a), cy + ry * math.sin(a)))
        start += step


def segments_to_paths(segs, shape_to_canvas, force_close, path_cls):
    """Subpaths of `segs` -> list of `path_cls(num_control_points, points, is_closed)`,
    points mapped through `shape_to_canvas` (pydiffvg/shape.py:63-172)."""
    out = []
    for sub in continuous_subpaths(segs):
        sub = list(sub)
        beg = sub[0].start
        if sub[-1].end == beg:
            if len(sub) > 1 and isinstance(sub[-1], Line) and sub[-1].length() < 1e-5:
                sub.pop()
                sub[-1].end = beg
        elif abs(sub[-1].end - beg) < 1e-5:
            sub[-1].end = beg
        elif force_close:
            sub.append(Line(sub[-1].end, beg))
        closed = sub[-1].end == sub[0].start
        ncp, points = [], []
        last = len(sub) - 1
        for i, e in enumerate(sub):
            if i == 0:
                points.append((e.start.real, e.start.imag))
            elif (e.start.real, e.start.imag) != points[-1]:
                raise ValueError('SVG path: a segment does not start where the previous one ends')
            if isinstance(e, Line):
                ncp.append(0)
            elif isinstance(e, QuadraticBezier):
                ncp.append(1)
                points.append((e.control.real, e.control.imag))
            elif isinstance(e, CubicBezier):
                ncp.append(2)
                points.append((e.control1.real, e.control1.imag))
                points.append((e.control2.real, e.control2.imag))
            else:
                _arc_to_cubics(e, ncp, points)
            if i != last or not closed:
                points.append((e.end.real, e.end.imag))
        pts = torch.tensor(points, dtype=torch.float32)
        h = torch.cat((pts, torch.ones([pts.shape[0], 1])), dim=1) @ torch.transpose(shape_to_canvas, 0, 1)
        pts = (h / h[:, 2:3])[:, :2].contiguous()
        out.append(path_cls(torch.tensor(ncp), pts, closed))
    return out
