// dvg_winding_fast.h -- decision-exact fast paths of the cubic winding and cubic stroke
// coverage tests.
//
// The winding contribution of a cubic segment (winding_number.h:119-157) is the sum,
// over the roots t of y(t) = pt.y that the double cubic solver (solve.h:29-64) returns,
// of [0 <= t <= 1] [x(t) - pt.x > 0] sign(y'(t)).  The reference evaluates this on
// power-basis coefficients rounded to float, so its polynomial differs from the exact
// Bezier curve by at most a few units in the last place of the coordinates.
//
// This file decides the same sum in float without the cubic solve whenever every one of
// those decisions is certain by a margin that covers that difference:
//   * the curve's y(t) - pt.y is split at its critical points; every breakpoint value
//     (ends and extrema) must be at least `ey` away from zero, so each piece crosses
//     the ray at most once, away from the piece ends (no root near t = 0 or 1, no
//     tangency);
//   * each crossing is bracketed by bisection to 2^-16 of its piece (fine enough that
//     points 1e-4 of the canvas from the curve, where the edge kernel shades both sides
//     of a boundary, are decided), and |y'| must stay above a margin over the bracket,
//     so the sign the reference computes at its root is the crossing direction;
//   * x(t) - pt.x must be certainly positive or certainly negative over the bracket.
// When any test is uncertain it returns false and the caller runs the reference's
// solve.  Host and device compile the same code (no floating-point contraction).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define DVG_WF __host__ __device__ inline
#else
#define DVG_WF inline
#endif

namespace dvg {
namespace wf {

DVG_WF float bern3(float c0, float c1, float c2, float c3, float t) {
    float s = 1.f - t;
    return (s * s * s) * c0 + (3.f * s * s * t) * c1 + (3.f * s * t * t) * c2 + (t * t * t) * c3;
}
// derivative / 3 of the cubic with control values c0..c3, from its differences
DVG_WF float bern2(float d0, float d1, float d2, float t) {
    float s = 1.f - t;
    return (s * s) * d0 + (2.f * s * t) * d1 + (t * t) * d2;
}

DVG_WF float max4(float a, float b, float c, float d) { return fmaxf(fmaxf(a, b), fmaxf(c, d)); }

// N halvings of the bracket [lo, hi] of a root of the cubic with control values c0..c3
// (rising: the cubic goes from negative to positive over the bracket).
template <int N>
DVG_WF void halve(float c0, float c1, float c2, float c3, bool rising, float &lo, float &hi) {
#pragma unroll
    for (int it = 0; it < N; it++) {
        float mid = 0.5f * (lo + hi);
        float fm = bern3(c0, c1, c2, c3, mid);
        if ((fm > 0.f) == rising)
            hi = mid;
        else
            lo = mid;
    }
}

// Over the bracket [lo, hi] of a crossing, each value lies within (derivative bound) x
// (half width) of its value at the midpoint: returns +1 / -1 when y' keeps the crossing's
// sign with margin ed and x(t) - pt.x is certainly positive / negative (margin ex), else 0.
DVG_WF int crossing_side(float p0x, float p1x, float p2x, float p3x, float d0, float d1, float d2, float ptx,
                         bool rising, float lo, float hi, float ddmax, float xdmax, float ed, float ex) {
    const float m = 0.5f * (lo + hi), half = 0.5f * (hi - lo);
    const float yd = 3.f * bern2(d0, d1, d2, m);
    if ((yd > 0.f) != rising || !(fabsf(yd) - 3.f * ddmax * half > ed)) return 0;
    const float xm = bern3(p0x, p1x, p2x, p3x, m) - ptx, xv = xdmax * half + ex;
    if (xm - xv > 0.f) return 1;
    if (xm + xv < 0.f) return -1;
    return 0;
}

// The parts of the test that depend on the segment only, computed once per segment (the
// renderer keeps them in the scene, refreshed whenever the points change).
struct CubicPrep {
    float my, mx;       // largest |y|, |x| of the control points
    float E;            // root accuracy of the reference's solver; NaN: not decidable here
    float ddy, ddx;     // largest control-point difference per coordinate
    float ddmax, xdmax; // |y''| / 3 and |x'| bounds
    float br1, br2;     // interior critical points of y in (0, 1), ascending; -1 when absent;
                        // br1 NaN: y nearly constant, leave the segment to the solver
};

DVG_WF CubicPrep cubic_prep(float p0x, float p0y, float p1x, float p1y, float p2x, float p2y, float p3x, float p3y) {
    CubicPrep q;
    q.my = max4(fabsf(p0y), fabsf(p1y), fabsf(p2y), fabsf(p3y));
    q.mx = max4(fabsf(p0x), fabsf(p1x), fabsf(p2x), fabsf(p3x));
    // The reference's solver works on these float coefficients (winding_number.h:134-138).
    // Its roots are accurate to about E in t: for a near-quadratic cubic (|b| >> |a|) the
    // trigonometric branch loses precision as |b / a| grows; below |a| = 1e-6 it solves
    // the quadratic instead.
    const float ra = -p0y + 3 * p1y - 3 * p2y + p3y, rb = 3 * p0y - 6 * p1y + 3 * p2y, rc = -3 * p0y + 3 * p1y;
    float E = 1e-9f;
    if (!(fabsf(ra) < 1e-6f)) {
        const float ia = 1.f / fabsf(ra);
        E = 2e-7f * (fabsf(rb) * ia + sqrtf(fabsf(rc) * ia) + 1.f);
    }
    q.E = E < 1e-3f ? E : __builtin_nanf("");
    q.ddy = fmaxf(fmaxf(fabsf(p1y - p0y), fabsf(p2y - p1y)), fabsf(p3y - p2y));
    q.ddx = fmaxf(fmaxf(fabsf(p1x - p0x), fabsf(p2x - p1x)), fabsf(p3x - p2x));
    // critical points of y
    const float d0 = p1y - p0y, d1 = p2y - p1y, d2 = p3y - p2y;
    const float qa = d0 - 2.f * d1 + d2, qb = 2.f * (d1 - d0), qc = d0;
    q.br1 = q.br2 = -1.f;
    // y is monotone on [0, 1] when its control differences share a strict sign (y'/3 is
    // their Bernstein combination): no interior critical point to find
    const bool mono = (d0 > 0.f && d1 > 0.f && d2 > 0.f) || (d0 < 0.f && d1 < 0.f && d2 < 0.f);
    if (!mono) {
        float r0 = -1.f, r1 = -1.f;
        const float sc = fabsf(qa) + fabsf(qb) + fabsf(qc);
        bool flat = false;
        if (fabsf(qa) > 1e-6f * sc) {
            float disc = qb * qb - 4.f * qa * qc;
            if (disc >= 0.f) {
                float sq = sqrtf(disc);
                float qq = -0.5f * (qb + (qb >= 0.f ? sq : -sq));
                r0 = qq / qa;
                r1 = qq != 0.f ? qc / qq : -1.f;
            }
        } else if (fabsf(qb) > 1e-6f * sc) {
            r0 = -qc / qb;
        } else {
            flat = true;  // y nearly constant: leave it to the solver
        }
        if (r0 > r1) {
            float tmp = r0;
            r0 = r1;
            r1 = tmp;
        }
        float b[2] = {-1.f, -1.f};
        int nb = 0;
        if (r0 > 0.f && r0 < 1.f) b[nb++] = r0;
        if (r1 > 0.f && r1 < 1.f && r1 != r0) b[nb++] = r1;
        q.br1 = flat ? __builtin_nanf("") : b[0];
        q.br2 = b[1];
    }
    q.ddmax = 2.f * fmaxf(fabsf(d1 - d0), fabsf(d2 - d1));  // |y''| / 3 <= ddmax
    const float dx0 = p1x - p0x, dx1 = p2x - p1x, dx2 = p3x - p2x;
    q.xdmax = 3.f * fmaxf(fmaxf(fabsf(dx0), fabsf(dx1)), fabsf(dx2));  // |x'| bound
    return q;
}

// Returns true and sets `w` when the contribution is decided; `q` is the segment's
// cubic_prep.
DVG_WF bool cubic_winding_fast_prep(const CubicPrep &q, float p0x, float p0y, float p1x, float p1y, float p2x,
                                    float p2y, float p3x, float p3y, float ptx, float pty, int &w) {
    const float M = fmaxf(q.my, fabsf(pty)) + 1.f;
    const float Mx = fmaxf(q.mx, fabsf(ptx)) + 1.f;
    const float E = q.E;
    if (!(E < 1e-3f)) return false;
    // value margins: the reference's coefficients are within ~2e-6 M of the curve's, and
    // its roots within E of the true ones
    const float ey = 1e-5f * M + 3.f * E * q.ddy;
    const float ex = 1e-5f * Mx + 3.f * E * q.ddx;
    const float c0 = p0y - pty, c1 = p1y - pty, c2 = p2y - pty, c3 = p3y - pty;
    // hull above or below the ray
    if ((c0 > ey && c1 > ey && c2 > ey && c3 > ey) || (c0 < -ey && c1 < -ey && c2 < -ey && c3 < -ey)) {
        w = 0;
        return true;
    }
    // hull entirely left of pt: every root has x(t) - pt.x < 0
    if (max4(p0x, p1x, p2x, p3x) < ptx - ex) {
        w = 0;
        return true;
    }
    if (q.br1 != q.br1) return false;  // y nearly constant
    // breakpoints: 0, critical points of y in (0, 1), 1
    float br[4];
    int nb = 0;
    br[nb++] = 0.f;
    if (q.br1 > 0.f) br[nb++] = q.br1;
    if (q.br2 > 0.f) br[nb++] = q.br2;
    br[nb++] = 1.f;
    float fv[4];
    for (int i = 0; i < nb; i++) {
        fv[i] = i == 0 ? c0 : (i == nb - 1 ? c3 : bern3(c0, c1, c2, c3, br[i]));
        if (!(fabsf(fv[i]) > ey)) return false;
    }
    // bounds for the bracket tests
    const float d0 = p1y - p0y, d1 = p2y - p1y, d2 = p3y - p2y;
    const float ddmax = q.ddmax, xdmax = q.xdmax;
    const float ed = 1e-4f * M * 3.f + 1e-6f + 6.f * E * ddmax;  // derivative margin
    int acc = 0;
    for (int i = 0; i + 1 < nb; i++) {
        const float fa = fv[i], fb = fv[i + 1];
        if ((fa > 0.f) == (fb > 0.f)) continue;  // no crossing on this monotone piece
        // Bisection to a bracket of width (b - a) / 65536 at most, in two unrolled runs:
        // after the first 4 halvings a crossing far from pt is usually decided already.
        float lo = br[i], hi = br[i + 1];
        const bool rising = fb > 0.f;
        halve<4>(c0, c1, c2, c3, rising, lo, hi);
        int xs = crossing_side(p0x, p1x, p2x, p3x, d0, d1, d2, ptx, rising, lo, hi, ddmax, xdmax, ed, ex);
        if (xs == 0) {
            halve<12>(c0, c1, c2, c3, rising, lo, hi);
            xs = crossing_side(p0x, p1x, p2x, p3x, d0, d1, d2, ptx, rising, lo, hi, ddmax, xdmax, ed, ex);
        }
        if (xs == 0) return false;
        if (xs > 0) acc += rising ? 1 : -1;
    }
    w = acc;
    return true;
}

// Returns true and sets `w` when the contribution is decided.
DVG_WF bool cubic_winding_fast(float p0x, float p0y, float p1x, float p1y, float p2x, float p2y, float p3x,
                               float p3y, float ptx, float pty, int &w) {
    return cubic_winding_fast_prep(cubic_prep(p0x, p0y, p1x, p1y, p2x, p2y, p3x, p3y), p0x, p0y, p1x, p1y, p2x,
                                   p2y, p3x, p3y, ptx, pty, w);
}

// Stroke coverage of a cubic segment (within_distance.h:119-272) accepts only when one
// of the curve points it evaluates (the ends and the stationary points of the distance)
// is closer than the stroke radius.  This returns true when every point of the curve is
// at least rmax + a rounding margin away from pt, so that test certainly rejects: the
// curve lies within (h^2 / 8) max|B''| of its chords over 8 equal steps of t, with
// |B''| <= 6 max|second differences of the control points| per coordinate.
// Whether the segment ab is farther than `need` (>= 0) from p, compared in squares with
// no division: the end points when p projects outside the segment, else the squared
// cross product against need^2 |b - a|^2.
DVG_WF bool seg_farther(float ax, float ay, float bx, float by, float px, float py, float need) {
    const float dx = bx - ax, dy = by - ay, wx = px - ax, wy = py - ay;
    const float l2 = dx * dx + dy * dy, dt = wx * dx + wy * dy;
    const float n2 = need * need;
    if (!(dt > 0.f) || !(l2 > 0.f)) return wx * wx + wy * wy > n2;
    if (dt >= l2) {
        const float ux = px - bx, uy = py - by;
        return ux * ux + uy * uy > n2;
    }
    const float cr = wx * dy - wy * dx;
    return cr * cr > n2 * l2;
}

// The segment-only part of the stroke test: the polyline points at t = k/8, k = 1..8
// (`poly`, x/y interleaved), the largest control-point coordinate magnitude `mc` and the
// chords' deviation bound `dev`.
DVG_WF void stroke_prep(float p0x, float p0y, float p1x, float p1y, float p2x, float p2y, float p3x, float p3y,
                        float *poly, float &mc, float &dev) {
    mc = fmaxf(max4(fabsf(p0x), fabsf(p1x), fabsf(p2x), fabsf(p3x)), max4(fabsf(p0y), fabsf(p1y), fabsf(p2y), fabsf(p3y)));
    const float sx = fmaxf(fabsf(p0x - 2.f * p1x + p2x), fabsf(p1x - 2.f * p2x + p3x));
    const float sy = fmaxf(fabsf(p0y - 2.f * p1y + p2y), fabsf(p1y - 2.f * p2y + p3y));
    dev = (6.f / 512.f) * sqrtf(sx * sx + sy * sy) * 1.001f;
    for (int k = 1; k <= 8; k++) {
        float t = 0.125f * (float)k;
        poly[2 * (k - 1)] = bern3(p0x, p1x, p2x, p3x, t);
        poly[2 * (k - 1) + 1] = bern3(p0y, p1y, p2y, p3y, t);
    }
}

// The test itself, given the segment's stroke_prep.
DVG_WF bool cubic_certainly_farther_prep(const float *poly, float mc, float dev, float p0x, float p0y, float ptx,
                                         float pty, float rmax) {
    const float M = fmaxf(mc, fmaxf(fabsf(ptx), fabsf(pty))) + 1.f;
    const float margin = 1e-4f * M;
    const float need = rmax + margin + dev;
    float ax = p0x, ay = p0y;
    for (int k = 0; k < 8; k++) {
        const float bx = poly[2 * k], by = poly[2 * k + 1];
        if (!seg_farther(ax, ay, bx, by, ptx, pty, need)) return false;
        ax = bx;
        ay = by;
    }
    return true;
}

DVG_WF bool cubic_certainly_farther(float p0x, float p0y, float p1x, float p1y, float p2x, float p2y, float p3x,
                                    float p3y, float ptx, float pty, float rmax) {
    float poly[16], mc, dev;
    stroke_prep(p0x, p0y, p1x, p1y, p2x, p2y, p3x, p3y, poly, mc, dev);
    return cubic_certainly_farther_prep(poly, mc, dev, p0x, p0y, ptx, pty, rmax);
}

}  // namespace wf
}  // namespace dvg
