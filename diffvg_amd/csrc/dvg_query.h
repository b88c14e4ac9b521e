// dvg_query.h -- per-sample geometry queries on the device scene.
//
// Each query takes the same decisions as the reference's BVH-traversing version:
//   is_inside        diffvg.cpp:33-87 + winding_number.h:9-202
//   within_distance  within_distance.h:8-446
//   compute_distance compute_distance.h:18-439 (closest point) and 442-949 (adjoint)
// The BVH levels below the top level are small (one group usually holds one shape; a
// path holds a few dozen segments), so instead of a per-lane stack the kernels walk
// the leaves in the reference's visit order (precomputed per path and per group) and
// apply the reference's reachability test to each leaf.  Ancestor boxes contain their
// children and carry the maximum radius, so a leaf is reached exactly when its own
// inflated box passes the test; with the leaves in visit order, order-dependent
// results (distance ties, the use_distance_approx early return) match too.
#pragma once

#include "dvg_math.h"
#include "dvg_winding_fast.h"
#include "dvg_scene.h"

namespace dvg {

struct EdgeQuery {
    int group, shape;
    bool hit;
};
struct PathInfo {
    int base_point_id, point_id;
    float t_root;
};

DVG_DEV V2 ld_pt(const DevScene &S, int i) {
    float2 p = S.points[i];
    return V2{p.x, p.y};
}
// Segment points; returns the number of control points k (k + 2 points).
DVG_DEV int seg_load(const DevScene &S, int seg, V2 *p, int4 &idx) {
    idx = S.seg_idx[seg];
    int k = S.seg_kind[seg];
    p[0] = ld_pt(S, idx.x);
    p[1] = ld_pt(S, idx.y);
    if (k >= 1) p[2] = ld_pt(S, idx.z);
    if (k >= 2) p[3] = ld_pt(S, idx.w);
    return k;
}
DVG_DEV V2 qeval(V2 p0, V2 p1, V2 p2, float t) {
    float tt = 1 - t;
    return (tt * tt) * p0 + (2 * tt * t) * p1 + (t * t) * p2;
}
DVG_DEV V2 ceval(V2 p0, V2 p1, V2 p2, V2 p3, float t) {
    float tt = 1 - t;
    return (tt * tt * tt) * p0 + (3 * tt * tt * t) * p1 + (3 * tt * t * t) * p2 + (t * t * t) * p3;
}

// Winding contribution of one segment (winding_number.h:58-167)
// `prep`: the segment's wf::CubicPrep as three float4 (cubic segments only)
DVG_DEV int segment_winding(int k, const V2 *p, V2 pt, const float4 *prep) {
    int w = 0;
    if (k == 0) {
        V2 p0 = p[0], p1 = p[1];
        if (p1.y != p0.y) {
            float t = (pt.y - p0.y) / (p1.y - p0.y);
            if (t >= 0 && t <= 1) {
                float tp = p0.x - pt.x + t * (p1.x - p0.x);
                if (tp >= 0) w += (p1.y - p0.y > 0) ? 1 : -1;
            }
        }
    } else if (k == 1) {
        V2 p0 = p[0], p1 = p[1], p2 = p[2];
        float t[2];
        if (solve_quadratic_f(p0.y - 2 * p1.y + p2.y, -2 * p0.y + 2 * p1.y, p0.y - pt.y, &t[0], &t[1])) {
#pragma unroll
            for (int j = 0; j < 2; j++) {
                if (t[j] >= 0 && t[j] <= 1) {
                    float tp = (p0.x - 2 * p1.x + p2.x) * t[j] * t[j] + (-2 * p0.x + 2 * p1.x) * t[j] + p0.x - pt.x;
                    if (tp >= 0) w += (2 * (p0.y - 2 * p1.y + p2.y) * t[j] + (-2 * p0.y + 2 * p1.y) > 0) ? 1 : -1;
                }
            }
        }
    } else {
        V2 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
        // decided in float whenever every decision of the reference is certain
        // (dvg_winding_fast.h); otherwise the reference's double solve below
        int wfast = 0;
        const float4 q0 = prep[0], q1 = prep[1], q2 = prep[2];
        const wf::CubicPrep q{q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x};
        if (wf::cubic_winding_fast_prep(q, p0.x, p0.y, p1.x, p1.y, p2.x, p2.y, p3.x, p3.y, pt.x, pt.y, wfast))
            return wfast;
        double t[3];
        double ca = (double)(-p0.y + 3 * p1.y - 3 * p2.y + p3.y), cb = (double)(3 * p0.y - 6 * p1.y + 3 * p2.y),
               cc = (double)(-3 * p0.y + 3 * p1.y), cd = (double)(p0.y - pt.y);
        int num_sol = solve_cubic_d(ca, cb, cc, cd, t);
        // a root at a segment end is counted by the side its last bit falls on: take
        // that bit from correctly rounded libm functions, as the host libm gives it
        bool edge = false;
        for (int j = 0; j < num_sol; j++) edge |= fabs(t[j]) < 1e-9 || fabs(t[j] - 1.0) < 1e-9;
        if (edge) num_sol = solve_cubic_d_cr(ca, cb, cc, cd, t);
        float ax = -p0.x + 3 * p1.x - 3 * p2.x + p3.x;
        float bx = 3 * p0.x - 6 * p1.x + 3 * p2.x;
        float cx = -3 * p0.x + 3 * p1.x;
        float ay3 = 3 * (-p0.y + 3 * p1.y - 3 * p2.y + p3.y);
        float by2 = 2 * (3 * p0.y - 6 * p1.y + 3 * p2.y);
        float cy = -3 * p0.y + 3 * p1.y;
        for (int j = 0; j < num_sol; j++) {
            double tj = t[j];
            if (tj >= 0 && tj <= 1) {
                double tp = ax * tj * tj * tj + bx * tj * tj + cx * tj + p0.x - pt.x;
                if (tp > 0) w += (ay3 * tj * tj + by2 * tj + cy > 0) ? 1 : -1;
            }
        }
    }
    return w;
}

DVG_DEV int winding_path(const DevScene &S, int s, V2 pt) {
    int b = S.seg_begin[s], n = S.nsegs[s];
    int wn = 0;
    for (int i = 0; i < n; i++) {
        int seg = b + i;  // the winding sum is order independent
        if (n > 1 && !ray_box(S.seg_box[seg], pt)) continue;
        V2 p[4];
        int4 idx;
        int k = seg_load(S, seg, p, idx);
        wn += segment_winding(k, p, pt, S.seg_wf + kSegPrep * (size_t)seg);
    }
    return wn;
}

DVG_DEV int winding_shape(const DevScene &S, int s, V2 pt) {
    int type = S.shape_type[s];
    if (type == 2) return winding_path(S, s, pt);
    const float *P = S.params + S.param_off[s];
    if (type == 0) {
        float r = P[0];
        return dist_sq(V2{P[1], P[2]}, pt) < r * r ? 1 : 0;
    } else if (type == 1) {
        float dx = P[2] - pt.x, dy = P[3] - pt.y;
        return (dx * dx) / (P[0] * P[0]) + (dy * dy) / (P[1] * P[1]) < 1 ? 1 : 0;
    }
    return (pt.x > P[0] && pt.x < P[2] && pt.y > P[1] && pt.y < P[3]) ? 1 : 0;
}

// is_inside (diffvg.cpp:33-87); `local` is canvas_to_shape(pt)
DVG_DEV bool is_inside_local(const DevScene &S, int g, V2 local, EdgeQuery *eq) {
    if (!in_box(S.g_lbox[g], local)) return false;
    int off = S.g_off[g], n = S.g_n[g];
    int wn = 0;
    bool eo = S.even_odd[g] != 0;
    for (int i = 0; i < n; i++) {
        int s = S.g_ids[off + i];
        if (n > 1 && !in_box(S.shape_box[s], local)) continue;
        int w = winding_shape(S, s, local);
        wn += w;
        if (eq && eq->group == g && eq->shape == s) {
            if ((eo && (w < 0 ? -w : w) % 2 == 1) || (!eo && w != 0)) eq->hit = true;
        }
    }
    if (eo) return (wn < 0 ? -wn : wn) % 2 == 1;
    return wn != 0;
}

DVG_DEV V2 quadratic_closest_pt_approx(V2 b0, V2 b1, V2 b2, V2 pt, float *t_ = nullptr) {
    b0 = b0 - pt;
    b1 = b1 - pt;
    b2 = b2 - pt;
    auto det2 = [](V2 a, V2 b) { return a.x * b.y - b.x * a.y; };
    float a = det2(b0, b2), b = 2 * det2(b1, b0), d = 2 * det2(b2, b1);
    float f = b * d - a * a;
    V2 d21 = b2 - b1, d10 = b1 - b0, d20 = b2 - b0;
    V2 gf = 2 * (b * d21 + d * d10 + a * d20);
    gf = V2{gf.y, -gf.x};
    V2 pp = vdiv(-f * gf, dot(gf, gf));
    V2 d0p = b0 - pp;
    float ap = det2(d0p, d20), bp = 2 * det2(d10, d0p);
    float t = (ap + bp) / (2 * a + b + d);
    t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
    float tt = 1 - t;
    if (t_) *t_ = t;
    return (tt * tt) * b0 + (2 * tt * t) * b1 + (t * t) * b2 + pt;
}

struct Quintic {
    double A, B, C, D, E, F;
};
DVG_DEV Quintic cubic_quintic(V2 p0, V2 p1, V2 p2, V2 p3, V2 pt) {
    V2 a = -p0 + 3 * p1 - 3 * p2 + p3;
    V2 b = 3 * p0 - 6 * p1 + 3 * p2;
    V2 c = -3 * p0 + 3 * p1;
    V2 d = p0 - pt;
    Quintic q;
    q.A = 3 * vsum(a * a);
    q.B = 5 * vsum(a * b);
    q.C = 4 * vsum(a * c) + 2 * vsum(b * b);
    q.D = 3 * (vsum(b * c) + vsum(a * d));
    q.E = vsum(c * c) + 2 * vsum(d * b);
    q.F = vsum(d * c);
    q.B /= q.A;
    q.C /= q.A;
    q.D /= q.A;
    q.E /= q.A;
    q.F /= q.A;
    return q;
}
DVG_DEV double quintic_eval(const Quintic &q, double t) {
    return t * t * t * t * t + q.B * t * t * t * t + q.C * t * t * t + q.D * t * t + q.E * t + q.F;
}
DVG_DEV double quintic_deriv(const Quintic &q, double t) {
    return 5 * t * t * t * t + 4 * q.B * t * t * t + 3 * q.C * t * t + 2 * q.D * t + q.E;
}
// Isolating intervals (within_distance.h:195-216); the reference's uninitialised
// intervals[0] (isolator root outside [0, 1]) is treated as absent.
DVG_DEV int quintic_intervals(const Quintic &q, float *iv) {
    double p1A = ((2 / 5.f) * q.C - (4 / 25.f) * q.B * q.B);
    double p1B = ((3 / 5.f) * q.D - (3 / 25.f) * q.B * q.C);
    double p1C = ((4 / 5.f) * q.E - (2 / 25.f) * q.B * q.D);
    double p1D = q.F - q.B * q.E / 25.f;
    double q_root = -q.B / 5.f;
    double pr[3];
    int num_sol = solve_cubic_d(p1A, p1B, p1C, p1D, pr);
    int n = 0;
    if (q_root >= 0 && q_root <= 1) iv[n++] = (float)q_root;
    for (int j = 0; j < num_sol; j++) iv[n++] = (float)pr[j];
    for (int j = 1; j < n; j++)
        for (int k = j; k > 0 && iv[k - 1] > iv[k]; k--) {
            float tmp = iv[k];
            iv[k] = iv[k - 1];
            iv[k - 1] = tmp;
        }
    return n;
}
DVG_DEV float quintic_root(const Quintic &q, float lb, float ub) {
    float t = 0.5f * (lb + ub);
    for (int it = 0; it < 20; it++) {
        if (!(t >= lb && t <= ub)) t = 0.5f * (lb + ub);
        double value = quintic_eval(q, t);
        if (fabs(value) < 1e-5f || it == 19) break;
        if (value > 0.f)
            ub = t;
        else
            lb = t;
        double derivative = quintic_deriv(q, t);
        t = (float)((double)t - value / derivative);
    }
    return t;
}

// Every point of cubic segment `seg` is certainly farther than r from pt (the chords of an
// 8-step polyline, the chords' deviation bound and a rounding margin, from the segment's
// prep); false for r = inf.
DVG_DEV bool cubic_farther_than(const DevScene &S, int seg, V2 p0, V2 pt, float r) {
    const float4 *q = S.seg_wf + kSegPrep * (size_t)seg + 3;
    float poly[16];
    for (int i = 0; i < 4; i++) {
        const float4 v = q[i];
        poly[4 * i] = v.x;
        poly[4 * i + 1] = v.y;
        poly[4 * i + 2] = v.z;
        poly[4 * i + 3] = v.w;
    }
    const float4 md = q[4];
    return wf::cubic_certainly_farther_prep(poly, md.x, md.y, p0.x, p0.y, pt.x, pt.y, r);
}

// One stroke segment test of within_distance (within_distance.h:32-276).
// Returns 1 (hit), 0 (miss) or -1 (quadratic distance-approximation miss, which ends
// the whole path query in the reference).
DVG_DEV int segment_within(const DevScene &S, int s, int seg, V2 pt, float r) {
    V2 p[4];
    int4 idx;
    int k = seg_load(S, seg, p, idx);
    bool thick = S.has_thick[s] != 0;
    if (k == 0) {
        V2 p0 = p[0], p1 = p[1];
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        float r0 = r, r1 = r;
        if (thick) {
            r0 = S.thickness[idx.x];
            r1 = S.thickness[idx.y];
        }
        if (t < 0) return dist_sq(p0, pt) < r0 * r0;
        if (t > 1) return dist_sq(p1, pt) < r1 * r1;
        float rr = r0 + t * (r1 - r0);
        return dist_sq(p0 + t * (p1 - p0), pt) < rr * rr;
    } else if (k == 1) {
        V2 p0 = p[0], p1 = p[1], p2 = p[2];
        if (S.approx[s]) {
            V2 cp = quadratic_closest_pt_approx(p0, p1, p2, pt);
            return dist_sq(cp, pt) < r * r ? 1 : -1;
        }
        float r0 = r, r1 = r, r2 = r;
        if (thick) {
            r0 = S.thickness[idx.x];
            r1 = S.thickness[idx.y];
            r2 = S.thickness[idx.z];
        }
        if (dist_sq(qeval(p0, p1, p2, 0), pt) < r0 * r0) return 1;
        if (dist_sq(qeval(p0, p1, p2, 1), pt) < r2 * r2) return 1;
        float A = vsum((p0 - 2 * p1 + p2) * (p0 - 2 * p1 + p2));
        float B = vsum(3 * (p0 - 2 * p1 + p2) * (-p0 + p1));
        float C = vsum(2 * (-p0 + p1) * (-p0 + p1) + (p0 - 2 * p1 + p2) * (p0 - pt));
        float D = vsum((-p0 + p1) * (p0 - pt));
        float t[3];
        int num_sol = solve_cubic_f(A, B, C, D, t);
        for (int j = 0; j < num_sol; j++) {
            if (t[j] >= 0 && t[j] <= 1) {
                float tt = 1 - t[j];
                float rr = (tt * tt) * r0 + (2 * tt * t[j]) * r1 + (t[j] * t[j]) * r2;
                if (dist_sq(qeval(p0, p1, p2, t[j]), pt) < rr * rr) return 1;
            }
        }
        return 0;
    }
    V2 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
    float r0 = r, r1 = r, r2 = r, r3 = r;
    if (thick) {
        r0 = S.thickness[idx.x];
        r1 = S.thickness[idx.y];
        r2 = S.thickness[idx.z];
        r3 = S.thickness[idx.w];
    }
    if (dist_sq(p0, pt) < r0 * r0) return 1;  // eval(0) == p0 exactly
    if (dist_sq(ceval(p0, p1, p2, p3, 1), pt) < r3 * r3) return 1;
    // every curve point certainly farther than the largest radius magnitude (the radius
    // at t is a convex combination of r0..r3): the quintic search below would find nothing
    {
        const float4 *q = S.seg_wf + kSegPrep * (size_t)seg + 3;
        float poly[16];
        for (int i = 0; i < 4; i++) {
            const float4 v = q[i];
            poly[4 * i] = v.x;
            poly[4 * i + 1] = v.y;
            poly[4 * i + 2] = v.z;
            poly[4 * i + 3] = v.w;
        }
        const float4 md = q[4];
        if (wf::cubic_certainly_farther_prep(poly, md.x, md.y, p0.x, p0.y, pt.x, pt.y,
                                             fmaxf(fmaxf(fabsf(r0), fabsf(r1)), fmaxf(fabsf(r2), fabsf(r3)))))
            return 0;
    }
    Quintic q = cubic_quintic(p0, p1, p2, p3, pt);
    float iv[4];
    int niv = quintic_intervals(q, iv);
    float lower = 0.f;
    for (int j = 0; j < niv + 1; j++) {
        if (j < niv && iv[j] < 0.f) continue;
        float upper = j < niv ? fmin_r(iv[j], 1.f) : 1.f;
        float lb = lower, ub = upper;
        double lbe = quintic_eval(q, lb), ube = quintic_eval(q, ub);
        if (lbe * ube > 0) continue;
        if (lbe > ube) {
            float tmp = lb;
            lb = ub;
            ub = tmp;
        }
        float t = quintic_root(q, lb, ub);
        float tt = 1 - t;
        float rr = (tt * tt * tt) * r0 + (3 * tt * tt * t) * r1 + (3 * tt * t * t) * r2 + (t * t * t) * r3;
        if (dist_sq(ceval(p0, p1, p2, p3, t), pt) < rr * rr) return 1;
        if (upper >= 1.f) break;
        lower = upper;
    }
    return 0;
}

DVG_DEV bool within_distance_path(const DevScene &S, int s, V2 pt, float r) {
    int b = S.seg_begin[s], n = S.nsegs[s];
    for (int i = 0; i < n; i++) {
        int seg = S.seg_dfs[b + i];
        if (n > 1 && !in_box(S.seg_box[seg], pt, S.seg_maxr[seg])) continue;
        int res = segment_within(S, s, seg, pt, r);
        if (res != 0) return res > 0;
    }
    return false;
}

DVG_DEV bool within_distance_shape(const DevScene &S, int s, V2 pt, float r) {
    int type = S.shape_type[s];
    if (type == 2) return within_distance_path(S, s, pt, r);
    const float *P = S.params + S.param_off[s];
    if (type == 0) return fabs((double)(dist(V2{P[1], P[2]}, pt) - P[0])) < r;
    if (type == 1) return false;  // the reference asserts on stroked ellipses
    V2 lt{P[0], P[1]}, rt{P[2], P[1]}, lb{P[0], P[3]}, rb{P[2], P[3]};
    auto test = [&](V2 p0, V2 p1) {
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        if (t < 0) return dist_sq(p0, pt) < r * r;
        if (t > 1) return dist_sq(p1, pt) < r * r;
        return dist_sq(p0 + t * (p1 - p0), pt) < r * r;
    };
    return test(lt, lb) || test(lt, rt) || test(rt, rb) || test(lb, rb);
}

// within_distance for a group (within_distance.h:355-446)
DVG_DEV bool within_distance_local(const DevScene &S, int g, V2 local, EdgeQuery *eq) {
    bool tracking = eq && eq->group == g;
    int off = S.g_off[g], n = S.g_n[g];
    bool ret = false;
    for (int i = 0; i < n; i++) {
        int s = S.g_ids[off + i];
        float r = S.stroke_width[s];
        if (n > 1 && !in_box(S.shape_box[s], local, r)) continue;
        if (within_distance_shape(S, s, local, r)) {
            if (!tracking) return true;
            ret = true;
            if (s == eq->shape) eq->hit = true;
        }
    }
    return ret;
}

// Closest point on one line or quadratic path segment (compute_distance.h:55-126)
DVG_DEV float segment_closest_lq(const DevScene &S, int s, int seg, V2 pt, V2 &cp, float &t_root) {
    V2 p[4];
    int4 idx;
    int k = seg_load(S, seg, p, idx);
    float dd;
    if (k == 0) {
        V2 p0 = p[0], p1 = p[1];
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        if (t < 0) {
            dd = dist(p0, pt);
            cp = p0;
            t_root = 0;
        } else if (t > 1) {
            dd = dist(p1, pt);
            cp = p1;
            t_root = 1;
        } else {
            dd = dist(p0 + t * (p1 - p0), pt);
            cp = p0 + t * (p1 - p0);
            t_root = t;
        }
        return dd;
    }
    if (k == 1) {
        V2 p0 = p[0], p1 = p[1], p2 = p[2];
        if (S.approx[s]) {
            cp = quadratic_closest_pt_approx(p0, p1, p2, pt, &t_root);
            return dist(cp, pt);
        }
        V2 pt0 = qeval(p0, p1, p2, 0), pt1 = qeval(p0, p1, p2, 1);
        float d0 = dist(pt0, pt), d1 = dist(pt1, pt);
        dd = d0;
        cp = pt0;
        t_root = 0;
        if (d1 < dd) {
            dd = d1;
            cp = pt1;
            t_root = 1;
        }
        float A = vsum((p0 - 2 * p1 + p2) * (p0 - 2 * p1 + p2));
        float B = vsum(3 * (p0 - 2 * p1 + p2) * (-p0 + p1));
        float C = vsum(2 * (-p0 + p1) * (-p0 + p1) + (p0 - 2 * p1 + p2) * (p0 - pt));
        float D = vsum((-p0 + p1) * (p0 - pt));
        float t[3];
        int num_sol = solve_cubic_f(A, B, C, D, t);
        for (int j = 0; j < num_sol; j++) {
            if (t[j] >= 0 && t[j] <= 1) {
                V2 pp = qeval(p0, p1, p2, t[j]);
                float dp = dist(pp, pt);
                if (dp < dd) {
                    dd = dp;
                    cp = pp;
                    t_root = t[j];
                }
            }
        }
        return dd;
    }
    return __builtin_huge_valf();  // cubics: cubic_closest
}

// Closest point on one cubic path segment (compute_distance.h:127-280)
DVG_DEV float cubic_closest(const DevScene &S, int seg, V2 pt, V2 &cp, float &t_root) {
    V2 p[4];
    int4 idx;
    seg_load(S, seg, p, idx);
    float dd;
    V2 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
    V2 pt0 = ceval(p0, p1, p2, p3, 0), pt1 = ceval(p0, p1, p2, p3, 1);
    float d0 = dist(pt0, pt), d1 = dist(pt1, pt);
    dd = d0;
    cp = pt0;
    t_root = 0;
    if (d1 < dd) {
        dd = d1;
        cp = pt1;
        t_root = 1;
    }
    Quintic q = cubic_quintic(p0, p1, p2, p3, pt);
    float iv[4];
    int niv = quintic_intervals(q, iv);
    float lower = 0.f;
    for (int j = 0; j < niv + 1; j++) {
        if (j < niv && iv[j] < 0.f) continue;
        float upper = j < niv ? fmin_r(iv[j], 1.f) : 1.f;
        float lb = lower, ub = upper;
        double lbe = quintic_eval(q, lb), ube = quintic_eval(q, ub);
        if (lbe * ube > 0) continue;
        if (lbe > ube) {
            float tmp = lb;
            lb = ub;
            ub = tmp;
        }
        float t = quintic_root(q, lb, ub);
        V2 pp = ceval(p0, p1, p2, p3, t);
        float dp = dist(pp, pt);
        if (dp < dd) {
            dd = dp;
            cp = pp;
            t_root = t;
        }
        if (upper >= 1.f) break;
        lower = upper;
    }
    return dd;
}

// closest_point for a path (compute_distance.h:27-300): the segments in the reference's
// visit order, the first strict minimum kept.  A cubic whose every point is certainly
// farther than the current minimum cannot lower it and is skipped before its quintic
// search (the bound is checked against the oracle's closest point on the host); the
// minima and ties are the reference's.
DVG_DEV bool closest_point_path(const DevScene &S, int s, V2 pt, float max_radius, PathInfo &info, V2 &result) {
    int b = S.seg_begin[s], n = S.nsegs[s];
    float min_dist = max_radius;
    V2 ret{0, 0};
    bool found = false;
    for (int i = 0; i < n; i++) {
        int seg = S.seg_dfs[b + i];
        if (n > 1 && !in_box(S.seg_box[seg], pt, min_dist)) continue;
        V2 cp;
        float tr;
        float dd;
        if (S.seg_kind[seg] == 2) {
            if (cubic_farther_than(S, seg, ld_pt(S, S.seg_idx[seg].x), pt, min_dist)) continue;
            dd = cubic_closest(S, seg, pt, cp, tr);
        } else {
            dd = segment_closest_lq(S, s, seg, pt, cp, tr);
        }
        if (dd < min_dist) {
            min_dist = dd;
            ret = cp;
            info.base_point_id = seg - b;
            info.point_id = S.seg_pid[seg];
            info.t_root = tr;
            found = true;
        }
    }
    result = ret;
    return found;
}

DVG_DEV bool closest_point_shape(const DevScene &S, int s, V2 pt, float max_radius, PathInfo &info, V2 &result) {
    int type = S.shape_type[s];
    if (type == 2) return closest_point_path(S, s, pt, max_radius, info, result);
    const float *P = S.params + S.param_off[s];
    if (type == 0) {
        V2 c{P[1], P[2]};
        result = c + P[0] * normalize(pt - c);
        return false;  // compute_distance.h:20-25 reports circles as not found
    }
    if (type == 1) return false;
    // rect, including the reference's return of p0 for interior projections
    float min_dist = 0.f;
    V2 closest{0, 0};
    auto update = [&](V2 p0, V2 p1, bool first) {
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        float dd;
        V2 c;
        if (t < 0) {
            dd = dist(p0, pt);
            c = p0;
        } else if (t > 1) {
            dd = dist(p1, pt);
            c = p1;
        } else {
            dd = dist(p0 + t * (p1 - p0), pt);
            c = p0;
        }
        if (first || dd < min_dist) {
            min_dist = dd;
            closest = c;
        }
    };
    V2 lt{P[0], P[1]}, rt{P[2], P[1]}, lb{P[0], P[3]}, rb{P[2], P[3]};
    update(lt, lb, true);
    update(lt, rt, false);
    update(rt, rb, false);
    update(lb, rb, false);
    result = closest;
    return true;
}

// compute_distance (compute_distance.h:375-439)
DVG_DEV bool compute_distance(const DevScene &S, int g, V2 pt, float max_radius, int &min_shape, V2 &closest,
                              PathInfo &info, float &result) {
    M3 c2s = m3_load(S.c2s + 9 * g);
    V2 local = xform_pt(c2s, pt);
    int off = S.g_off[g], n = S.g_n[g];
    float min_dist = max_radius;
    bool found = false;
    M3 s2c;
    bool s2c_loaded = false;
    for (int i = 0; i < n; i++) {
        int s = S.g_ids[off + i];
        if (n > 1 && !in_box(S.shape_box[s], local, max_radius)) continue;
        PathInfo li{-1, -1, 0.f};
        V2 lcp{0, 0};
        if (closest_point_shape(S, s, local, max_radius, li, lcp)) {
            if (!s2c_loaded) {
                s2c = m3_load(S.s2c + 9 * g);
                s2c_loaded = true;
            }
            V2 cp = xform_pt(s2c, lcp);
            float dd = dist(cp, pt);
            if (!found || dd < min_dist) {
                found = true;
                min_dist = dd;
                min_shape = s;
                closest = cp;
                info = li;
            }
        }
    }
    result = min_dist;
    return found;
}

// ---------------------------------------------------------------- adjoints
// d_closest_point (compute_distance.h:442-897).  The quadratic interior branch of the
// reference declares new d_p0..d_p2 that shadow the accumulated ones, so only d_pt
// receives its adjoint there.
// Control-point adjoints of one segment: slot j adds d[j] to point idx[j] when use[j].
struct SegPointAdj {
    V2 d[4];
    int idx[4];
    bool use[4];
};

DVG_DEV int d_closest_point_path_vals(const DevScene &S, int s, V2 pt, V2 d_cp, const PathInfo &info, V2 &d_pt,
                                      SegPointAdj &out) {
    int seg = S.seg_begin[s] + info.base_point_id;
    V2 p[4];
    int4 idx;
    int k = seg_load(S, seg, p, idx);
    out.idx[0] = idx.x;
    out.idx[1] = idx.y;
    out.idx[2] = idx.z;
    out.idx[3] = idx.w;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        out.d[j] = V2{0, 0};
        out.use[j] = false;
    }
    if (k == 0) {
        V2 p0 = p[0], p1 = p[1];
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        V2 d_p0{0, 0}, d_p1{0, 0};
        if (t < 0) {
            d_p0 += d_cp;
        } else if (t > 1) {
            d_p1 += d_cp;
        } else {
            d_p0 += d_cp * (1 - t);
            d_p1 += d_cp * t;
        }
        out.d[0] = d_p0;
        out.d[1] = d_p1;
        out.use[0] = out.use[1] = true;
    } else if (k == 1) {
        V2 p0 = p[0], p1 = p[1], p2 = p[2];
        V2 d_p0{0, 0}, d_p2{0, 0};
        float t = info.t_root;
        if (t == 0) {
            d_p0 += d_cp;
        } else if (t == 1) {
            d_p2 += d_cp;
        } else {
            float A = vsum((p0 - 2 * p1 + p2) * (p0 - 2 * p1 + p2));
            float B = vsum(3 * (p0 - 2 * p1 + p2) * (-p0 + p1));
            float C = vsum(2 * (-p0 + p1) * (-p0 + p1) + (p0 - 2 * p1 + p2) * (p0 - pt));
            float tt = 1 - t;
            float d_tt = 2 * tt * dot(d_cp, p0) + 2 * t * dot(d_cp, p1);
            float d_t = -d_tt + 2 * tt * dot(d_cp, p1) + 2 * t * dot(d_cp, p2);
            float poly_deriv_t = 3 * A * t * t + 2 * B * t + C;
            if (fabs((double)poly_deriv_t) > 1e-6f) {
                float d_C = -(d_t / poly_deriv_t) * t;
                float d_D = -(d_t / poly_deriv_t);
                d_pt += d_C * (-(p0 - 2 * p1 + p2)) + d_D * (-(-p0 + p1));
            }
        }
        out.d[0] = d_p0;
        out.d[2] = d_p2;
        out.use[0] = out.use[2] = true;
    } else {
        V2 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
        V2 d_p0{0, 0}, d_p1{0, 0}, d_p2{0, 0}, d_p3{0, 0};
        float t = info.t_root;
        if (t == 0) {
            d_p0 += d_cp;
        } else if (t == 1) {
            d_p3 += d_cp;
        } else {
            Quintic q = cubic_quintic(p0, p1, p2, p3, pt);
            double A = q.A;
            V2 d_p = d_cp;
            float tt = 1 - t;
            float d_tt = 3 * tt * tt * dot(d_p, p0) + 6 * tt * t * dot(d_p, p1) + 3 * t * t * dot(d_p, p2);
            float d_t = -d_tt + 3 * tt * tt * dot(d_p, p1) + 6 * tt * t * dot(d_p, p2) + 3 * t * t * dot(d_p, p3);
            d_p0 += d_p * (tt * tt * tt);
            d_p1 += d_p * (3 * tt * tt * t);
            d_p2 += d_p * (3 * tt * t * t);
            d_p3 += d_p * (t * t * t);
            double poly_deriv_t = quintic_deriv(q, t);
            if (fabs(poly_deriv_t) > 1e-10f) {
                double r = d_t / poly_deriv_t;
                double d_B = -r * t * t * t * t;
                double d_C = -r * t * t * t;
                double d_D = -r * t * t;
                double d_E = -r * t;
                double d_F = -r;
                double d_A = -d_B * q.B / A - d_C * q.C / A - d_D * q.D / A - d_E * q.E / A - d_F * q.F / A;
                d_B /= A;
                d_C /= A;
                d_D /= A;
                d_E /= A;
                d_F /= A;
                V2 a = -p0 + 3 * p1 - 3 * p2 + p3;
                V2 b = 3 * p0 - 6 * p1 + 3 * p2;
                V2 c = -3 * p0 + 3 * p1;
                V2 d = p0 - pt;
                d_p0 += dmul(d_A * 3 * (-1) * 2, a);
                d_p1 += dmul(d_A * 3 * 3 * 2, a);
                d_p2 += dmul(d_A * 3 * (-3) * 2, a);
                d_p3 += dmul(d_A * 3 * 1 * 2, a);
                d_p0 += dmul(d_B * 5, (-1) * b + 3 * a);
                d_p1 += dmul(d_B * 5, 3 * b + (-6) * a);
                d_p2 += dmul(d_B * 5, (-3) * b + 3 * a);
                d_p3 += dmul(d_B * 5, b);
                d_p0 += dmul(d_C * 4, (-1) * c + (-3) * a) + dmul(d_C * 2, 3 * 2 * b);
                d_p1 += dmul(d_C * 4, 3 * c + 3 * a) + dmul(d_C * 2, (-6) * 2 * b);
                d_p2 += dmul(d_C * 4, (-3) * c) + dmul(d_C * 2, 3 * 2 * b);
                d_p3 += dmul(d_C * 4, c);
                d_p0 += dmul(d_D * 3, 3 * c + (-3) * b) + dmul(d_D * 3, (-1) * d + 1 * a);
                d_p1 += dmul(d_D * 3, (-6) * c + (3) * b) + dmul(d_D * 3, 3 * d);
                d_p2 += dmul(d_D * 3, 3 * c) + dmul(d_D * 3, (-3) * d);
                d_pt += dmul(d_D * 3, (-1) * a);
                d_p0 += dmul(d_E, (-3) * 2 * c) + dmul(d_E * 2, 1 * b + 3 * d);
                d_p1 += dmul(d_E, 3 * 2 * c) + dmul(d_E * 2, (-6) * d);
                d_p2 += dmul(d_E * 2, 3 * d);
                d_pt += dmul(d_E * 2, (-1) * b);
                d_p0 += dmul(d_F, 1 * c) + dmul(d_F, (-3) * d);
                d_p1 += dmul(d_F, 3 * d);
                d_pt += dmul(d_F, (-1) * c);
            }
        }
        out.d[0] = d_p0;
        out.d[1] = d_p1;
        out.d[2] = d_p2;
        out.d[3] = d_p3;
        out.use[0] = out.use[1] = out.use[2] = out.use[3] = true;
    }
    return seg;
}

DVG_DEV void d_closest_point_path(const DevScene &S, int s, V2 pt, V2 d_cp, const PathInfo &info, V2 &d_pt) {
    SegPointAdj a;
    d_closest_point_path_vals(S, s, pt, d_cp, info, d_pt, a);
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (a.use[j]) aadd2(S.d_points + 2 * a.idx[j], a.d[j]);
}

// d_pmin (v[0..1]) and d_pmax (v[2..3]) of a rect
DVG_DEV void d_closest_point_rect_vals(const DevScene &S, int s, V2 pt, V2 d_cp, V2 &d_pt, float (&v)[4]) {
    const float *P = S.params + S.param_off[s];
    auto seg_dist = [&](V2 p0, V2 p1) -> float {
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        if (t < 0) return dist(p0, pt);
        if (t > 1) return dist(p1, pt);
        return dist(p0 + t * (p1 - p0), pt);
    };
    V2 lt{P[0], P[1]}, rt{P[2], P[1]}, lb{P[0], P[3]}, rb{P[2], P[3]};
    float ld = seg_dist(lt, lb), td = seg_dist(lt, rt), rd = seg_dist(rt, rb), bd = seg_dist(lb, rb);
    int min_id = 0;
    float md = ld;
    if (td < md) {
        md = td;
        min_id = 1;
    }
    if (rd < md) {
        md = rd;
        min_id = 2;
    }
    if (bd < md) {
        md = bd;
        min_id = 3;
    }
    V2 d_lt{0, 0}, d_rt{0, 0}, d_lb{0, 0}, d_rb{0, 0};
    auto d_update = [&](V2 p0, V2 p1, V2 &d_p0, V2 &d_p1) {
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        if (t < 0) {
            d_p0 += d_cp;
        } else if (t > 1) {
            d_p1 += d_cp;
        } else {
            d_p0 += d_cp * (1 - t);
            d_p1 += d_cp * t;
            float d_t = vsum(d_cp * (p1 - p0));
            float d_num = d_t / dot(p1 - p0, p1 - p0);
            float d_den = d_t * (-t) / dot(p1 - p0, p1 - p0);
            d_pt += (p1 - p0) * d_num;
            d_p1 += (pt - p0) * d_num;
            d_p0 += ((p0 - p1) + (p0 - pt)) * d_num;
            d_p1 += 2 * (p1 - p0) * d_den;
            d_p0 += 2 * (p0 - p1) * d_den;
        }
    };
    if (min_id == 0)
        d_update(lt, lb, d_lt, d_lb);
    else if (min_id == 1)
        d_update(lt, rt, d_lt, d_rt);
    else if (min_id == 2)
        d_update(rt, rb, d_rt, d_rb);
    else
        d_update(lb, rb, d_lb, d_rb);
    V2 d_pmin{0, 0}, d_pmax{0, 0};
    d_pmin += d_lt;
    d_pmax.x += d_rt.x;
    d_pmin.y += d_rt.y;
    d_pmin.x += d_lb.x;
    d_pmax.y += d_lb.y;
    d_pmax += d_rb;
    v[0] = d_pmin.x;
    v[1] = d_pmin.y;
    v[2] = d_pmax.x;
    v[3] = d_pmax.y;
}

DVG_DEV void d_closest_point_rect(const DevScene &S, int s, V2 pt, V2 d_cp, V2 &d_pt) {
    float v[4];
    d_closest_point_rect_vals(S, s, pt, d_cp, d_pt, v);
    float *dp = S.d_params + S.param_off[s];
    aadd2(dp, V2{v[0], v[1]});
    aadd2(dp + 2, V2{v[2], v[3]});
}

// d_radius (v[0]) and d_center (v[1..2]) of a circle
DVG_DEV void d_closest_point_circle_vals(const DevScene &S, int s, V2 pt, V2 d_cp, float (&v)[4]) {
    const float *P = S.params + S.param_off[s];
    V2 c{P[1], P[2]};
    V2 dn = d_normalize(pt - c, P[0] * d_cp);
    V2 dc = d_cp * V2{1 + dn.x, 1 + dn.y};
    v[0] = dot(d_cp, normalize(pt - c));
    v[1] = dc.x;
    v[2] = dc.y;
    v[3] = 0.f;
}

DVG_DEV void d_closest_point_circle(const DevScene &S, int s, V2 pt, V2 d_cp) {
    float v[4];
    d_closest_point_circle_vals(S, s, pt, d_cp, v);
    float *dp = S.d_params + S.param_off[s];
    aadd2(dp + 1, V2{v[1], v[2]});
    aadd(dp, v[0]);
}

// Adjoint values of one d_compute_distance call, before they are scattered.
struct DistAdj {
    float xform[9];   // shape_to_canvas of the group
    SegPointAdj pts;  // path: control points
    int seg;          // path: segment (the key of its points)
    float par[4];     // circle (3) or rect (4) parameters
    int npar;
    V2 d_pt;          // canvas point (-> d_translation)
};

// d_compute_distance (compute_distance.h:899-949) up to the scatter
DVG_DEV void d_compute_distance_vals(const DevScene &S, int g, int s, V2 pt, V2 closest, const PathInfo &info,
                                     float d_dist, DistAdj &out) {
    M3 c2s = m3_load(S.c2s + 9 * g);
    M3 s2c = m3_load(S.s2c + 9 * g);
    V2 local = xform_pt(c2s, pt);
    V2 local_cp = xform_pt(c2s, closest);
    V2 d_pt{0, 0}, d_cp{0, 0};
    d_distance(closest, pt, d_dist, d_cp, d_pt);
    V2 d_local_cp{0, 0};
    M3 d_s2c;
#pragma unroll
    for (int i = 0; i < 9; i++) d_s2c.m[i] = 0.f;
    d_xform_pt(s2c, local_cp, d_cp, d_s2c, d_local_cp);
    V2 d_local{0, 0};
    int type = S.shape_type[s];
    out.npar = 0;
    out.seg = -1;
#pragma unroll
    for (int j = 0; j < 4; j++) out.pts.use[j] = false;
    if (type == 2) {
        out.seg = d_closest_point_path_vals(S, s, local, d_local_cp, info, d_local, out.pts);
    } else if (type == 3) {
        d_closest_point_rect_vals(S, s, local, d_local_cp, d_local, out.par);
        out.npar = 4;
    } else if (type == 0) {
        d_closest_point_circle_vals(S, s, local, d_local_cp, out.par);
        out.npar = 3;
    }
    M3 d_c2s;
#pragma unroll
    for (int i = 0; i < 9; i++) d_c2s.m[i] = 0.f;
    d_xform_pt(c2s, pt, d_local, d_c2s, d_pt);
    // d_s2c += -c2s^T * d_c2s * c2s^T  (matrix.h products, row-major)
    float A[9], B[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 3; k++) acc += (-c2s.m[3 * k + i]) * d_c2s.m[3 * k + j];
            A[3 * i + j] = acc;
        }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < 3; k++) acc += A[3 * i + k] * c2s.m[3 * j + k];
            B[3 * i + j] = acc;
        }
#pragma unroll
    for (int i = 0; i < 9; i++) out.xform[i] = d_s2c.m[i] + B[i];
    out.d_pt = d_pt;
}

DVG_DEV void d_compute_distance(const DevScene &S, int g, int s, V2 pt, V2 closest, const PathInfo &info, float d_dist,
                                float *d_translation) {
    if (dist_sq(pt, closest) < 1e-10f) return;
    DistAdj a;
    d_compute_distance_vals(S, g, s, pt, closest, info, d_dist, a);
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (a.pts.use[j]) aadd2(S.d_points + 2 * a.pts.idx[j], a.pts.d[j]);
    float *dp = S.d_params + (a.npar ? S.param_off[s] : 0);
    if (a.npar == 4) {
        aadd2(dp, V2{a.par[0], a.par[1]});
        aadd2(dp + 2, V2{a.par[2], a.par[3]});
    } else if (a.npar == 3) {
        aadd2(dp + 1, V2{a.par[1], a.par[2]});
        aadd(dp, a.par[0]);
    }
    float *dx = S.d_xform + 9 * g;
#pragma unroll
    for (int i = 0; i < 9; i++) aadd(dx + i, a.xform[i]);
    if (d_translation) aadd2(d_translation, -a.d_pt);
}

}  // namespace dvg
