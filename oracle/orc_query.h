// orc_query.h -- oracle per-sample geometry queries (test infrastructure only).
//
// winding numbers (winding_number.h:9-202), is_inside (diffvg.cpp:33-87),
// within_distance (within_distance.h:8-446), closest_point / compute_distance
// (compute_distance.h:18-439) and their adjoints (compute_distance.h:442-949).
// BVHs are traversed depth first with the reference's stacks and child order.
#pragma once

#include <cstdlib>

#include "orc_scene.h"

namespace orc {

struct EdgeQuery {  // edge_query.h:3-7
    int shape_group_id;
    int shape_id;
    bool hit;
};

struct PathInfo {  // ClosestPointPathInfo (compute_distance.h:12-16)
    int base_point_id;
    int point_id;
    float t_root;
};

struct Counters {
    int64_t c[12];
};
extern thread_local Counters *g_tc;
inline void bump(int i, int64_t v = 1) {
    if (g_tc) g_tc->c[i] += v;
}
// Set when a query of the current sample reaches the isolating-interval slot that the
// reference leaves uninitialised (see quintic_intervals); read by the render loops
// to tally the samples whose result may depend on that slot.
extern thread_local bool t_q8_slot;
enum {
    kCntSamples = 0,
    kCntTopBox,
    kCntGroupBox,
    kCntSegBox,
    kCntLine,
    kCntQuad,
    kCntCubicWinding,
    kCntCubicDist,
    kCntNewton,
    kCntClosest,
    kCntFragments,
    kCntBoundary
};

// Segment point indices of path `s` for a BVH leaf (winding_number.h:58-131).
inline int seg_fetch(const Scene &S, int s, int base_point_id, int point_id, V2 *p, int *idx) {
    int n = S.npts[s];
    int k = S.ncp[S.ctrl_off[s] + base_point_id];
    idx[0] = point_id;
    if (k == 0) {
        idx[1] = (point_id + 1) % n;
    } else if (k == 1) {
        idx[1] = point_id + 1;
        idx[2] = (point_id + 2) % n;
    } else {
        idx[1] = point_id + 1;
        idx[2] = point_id + 2;
        idx[3] = (point_id + 3) % n;
    }
    for (int j = 0; j <= k + 1; j++) p[j] = S.pt(s, idx[j]);
    return k;
}

inline bool ray_box(const Box &b, V2 pt) {  // intersect (winding_number.h:33-42)
    if (pt.y < b.miny || pt.y > b.maxy) return false;
    if (pt.x > b.maxx) return false;
    return true;
}

// Winding contribution of one path segment (winding_number.h:58-167).
inline int segment_winding(int k, const V2 *p, V2 pt) {
    int w = 0;
    if (k == 0) {
        bump(kCntLine);
        V2 p0 = p[0], p1 = p[1];
        if (p1.y != p0.y) {
            float t = (pt.y - p0.y) / (p1.y - p0.y);
            if (t >= 0 && t <= 1) {
                float tp = p0.x - pt.x + t * (p1.x - p0.x);
                if (tp >= 0) w += (p1.y - p0.y > 0) ? 1 : -1;
            }
        }
    } else if (k == 1) {
        bump(kCntQuad);
        V2 p0 = p[0], p1 = p[1], p2 = p[2];
        float t[2];
        if (solve_quadratic_f(p0.y - 2 * p1.y + p2.y, -2 * p0.y + 2 * p1.y, p0.y - pt.y, &t[0], &t[1])) {
            for (int j = 0; j < 2; j++) {
                if (t[j] >= 0 && t[j] <= 1) {
                    float tp = (p0.x - 2 * p1.x + p2.x) * t[j] * t[j] + (-2 * p0.x + 2 * p1.x) * t[j] +
                               p0.x - pt.x;
                    if (tp >= 0) {
                        if (2 * (p0.y - 2 * p1.y + p2.y) * t[j] + (-2 * p0.y + 2 * p1.y) > 0)
                            w += 1;
                        else
                            w -= 1;
                    }
                }
            }
        }
    } else {
        bump(kCntCubicWinding);
        V2 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
        double t[3];
        int num_sol = solve_cubic_d((double)(-p0.y + 3 * p1.y - 3 * p2.y + p3.y),
                                    (double)(3 * p0.y - 6 * p1.y + 3 * p2.y),
                                    (double)(-3 * p0.y + 3 * p1.y), (double)(p0.y - pt.y), t);
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
                if (tp > 0) {
                    if (ay3 * tj * tj + by2 * tj + cy > 0)
                        w += 1;
                    else
                        w -= 1;
                }
            }
        }
    }
    return w;
}

// compute_winding_number for a path (winding_number.h:44-174)
inline int winding_path(const Scene &S, int s, V2 pt) {
    const Node *nodes = S.pbvh(s);
    int num_segments = S.nbase[s];
    int stack[128];
    int sp = 0;
    int wn = 0;
    stack[sp++] = 2 * num_segments - 2;
    while (sp > 0) {
        const Node &node = nodes[stack[--sp]];
        if (node.child1 < 0) {
            V2 p[4];
            int idx[4];
            int k = seg_fetch(S, s, node.child0, -node.child1 - 1, p, idx);
            wn += segment_winding(k, p, pt);
        } else {
            bump(kCntSegBox, 2);
            if (ray_box(nodes[node.child0].box, pt)) stack[sp++] = node.child0;
            if (ray_box(nodes[node.child1].box, pt)) stack[sp++] = node.child1;
            if (sp > 128) std::abort();
        }
    }
    return wn;
}

// compute_winding_number dispatch (winding_number.h:9-31, 176-202)
inline int winding_shape(const Scene &S, int s, V2 pt) {
    const float *P = S.param_off[s] >= 0 ? &S.params[S.param_off[s]] : nullptr;
    switch (S.type[s]) {
        case kCircle: {
            float r = P[0];
            return dist_sq(V2{P[1], P[2]}, pt) < r * r ? 1 : 0;
        }
        case kEllipse: {
            float dx = P[2] - pt.x, dy = P[3] - pt.y;
            return (dx * dx) / (P[0] * P[0]) + (dy * dy) / (P[1] * P[1]) < 1 ? 1 : 0;
        }
        case kPath:
            return winding_path(S, s, pt);
        case kRect:
            return (pt.x > P[0] && pt.x < P[2] && pt.y > P[1] && pt.y < P[3]) ? 1 : 0;
    }
    return 0;
}

// is_inside (diffvg.cpp:33-87)
inline bool is_inside(const Scene &S, int g, V2 pt, EdgeQuery *eq) {
    V2 local_pt = xform_pt(S.canvas_to_shape[g], pt);
    const Node *nodes = S.gbvh(g);
    int n = S.g_n[g];
    if (!inside(nodes[2 * n - 2].box, local_pt)) return false;
    int wn = 0;
    int stack[64];
    int sp = 0;
    stack[sp++] = 2 * n - 2;
    while (sp > 0) {
        const Node &node = nodes[stack[--sp]];
        if (node.child1 < 0) {
            int shape_id = node.child0;
            int w = winding_shape(S, shape_id, local_pt);
            wn += w;
            if (eq != nullptr && eq->shape_group_id == g && eq->shape_id == shape_id) {
                if ((S.even_odd[g] && std::abs(w) % 2 == 1) || (!S.even_odd[g] && w != 0)) eq->hit = true;
            }
        } else {
            bump(kCntGroupBox, 2);
            if (inside(nodes[node.child0].box, local_pt)) stack[sp++] = node.child0;
            if (inside(nodes[node.child1].box, local_pt)) stack[sp++] = node.child1;
            if (sp > 64) std::abort();
        }
    }
    if (S.even_odd[g]) return std::abs(wn) % 2 == 1;
    return wn != 0;
}

// quadratic_closest_pt_approx (vector.h:785-817)
inline float det2(V2 a, V2 b) { return a.x * b.y - b.x * a.y; }
inline V2 quadratic_closest_pt_approx0(V2 b0, V2 b1, V2 b2, float *t_) {
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
    if (t_ != nullptr) *t_ = t;
    return (tt * tt) * b0 + (2 * tt * t) * b1 + (t * t) * b2;
}
inline V2 quadratic_closest_pt_approx(V2 b0, V2 b1, V2 b2, V2 pt, float *t = nullptr) {
    return quadratic_closest_pt_approx0(b0 - pt, b1 - pt, b2 - pt, t) + pt;
}

inline V2 qeval(V2 p0, V2 p1, V2 p2, float t) {
    float tt = 1 - t;
    return (tt * tt) * p0 + (2 * tt * t) * p1 + (t * t) * p2;
}
inline V2 ceval(V2 p0, V2 p1, V2 p2, V2 p3, float t) {
    float tt = 1 - t;
    return (tt * tt * tt) * p0 + (3 * tt * tt * t) * p1 + (3 * tt * t * t) * p2 + (t * t * t) * p3;
}

// Normalised quintic of the closest-point condition for a cubic (within_distance.h:174-194,
// compute_distance.h:166-185): coefficients B..F after division by A, all double; the
// float sums are formed first and then widened, as the reference does.
struct Quintic {
    double A, B, C, D, E, F;
};
inline Quintic cubic_quintic(V2 p0, V2 p1, V2 p2, V2 p3, V2 pt) {
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
inline double quintic_eval(const Quintic &q, double t) {
    return t * t * t * t * t + q.B * t * t * t * t + q.C * t * t * t + q.D * t * t + q.E * t + q.F;
}
inline double quintic_deriv(const Quintic &q, double t) {
    return 5 * t * t * t * t + 4 * q.B * t * t * t + 3 * q.C * t * t + 2 * q.D * t + q.E;
}
// Isolating intervals (within_distance.h:195-216).  The reference leaves intervals[0]
// uninitialised when the isolator root -B/5 is outside [0, 1] and still counts it; the
// oracle treats that slot as absent.
inline int quintic_intervals(const Quintic &q, float *intervals) {
    double p1A = ((2 / 5.f) * q.C - (4 / 25.f) * q.B * q.B);
    double p1B = ((3 / 5.f) * q.D - (3 / 25.f) * q.B * q.C);
    double p1C = ((4 / 5.f) * q.E - (2 / 25.f) * q.B * q.D);
    double p1D = q.F - q.B * q.E / 25.f;
    double q_root = -q.B / 5.f;
    double p_roots[3];
    int num_sol = solve_cubic_d(p1A, p1B, p1C, p1D, p_roots);
    int n = 0;
    if (q_root >= 0 && q_root <= 1)
        intervals[n++] = (float)q_root;
    else
        t_q8_slot = true;
    for (int j = 0; j < num_sol; j++) intervals[n++] = (float)p_roots[j];
    for (int j = 1; j < n; j++) {
        for (int k = j; k > 0 && intervals[k - 1] > intervals[k]; k--) {
            float tmp = intervals[k];
            intervals[k] = intervals[k - 1];
            intervals[k - 1] = tmp;
        }
    }
    return n;
}
// Newton + bisection on one bracket (within_distance.h:239-257); returns t.
inline float quintic_root(const Quintic &q, float lb, float ub) {
    float t = 0.5f * (lb + ub);
    for (int it = 0; it < 20; it++) {
        if (!(t >= lb && t <= ub)) t = 0.5f * (lb + ub);
        double value = quintic_eval(q, t);
        bump(kCntNewton);
        if (std::fabs(value) < 1e-5f || it == 19) break;
        if (value > 0.f)
            ub = t;
        else
            lb = t;
        double derivative = quintic_deriv(q, t);
        t = (float)((double)t - value / derivative);
    }
    return t;
}

// within_distance for a path (within_distance.h:18-290)
inline bool within_distance_path(const Scene &S, int s, V2 pt, float r) {
    const Node *nodes = S.pbvh(s);
    int num_segments = S.nbase[s];
    int stack[128];
    int sp = 0;
    stack[sp++] = 2 * num_segments - 2;
    bool thick = S.has_thick[s];
    while (sp > 0) {
        const Node &node = nodes[stack[--sp]];
        if (node.child1 < 0) {
            V2 p[4];
            int idx[4];
            int k = seg_fetch(S, s, node.child0, -node.child1 - 1, p, idx);
            if (k == 0) {
                bump(kCntLine);
                V2 p0 = p[0], p1 = p[1];
                float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
                float r0 = r, r1 = r;
                if (thick) {
                    r0 = S.thick(s, idx[0]);
                    r1 = S.thick(s, idx[1]);
                }
                if (t < 0) {
                    if (dist_sq(p0, pt) < r0 * r0) return true;
                } else if (t > 1) {
                    if (dist_sq(p1, pt) < r1 * r1) return true;
                } else {
                    float rr = r0 + t * (r1 - r0);
                    if (dist_sq(p0 + t * (p1 - p0), pt) < rr * rr) return true;
                }
            } else if (k == 1) {
                bump(kCntQuad);
                V2 p0 = p[0], p1 = p[1], p2 = p[2];
                if (S.approx[s]) {
                    V2 cp = quadratic_closest_pt_approx(p0, p1, p2, pt);
                    return dist_sq(cp, pt) < r * r;  // returns from the whole path query
                }
                float r0 = r, r1 = r, r2 = r;
                if (thick) {
                    r0 = S.thick(s, idx[0]);
                    r1 = S.thick(s, idx[1]);
                    r2 = S.thick(s, idx[2]);
                }
                if (dist_sq(qeval(p0, p1, p2, 0), pt) < r0 * r0) return true;
                if (dist_sq(qeval(p0, p1, p2, 1), pt) < r2 * r2) return true;
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
                        V2 pp = qeval(p0, p1, p2, t[j]);
                        if (dist_sq(pp, pt) < rr * rr) return true;
                    }
                }
            } else {
                bump(kCntCubicDist);
                V2 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
                float r0 = r, r1 = r, r2 = r, r3 = r;
                if (thick) {
                    r0 = S.thick(s, idx[0]);
                    r1 = S.thick(s, idx[1]);
                    r2 = S.thick(s, idx[2]);
                    r3 = S.thick(s, idx[3]);
                }
                if (dist_sq(ceval(p0, p1, p2, p3, 0), pt) < r0 * r0) return true;
                if (dist_sq(ceval(p0, p1, p2, p3, 1), pt) < r3 * r3) return true;
                Quintic q = cubic_quintic(p0, p1, p2, p3, pt);
                float intervals[4];
                int num_intervals = quintic_intervals(q, intervals);
                float lower_bound = 0.f;
                for (int j = 0; j < num_intervals + 1; j++) {
                    if (j < num_intervals && intervals[j] < 0.f) continue;
                    float upper_bound = j < num_intervals ? fmin_r(intervals[j], 1.f) : 1.f;
                    float lb = lower_bound, ub = upper_bound;
                    double lb_eval = quintic_eval(q, lb);
                    double ub_eval = quintic_eval(q, ub);
                    if (lb_eval * ub_eval > 0) continue;
                    if (lb_eval > ub_eval) {
                        float tmp = lb;
                        lb = ub;
                        ub = tmp;
                    }
                    float t = quintic_root(q, lb, ub);
                    float tt = 1 - t;
                    float rr = (tt * tt * tt) * r0 + (3 * tt * tt * t) * r1 + (3 * tt * t * t) * r2 +
                               (t * t * t) * r3;
                    if (dist_sq(ceval(p0, p1, p2, p3, t), pt) < rr * rr) return true;
                    if (upper_bound >= 1.f) break;
                    lower_bound = upper_bound;
                }
            }
        } else {
            bump(kCntSegBox, 2);
            if (inside(nodes[node.child0].box, pt, nodes[node.child0].max_radius)) stack[sp++] = node.child0;
            if (inside(nodes[node.child1].box, pt, nodes[node.child1].max_radius)) stack[sp++] = node.child1;
            if (sp > 128) std::abort();
        }
    }
    return false;
}

// within_distance for rect (within_distance.h:292-334)
inline bool within_distance_rect(const float *P, V2 pt, float r) {
    auto test = [&](V2 p0, V2 p1) {
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        if (t < 0) return dist_sq(p0, pt) < r * r;
        if (t > 1) return dist_sq(p1, pt) < r * r;
        return dist_sq(p0 + t * (p1 - p0), pt) < r * r;
    };
    V2 lt{P[0], P[1]}, rt{P[2], P[1]}, lb{P[0], P[3]}, rb{P[2], P[3]};
    if (test(lt, lb)) return true;
    if (test(lt, rt)) return true;
    if (test(rt, rb)) return true;
    if (test(lb, rb)) return true;
    return false;
}

inline bool within_distance_shape(const Scene &S, int s, V2 pt, float r) {  // within_distance.h:336-353
    const float *P = S.param_off[s] >= 0 ? &S.params[S.param_off[s]] : nullptr;
    switch (S.type[s]) {
        case kCircle:
            return std::fabs((double)(dist(V2{P[1], P[2]}, pt) - P[0])) < r;
        case kEllipse:
            return false;  // the reference asserts here (within_distance.h:342-345)
        case kPath:
            return within_distance_path(S, s, pt, r);
        case kRect:
            return within_distance_rect(P, pt, r);
    }
    return false;
}

// within_distance for a group, with the EdgeQuery variant (within_distance.h:355-446)
inline bool within_distance_group(const Scene &S, int g, V2 pt, EdgeQuery *eq) {
    bool tracking = eq != nullptr && g == eq->shape_group_id;
    V2 local_pt = xform_pt(S.canvas_to_shape[g], pt);
    const Node *nodes = S.gbvh(g);
    int stack[64];
    int sp = 0;
    stack[sp++] = 2 * S.g_n[g] - 2;
    bool ret = false;
    while (sp > 0) {
        const Node &node = nodes[stack[--sp]];
        if (node.child1 < 0) {
            int shape_id = node.child0;
            if (within_distance_shape(S, shape_id, local_pt, S.stroke_width[shape_id])) {
                if (!tracking) return true;
                ret = true;
                if (shape_id == eq->shape_id) eq->hit = true;
            }
        } else {
            bump(kCntGroupBox, 2);
            if (inside(nodes[node.child0].box, local_pt, nodes[node.child0].max_radius)) stack[sp++] = node.child0;
            if (inside(nodes[node.child1].box, local_pt, nodes[node.child1].max_radius)) stack[sp++] = node.child1;
            if (sp > 64) std::abort();
        }
    }
    return ret;
}

// closest_point for a path (compute_distance.h:27-300)
inline bool closest_point_path(const Scene &S, int s, V2 pt, float max_radius, PathInfo *info, V2 *result) {
    const Node *nodes = S.pbvh(s);
    float min_dist = max_radius;
    V2 ret_pt{0, 0};
    bool found = false;
    int stack[128];
    int sp = 0;
    stack[sp++] = 2 * S.nbase[s] - 2;
    while (sp > 0) {
        const Node &node = nodes[stack[--sp]];
        if (node.child1 < 0) {
            int base_point_id = node.child0;
            int point_id = -node.child1 - 1;
            V2 p[4];
            int idx[4];
            int k = seg_fetch(S, s, base_point_id, point_id, p, idx);
            float dd = 0.f;
            V2 cp{0, 0};
            float t_root = 0.f;
            bump(kCntClosest);
            bump(k == 0 ? kCntLine : (k == 1 ? kCntQuad : kCntCubicDist));
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
            } else if (k == 1) {
                V2 p0 = p[0], p1 = p[1], p2 = p[2];
                if (S.approx[s]) {
                    cp = quadratic_closest_pt_approx(p0, p1, p2, pt, &t_root);
                    dd = dist(cp, pt);
                } else {
                    V2 pt0 = qeval(p0, p1, p2, 0), pt1 = qeval(p0, p1, p2, 1);
                    float dist0 = dist(pt0, pt), dist1 = dist(pt1, pt);
                    dd = dist0;
                    cp = pt0;
                    t_root = 0;
                    if (dist1 < dd) {
                        dd = dist1;
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
                            float distp = dist(pp, pt);
                            if (distp < dd) {
                                dd = distp;
                                cp = pp;
                                t_root = t[j];
                            }
                        }
                    }
                }
            } else {
                V2 p0 = p[0], p1 = p[1], p2 = p[2], p3 = p[3];
                V2 pt0 = ceval(p0, p1, p2, p3, 0), pt1 = ceval(p0, p1, p2, p3, 1);
                float dist0 = dist(pt0, pt), dist1 = dist(pt1, pt);
                dd = dist0;
                cp = pt0;
                t_root = 0;
                if (dist1 < dd) {
                    dd = dist1;
                    cp = pt1;
                    t_root = 1;
                }
                Quintic q = cubic_quintic(p0, p1, p2, p3, pt);
                float intervals[4];
                int num_intervals = quintic_intervals(q, intervals);
                float lower_bound = 0.f;
                for (int j = 0; j < num_intervals + 1; j++) {
                    if (j < num_intervals && intervals[j] < 0.f) continue;
                    float upper_bound = j < num_intervals ? fmin_r(intervals[j], 1.f) : 1.f;
                    float lb = lower_bound, ub = upper_bound;
                    double lb_eval = quintic_eval(q, lb);
                    double ub_eval = quintic_eval(q, ub);
                    if (lb_eval * ub_eval > 0) continue;
                    if (lb_eval > ub_eval) {
                        float tmp = lb;
                        lb = ub;
                        ub = tmp;
                    }
                    float t = quintic_root(q, lb, ub);
                    V2 pp = ceval(p0, p1, p2, p3, t);
                    float distp = dist(pp, pt);
                    if (distp < dd) {
                        dd = distp;
                        cp = pp;
                        t_root = t;
                    }
                    if (upper_bound >= 1.f) break;
                    lower_bound = upper_bound;
                }
            }
            if (dd < min_dist) {
                min_dist = dd;
                ret_pt = cp;
                info->base_point_id = base_point_id;
                info->point_id = point_id;
                info->t_root = t_root;
                found = true;
            }
        } else {
            bump(kCntSegBox, 2);
            if (inside(nodes[node.child0].box, pt, min_dist)) stack[sp++] = node.child0;
            if (inside(nodes[node.child1].box, pt, min_dist)) stack[sp++] = node.child1;
            if (sp > 128) std::abort();
        }
    }
    *result = ret_pt;
    return found;
}

// closest_point for rect, including the reference's p0 return (compute_distance.h:305-345)
inline bool closest_point_rect(const float *P, V2 pt, V2 *result) {
    float min_dist = 0.f;
    V2 closest{0, 0};
    auto update = [&](V2 p0, V2 p1, bool first) {
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        if (t < 0) {
            float d = dist(p0, pt);
            if (first || d < min_dist) {
                min_dist = d;
                closest = p0;
            }
        } else if (t > 1) {
            float d = dist(p1, pt);
            if (first || d < min_dist) {
                min_dist = d;
                closest = p1;
            }
        } else {
            V2 pp = p0 + t * (p1 - p0);
            float d = dist(pp, pt);
            if (first || d < min_dist) {
                min_dist = d;
                closest = p0;
            }
        }
    };
    V2 lt{P[0], P[1]}, rt{P[2], P[1]}, lb{P[0], P[3]}, rb{P[2], P[3]};
    update(lt, lb, true);
    update(lt, rt, false);
    update(rt, rb, false);
    update(lb, rb, false);
    *result = closest;
    return true;
}

inline bool closest_point_shape(const Scene &S, int s, V2 pt, float max_radius, PathInfo *info, V2 *result) {
    const float *P = S.param_off[s] >= 0 ? &S.params[S.param_off[s]] : nullptr;
    switch (S.type[s]) {
        case kCircle: {
            V2 c{P[1], P[2]};
            *result = c + P[0] * normalize(pt - c);
            return false;  // compute_distance.h:20-25 reports "not found" for circles
        }
        case kEllipse:
            return false;
        case kPath:
            return closest_point_path(S, s, pt, max_radius, info, result);
        case kRect:
            return closest_point_rect(P, pt, result);
    }
    return false;
}

// compute_distance (compute_distance.h:375-439)
inline bool compute_distance(const Scene &S, int g, V2 pt, float max_radius, int *min_shape_id, V2 *closest_pt_,
                             PathInfo *path_info, float *result) {
    V2 local_pt = xform_pt(S.canvas_to_shape[g], pt);
    const Node *nodes = S.gbvh(g);
    int stack[64];
    int sp = 0;
    stack[sp++] = 2 * S.g_n[g] - 2;
    float min_dist = max_radius;
    bool found = false;
    while (sp > 0) {
        const Node &node = nodes[stack[--sp]];
        if (node.child1 < 0) {
            int shape_id = node.child0;
            PathInfo local_info{-1, -1, 0.f};
            V2 lcp{0, 0};
            if (closest_point_shape(S, shape_id, local_pt, max_radius, &local_info, &lcp)) {
                V2 cp = xform_pt(S.shape_to_canvas[g], lcp);
                float dd = dist(cp, pt);
                if (!found || dd < min_dist) {
                    found = true;
                    min_dist = dd;
                    if (min_shape_id) *min_shape_id = shape_id;
                    if (closest_pt_) *closest_pt_ = cp;
                    if (path_info) *path_info = local_info;
                }
            }
        } else {
            bump(kCntGroupBox, 2);
            if (inside(nodes[node.child0].box, local_pt, max_radius)) stack[sp++] = node.child0;
            if (inside(nodes[node.child1].box, local_pt, max_radius)) stack[sp++] = node.child1;
            if (sp > 64) std::abort();
        }
    }
    *result = min_dist;
    return found;
}

}  // namespace orc
