// orc_render.cpp -- oracle render passes and C API (test infrastructure only).
//
// Restates diffvg.cpp: accumulate_boundary_gradient (89-274), sample_color for colour
// types (276-368) and its adjoint (370-504), the fragment compositor sample_color
// (525-707), sample_distance (709-775), gather_d_color (777-815),
// sample_color_prefiltered (835-1113), weight_kernel (1115-1158), render_kernel
// (1160-1312), sample_boundary_kernel (1325-1386), render_edge_kernel (1388-1475) and
// render() (1477-1649); d_closest_point / d_compute_distance (compute_distance.h:442-949)
// and sample_boundary (sample_boundary.h:20-453).  The CPU schedule follows the
// reference's parallel_for (parallel.h:59-91): ceil(count/256) tasks of 256 consecutive
// samples on hardware_concurrency() threads, float atomics by compare-and-swap.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "orc_query.h"

namespace orc {

thread_local Counters *g_tc = nullptr;

static thread_local std::string g_err;
static int g_num_threads = 0;
static bool g_counting = false;
static Counters g_counters[2];  // phase 0: weight + render kernels, 1: boundary sampling + edges
static int g_phase = 0;
static int g_stride = 1;        // evaluate only every g_stride-th sample (cost counting)
static std::mutex g_counter_mu;

static bool g_multi = false;  // atomics needed only when several threads run

static inline void aadd(float *p, float v) {
    if (!g_multi) {
        *p += v;
        return;
    }
    float old_v, new_v;
    __atomic_load(p, &old_v, __ATOMIC_RELAXED);
    do {
        new_v = old_v + v;
    } while (!__atomic_compare_exchange(p, &old_v, &new_v, true, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
}
static inline void aadd2(float *p, V2 v) {
    aadd(p, v.x);
    aadd(p + 1, v.y);
}
static inline void aadd4(float *p, V4 v) {
    aadd(p, v.x);
    aadd(p + 1, v.y);
    aadd(p + 2, v.z);
    aadd(p + 3, v.w);
}
static inline void aaddm(float *p, const M3 &m) {
    for (int i = 0; i < 9; i++) aadd(p + i, m.m[i / 3][i % 3]);
}

struct V2d {
    double x, y;
};
static inline V2d dmul(double s, V2 v) { return {s * v.x, s * v.y}; }
static inline V2d operator+(V2d a, V2d b) { return {a.x + b.x, a.y + b.y}; }
static inline void operator+=(V2 &a, V2d b) {
    a.x = (float)((double)a.x + b.x);
    a.y = (float)((double)a.y + b.y);
}

// ---------------------------------------------------------------- colours
static V4 stop_color(const float *C, int i) { return V4{C[4 * i], C[4 * i + 1], C[4 * i + 2], C[4 * i + 3]}; }

// sample_color(ColorType, ...) (diffvg.cpp:276-368)
static V4 sample_paint(const Scene &S, int type, int off, int nstops, V2 pt) {
    const float *c = &S.color[off];
    if (type == kConstant) return V4{c[0], c[1], c[2], c[3]};
    const float *offsets = c + 4;
    const float *colors = c + 4 + nstops;
    float t;
    if (type == kLinear) {
        V2 beg{c[0], c[1]}, end{c[2], c[3]};
        t = dot(pt - beg, end - beg) / fmax_r(dot(end - beg, end - beg), 1e-3f);
    } else {
        V2 center{c[0], c[1]}, radius{c[2], c[3]};
        V2 offset = pt - center;
        V2 normalized_offset = vdivv(offset, radius);
        t = len(normalized_offset);
    }
    if (t < offsets[0]) return stop_color(colors, 0);
    for (int i = 0; i < nstops - 1; i++) {
        float oc = offsets[i], on = offsets[i + 1];
        if (t >= oc && t < on) {
            float tt = (t - oc) / (on - oc);
            return stop_color(colors, i) * (1 - tt) + stop_color(colors, i + 1) * tt;
        }
    }
    return stop_color(colors, nstops - 1);
}

// d_sample_color (diffvg.cpp:370-504); the radial branch keeps the reference's missing
// `return` and therefore also adds d_color to the last stop.
static void d_sample_paint(Scene &S, int type, int off, int nstops, V2 pt, V4 d_color, float *d_translation) {
    const float *c = &S.color[off];
    float *dc = &S.d_color[off];
    if (type == kConstant) {
        aadd4(dc, d_color);
        return;
    }
    const float *offsets = c + 4;
    const float *colors = c + 4 + nstops;
    float *d_offsets = dc + 4;
    float *d_colors = dc + 4 + nstops;
    if (type == kLinear) {
        V2 beg{c[0], c[1]}, end{c[2], c[3]};
        float t = dot(pt - beg, end - beg) / fmax_r(dot(end - beg, end - beg), 1e-3f);
        if (t < offsets[0]) {
            aadd4(d_colors, d_color);
            return;
        }
        for (int i = 0; i < nstops - 1; i++) {
            float oc = offsets[i], on = offsets[i + 1];
            if (t >= oc && t < on) {
                V4 cc = stop_color(colors, i), cn = stop_color(colors, i + 1);
                float tt = (t - oc) / (on - oc);
                V4 d_cc = d_color * (1 - tt);
                V4 d_cn = d_color * tt;
                float d_tt = vsum(d_color * (cn - cc));
                float d_on = -d_tt * tt / (on - oc);
                float d_oc = d_tt * ((tt - 1.f) / (on - oc));
                float d_t = d_tt / (on - oc);
                aadd4(d_colors + 4 * i, d_cc);
                aadd4(d_colors + 4 * (i + 1), d_cn);
                aadd(d_offsets + i, d_oc);
                aadd(d_offsets + i + 1, d_on);
                float l = fmax_r(dot(end - beg, end - beg), 1e-3f);
                V2 d_beg = vdiv(d_t * (-(pt - beg) - (end - beg)), l);
                V2 d_end = vdiv(d_t * (pt - beg), l);
                float d_l = -d_t * t / l;
                if (dot(end - beg, end - beg) > 1e-3f) {
                    d_beg += 2 * d_l * (beg - end);
                    d_end += 2 * d_l * (end - beg);
                }
                aadd2(dc, d_beg);
                aadd2(dc + 2, d_end);
                if (d_translation) aadd2(d_translation, d_beg + d_end);
                return;
            }
        }
        aadd4(d_colors + 4 * (nstops - 1), d_color);
        return;
    }
    V2 center{c[0], c[1]}, radius{c[2], c[3]};
    V2 offset = pt - center;
    V2 normalized_offset = vdivv(offset, radius);
    float t = len(normalized_offset);
    if (t < offsets[0]) {
        aadd4(d_colors, d_color);
        return;
    }
    for (int i = 0; i < nstops - 1; i++) {
        float oc = offsets[i], on = offsets[i + 1];
        if (t >= oc && t < on) {
            V4 cc = stop_color(colors, i), cn = stop_color(colors, i + 1);
            float tt = (t - oc) / (on - oc);
            V4 d_cc = d_color * (1 - tt);
            V4 d_cn = d_color * tt;
            float d_tt = vsum(d_color * (cn - cc));
            float d_on = -d_tt * tt / (on - oc);
            float d_oc = d_tt * ((tt - 1.f) / (on - oc));
            float d_t = d_tt / (on - oc);
            aadd4(d_colors + 4 * i, d_cc);
            aadd4(d_colors + 4 * (i + 1), d_cn);
            aadd(d_offsets + i, d_oc);
            aadd(d_offsets + i + 1, d_on);
            V2 d_no = d_length(normalized_offset, d_t);
            V2 d_offset = vdivv(d_no, radius);
            V2 d_radius = vdivv((-d_no) * offset, radius * radius);
            V2 d_center = -d_offset;
            aadd2(dc, d_center);
            aadd2(dc + 2, d_radius);
            if (d_translation) aadd2(d_translation, d_center);
        }
    }
    aadd4(d_colors + 4 * (nstops - 1), d_color);
}

// ---------------------------------------------------------------- distance adjoints
// d_closest_point (compute_distance.h:442-897).  The quadratic interior branch declares
// new d_p0..d_p2 that shadow the accumulated ones, so only d_pt receives its adjoint.
static void d_closest_point_path(Scene &S, int s, V2 pt, V2 d_cp, const PathInfo &info, V2 &d_pt) {
    V2 p[4];
    int idx[4];
    int k = seg_fetch(S, s, info.base_point_id, info.point_id, p, idx);
    float *dpts = &S.d_points[2 * (size_t)S.pt_off[s]];
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
        aadd2(dpts + 2 * idx[0], d_p0);
        aadd2(dpts + 2 * idx[1], d_p1);
    } else if (k == 1) {
        V2 p0 = p[0], p1 = p[1], p2 = p[2];
        V2 d_p0{0, 0}, d_p1{0, 0}, d_p2{0, 0};
        float t = info.t_root;
        if (t == 0) {
            d_p0 += d_cp;
        } else if (t == 1) {
            d_p2 += d_cp;
        } else {
            float A = vsum((p0 - 2 * p1 + p2) * (p0 - 2 * p1 + p2));
            float B = vsum(3 * (p0 - 2 * p1 + p2) * (-p0 + p1));
            float C = vsum(2 * (-p0 + p1) * (-p0 + p1) + (p0 - 2 * p1 + p2) * (p0 - pt));
            V2 d_p = d_cp;
            float tt = 1 - t;
            float d_tt = 2 * tt * dot(d_p, p0) + 2 * t * dot(d_p, p1);
            float d_t = -d_tt + 2 * tt * dot(d_p, p1) + 2 * t * dot(d_p, p2);
            float poly_deriv_t = 3 * A * t * t + 2 * B * t + C;
            if (std::fabs((double)poly_deriv_t) > 1e-6f) {
                float d_C = -(d_t / poly_deriv_t) * t;
                float d_D = -(d_t / poly_deriv_t);
                d_pt += d_C * (-(p0 - 2 * p1 + p2)) + d_D * (-(-p0 + p1));
            }
        }
        aadd2(dpts + 2 * idx[0], d_p0);
        aadd2(dpts + 2 * idx[1], d_p1);
        aadd2(dpts + 2 * idx[2], d_p2);
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
            if (std::fabs(poly_deriv_t) > 1e-10f) {
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
                // A = 3*sum(a*a)
                d_p0 += dmul(d_A * 3 * (-1) * 2, a);
                d_p1 += dmul(d_A * 3 * 3 * 2, a);
                d_p2 += dmul(d_A * 3 * (-3) * 2, a);
                d_p3 += dmul(d_A * 3 * 1 * 2, a);
                // B = 5*sum(a*b)
                d_p0 += dmul(d_B * 5, (-1) * b + 3 * a);
                d_p1 += dmul(d_B * 5, 3 * b + (-6) * a);
                d_p2 += dmul(d_B * 5, (-3) * b + 3 * a);
                d_p3 += dmul(d_B * 5, b);
                // C = 4*sum(a*c) + 2*sum(b*b)
                d_p0 += dmul(d_C * 4, (-1) * c + (-3) * a) + dmul(d_C * 2, 3 * 2 * b);
                d_p1 += dmul(d_C * 4, 3 * c + 3 * a) + dmul(d_C * 2, (-6) * 2 * b);
                d_p2 += dmul(d_C * 4, (-3) * c) + dmul(d_C * 2, 3 * 2 * b);
                d_p3 += dmul(d_C * 4, c);
                // D = 3*(sum(b*c) + sum(a*d))
                d_p0 += dmul(d_D * 3, 3 * c + (-3) * b) + dmul(d_D * 3, (-1) * d + 1 * a);
                d_p1 += dmul(d_D * 3, (-6) * c + (3) * b) + dmul(d_D * 3, 3 * d);
                d_p2 += dmul(d_D * 3, 3 * c) + dmul(d_D * 3, (-3) * d);
                d_pt += dmul(d_D * 3, (-1) * a);
                // E = sum(c*c) + 2*sum(d*b)
                d_p0 += dmul(d_E, (-3) * 2 * c) + dmul(d_E * 2, 1 * b + 3 * d);
                d_p1 += dmul(d_E, 3 * 2 * c) + dmul(d_E * 2, (-6) * d);
                d_p2 += dmul(d_E * 2, 3 * d);
                d_pt += dmul(d_E * 2, (-1) * b);
                // F = sum(d*c)
                d_p0 += dmul(d_F, 1 * c) + dmul(d_F, (-3) * d);
                d_p1 += dmul(d_F, 3 * d);
                d_pt += dmul(d_F, (-1) * c);
            }
        }
        aadd2(dpts + 2 * idx[0], d_p0);
        aadd2(dpts + 2 * idx[1], d_p1);
        aadd2(dpts + 2 * idx[2], d_p2);
        aadd2(dpts + 2 * idx[3], d_p3);
    }
}

// d_closest_point for rect (compute_distance.h:783-873)
static void d_closest_point_rect(Scene &S, int s, V2 pt, V2 d_cp, V2 &d_pt) {
    const float *P = &S.params[S.param_off[s]];
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
    if (td < md) { md = td; min_id = 1; }
    if (rd < md) { md = rd; min_id = 2; }
    if (bd < md) { md = bd; min_id = 3; }
    auto d_update = [&](V2 p0, V2 p1, V2 &d_p0, V2 &d_p1) {
        float t = dot(pt - p0, p1 - p0) / dot(p1 - p0, p1 - p0);
        if (t < 0) {
            d_p0 += d_cp;
        } else if (t > 1) {
            d_p1 += d_cp;
        } else {
            V2 d_p = d_cp;
            d_p0 += d_p * (1 - t);
            d_p1 += d_p * t;
            float d_t = vsum(d_p * (p1 - p0));
            float d_num = d_t / dot(p1 - p0, p1 - p0);
            float d_den = d_t * (-t) / dot(p1 - p0, p1 - p0);
            d_pt += (p1 - p0) * d_num;
            d_p1 += (pt - p0) * d_num;
            d_p0 += ((p0 - p1) + (p0 - pt)) * d_num;
            d_p1 += 2 * (p1 - p0) * d_den;
            d_p0 += 2 * (p0 - p1) * d_den;
        }
    };
    V2 d_lt{0, 0}, d_rt{0, 0}, d_lb{0, 0}, d_rb{0, 0};
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
    float *dp = &S.d_params[S.param_off[s]];
    aadd2(dp, d_pmin);
    aadd2(dp + 2, d_pmax);
}

// d_closest_point for circle (compute_distance.h:442-452)
static void d_closest_point_circle(Scene &S, int s, V2 pt, V2 d_cp) {
    const float *P = &S.params[S.param_off[s]];
    float *dp = &S.d_params[S.param_off[s]];
    V2 c{P[1], P[2]};
    V2 dn = d_normalize(pt - c, P[0] * d_cp);
    V2 d_center = d_cp * V2{1 + dn.x, 1 + dn.y};
    aadd2(dp + 1, d_center);
    aadd(dp, dot(d_cp, normalize(pt - c)));
}

// d_compute_distance (compute_distance.h:899-949)
static void d_compute_distance(Scene &S, int g, int s, V2 pt, V2 closest_pt, const PathInfo &info, float d_dist,
                               float *d_translation) {
    if (dist_sq(pt, closest_pt) < 1e-10f) return;
    const M3 &c2s = S.canvas_to_shape[g];
    const M3 &s2c = S.shape_to_canvas[g];
    V2 local_pt = xform_pt(c2s, pt);
    V2 local_cp = xform_pt(c2s, closest_pt);
    V2 d_pt{0, 0}, d_cp{0, 0};
    d_distance(closest_pt, pt, d_dist, d_cp, d_pt);
    V2 d_local_cp{0, 0};
    M3 d_s2c = m3_zero();
    d_xform_pt(s2c, local_cp, d_cp, d_s2c, d_local_cp);
    V2 d_local_pt{0, 0};
    switch (S.type[s]) {
        case kCircle:
            d_closest_point_circle(S, s, local_pt, d_local_cp);
            break;
        case kPath:
            d_closest_point_path(S, s, local_pt, d_local_cp, info, d_local_pt);
            break;
        case kRect:
            d_closest_point_rect(S, s, local_pt, d_local_cp, d_local_pt);
            break;
        default:
            break;
    }
    M3 d_c2s = m3_zero();
    d_xform_pt(c2s, pt, d_local_pt, d_c2s, d_pt);
    M3 tc2s = m3_transpose(c2s);
    M3 corr = m3_mul(m3_mul(m3_neg(tc2s), d_c2s), tc2s);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) d_s2c.m[i][j] += corr.m[i][j];
    aaddm(&S.d_xform[9 * (size_t)g], d_s2c);
    if (d_translation) aadd2(d_translation, -d_pt);
}

// ---------------------------------------------------------------- compositing
struct Frag {
    V3 color;
    float alpha;
    int group_id;
    bool is_stroke;
};
static const int kMaxFrags = 256;
static thread_local bool t_overflow = false;
thread_local bool t_q8_slot = false;
// Tallies (while counting) of the samples whose queries reached the interval slot that
// the reference leaves uninitialised: [0] interior samples, [1] contributing edge
// samples; and the contributing edge samples by the shape they were drawn on.
static int64_t g_slot_samples[2];
static std::vector<int64_t> g_slot_edge_by_shape;
static void tally_slot(int phase, int shape_id) {
    if (!g_counting || !t_q8_slot) return;
    __atomic_fetch_add(&g_slot_samples[phase], 1, __ATOMIC_RELAXED);
    if (shape_id >= 0 && shape_id < (int)g_slot_edge_by_shape.size())
        __atomic_fetch_add(&g_slot_edge_by_shape[shape_id], 1, __ATOMIC_RELAXED);
}

// sample_color(scene, ...) (diffvg.cpp:525-707)
static V4 sample_color_scene(Scene &S, const V4 *bg, V2 screen_pt, const V4 *d_color, EdgeQuery *eq, V4 *d_bg,
                             float *d_translation) {
    if (eq) eq->hit = false;
    V2 pt = screen_pt;
    pt.x *= S.cw;
    pt.y *= S.ch;
    Frag frags[kMaxFrags];
    int nf = 0;
    if (S.ng > 0) {
        int stack[64];
        int sp = 0;
        stack[sp++] = 2 * S.ng - 2;
        while (sp > 0) {
            const Node &node = S.top_bvh[stack[--sp]];
            if (node.child1 < 0) {
                int g = node.child0;
                if (S.stroke_type[g] != kNoColor) {
                    if (within_distance_group(S, g, pt, eq)) {
                        V4 ca = sample_paint(S, S.stroke_type[g], S.stroke_off[g], S.stroke_ns[g], pt);
                        if (nf < kMaxFrags)
                            frags[nf++] = Frag{V3{ca.x, ca.y, ca.z}, ca.w, g, true};
                        else
                            t_overflow = true;
                    }
                }
                if (S.fill_type[g] != kNoColor) {
                    if (is_inside(S, g, pt, eq)) {
                        V4 ca = sample_paint(S, S.fill_type[g], S.fill_off[g], S.fill_ns[g], pt);
                        if (nf < kMaxFrags)
                            frags[nf++] = Frag{V3{ca.x, ca.y, ca.z}, ca.w, g, false};
                        else
                            t_overflow = true;
                    }
                }
            } else {
                bump(kCntTopBox, 2);
                const Node &c0 = S.top_bvh[node.child0];
                const Node &c1 = S.top_bvh[node.child1];
                if (inside(c0.box, pt, c0.max_radius)) stack[sp++] = node.child0;
                if (inside(c1.box, pt, c1.max_radius)) stack[sp++] = node.child1;
                if (sp > 64) std::abort();
            }
        }
    }
    bump(kCntFragments, nf);
    if (nf <= 0) {
        if (bg) {
            if (d_bg) *d_bg = *d_color;
            return *bg;
        }
        return V4{0, 0, 0, 0};
    }
    for (int i = 1; i < nf; i++) {  // stable insertion sort by group id (diffvg.cpp:605-615)
        int j = i;
        Frag temp = frags[j];
        while (j > 0 && frags[j - 1].group_id > temp.group_id) {
            frags[j] = frags[j - 1];
            j--;
        }
        frags[j] = temp;
    }
    V3 accum_color[kMaxFrags];
    float accum_alpha[kMaxFrags];
    float first_alpha = 0.f;
    V3 first_color{0, 0, 0};
    if (bg) {
        first_alpha = bg->w;
        first_color = V3{bg->x, bg->y, bg->z};
    }
    for (int i = 0; i < nf; i++) {
        const Frag &f = frags[i];
        float prev_alpha = i > 0 ? accum_alpha[i - 1] : first_alpha;
        V3 prev_color = i > 0 ? accum_color[i - 1] : first_color;
        if (eq) {
            if (f.alpha >= 1.f && eq->hit) eq->hit = false;
            if (eq->shape_group_id == f.group_id) eq->hit = true;
        }
        accum_color[i] = prev_color * (1 - f.alpha) + f.alpha * f.color;
        accum_alpha[i] = prev_alpha * (1 - f.alpha) + f.alpha;
    }
    V3 final_color = accum_color[nf - 1];
    float final_alpha = accum_alpha[nf - 1];
    if (final_alpha > 1e-6f) final_color = vdiv(final_color, final_alpha);
    if (d_color) {
        V3 d_final_color{d_color->x, d_color->y, d_color->z};
        float d_final_alpha = d_color->w;
        V3 d_curr_color = d_final_color;
        float d_curr_alpha = d_final_alpha;
        if (final_alpha > 1e-6f) {
            d_curr_color = vdiv(d_final_color, final_alpha);
            d_curr_alpha -= vsum(d_final_color * final_color) / final_alpha;
        }
        for (int i = nf - 1; i >= 0; i--) {
            float prev_alpha = i > 0 ? accum_alpha[i - 1] : first_alpha;
            V3 prev_color = i > 0 ? accum_color[i - 1] : first_color;
            float d_prev_alpha = d_curr_alpha * (1.f - frags[i].alpha);
            float d_alpha_i = d_curr_alpha * (1.f - prev_alpha);
            d_alpha_i += vsum(d_curr_color * (frags[i].color - prev_color));
            V3 d_prev_color = d_curr_color * (1 - frags[i].alpha);
            V3 d_color_i = d_curr_color * frags[i].alpha;
            int g = frags[i].group_id;
            V4 dci{d_color_i.x, d_color_i.y, d_color_i.z, d_alpha_i};
            if (frags[i].is_stroke)
                d_sample_paint(S, S.stroke_type[g], S.stroke_off[g], S.stroke_ns[g], pt, dci, d_translation);
            else
                d_sample_paint(S, S.fill_type[g], S.fill_off[g], S.fill_ns[g], pt, dci, d_translation);
            d_curr_color = d_prev_color;
            d_curr_alpha = d_prev_alpha;
        }
        if (d_bg) {
            d_bg->x += d_curr_color.x;
            d_bg->y += d_curr_color.y;
            d_bg->z += d_curr_color.z;
            d_bg->w += d_curr_alpha;
        }
    }
    return V4{final_color.x, final_color.y, final_color.z, final_alpha};
}

// sample_distance (diffvg.cpp:709-775)
static float sample_distance(Scene &S, V2 screen_pt, float weight, const float *d_dist, float *d_translation) {
    V2 pt = screen_pt;
    pt.x *= S.cw;
    pt.y *= S.ch;
    int min_group = -1;
    float min_distance = 0.f;
    int min_shape = -1;
    V2 closest{0, 0};
    PathInfo min_info{-1, -1, 0};
    for (int g = S.ng - 1; g >= 0; g--) {
        int s = -1;
        V2 p{0, 0};
        PathInfo li{-1, -1, 0};
        float d = std::numeric_limits<float>::infinity();
        if (compute_distance(S, g, pt, std::numeric_limits<float>::infinity(), &s, &p, &li, &d)) {
            if (min_group == -1 || d < min_distance) {
                min_distance = d;
                min_group = g;
                min_shape = s;
                closest = p;
                min_info = li;
            }
        }
    }
    if (min_group == -1) return min_distance;
    min_distance *= weight;
    bool in = false;
    if (S.fill_type[min_group] != kNoColor) {
        in = is_inside(S, min_group, pt, nullptr);
        if (in) min_distance = -min_distance;
    }
    if (d_dist) {
        float d_abs = in ? -(*d_dist) : (*d_dist);
        d_compute_distance(S, min_group, min_shape, pt, closest, min_info, d_abs, d_translation);
    }
    return min_distance;
}

// gather_d_color (diffvg.cpp:777-815)
static V4 gather_d_color(const Scene &S, const float *d_img, const float *w_img, int W, int H, V2 pt) {
    int x = (int)pt.x, y = (int)pt.y;
    float radius = S.filter_radius;
    int ri = (int)std::ceil((double)radius);
    V4 d_color{0, 0, 0, 0};
    for (int dy = -ri; dy <= ri; dy++) {
        for (int dx = -ri; dx <= ri; dx++) {
            int xx = x + dx, yy = y + dy;
            if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
                float xc = xx + 0.5f, yc = yy + 0.5f;
                float fw = filter_weight(S.filter_type, radius, xc - pt.x, yc - pt.y);
                float ws = w_img[yy * W + xx];
                if (ws > 0) {
                    const float *d = &d_img[4 * (yy * W + xx)];
                    d_color += (fw / ws) * V4{d[0], d[1], d[2], d[3]};
                }
            }
        }
    }
    return d_color;
}

// ---------------------------------------------------------------- prefiltered
struct PFrag {
    V3 color;
    float alpha;
    int group_id;
    bool is_stroke;
    int shape_id;
    float distance;
    V2 closest_pt;
    PathInfo info;
    bool within;
};
static const int kMaxPFrags = 64;

// sample_color_prefiltered (diffvg.cpp:835-1113)
static V4 sample_color_prefiltered(Scene &S, const V4 *bg, V2 screen_pt, const V4 *d_color, V4 *d_bg,
                                   float *d_translation) {
    V2 pt = screen_pt;
    pt.x *= S.cw;
    pt.y *= S.ch;
    PFrag frags[kMaxPFrags];
    int nf = 0;
    const float inf = std::numeric_limits<float>::infinity();
    if (S.ng > 0) {
        int stack[64];
        int sp = 0;
        stack[sp++] = 2 * S.ng - 2;
        while (sp > 0) {
            const Node &node = S.top_bvh[stack[--sp]];
            if (node.child1 < 0) {
                int g = node.child0;
                if (S.stroke_type[g] != kNoColor) {
                    int ms = -1;
                    V2 cp{0, 0};
                    PathInfo li{-1, -1, 0};
                    float d = inf;
                    compute_distance(S, g, pt, inf, &ms, &cp, &li, &d);
                    if (ms >= 0) {
                        float sw = S.stroke_width[ms];
                        float ad = (float)std::fabs((double)d);
                        float w = smoothstep(ad + sw) - smoothstep(ad - sw);
                        if (w > 0) {
                            V4 ca = sample_paint(S, S.stroke_type[g], S.stroke_off[g], S.stroke_ns[g], pt);
                            ca.w *= w;
                            if (nf < kMaxPFrags)
                                frags[nf++] = PFrag{V3{ca.x, ca.y, ca.z}, ca.w, g, true, ms, d, cp, li, true};
                            else
                                t_overflow = true;
                        }
                    } else {
                        t_overflow = true;  // the reference asserts min_shape_id != -1 (diffvg.cpp:868)
                    }
                }
                if (S.fill_type[g] != kNoColor) {
                    int ms = -1;
                    V2 cp{0, 0};
                    PathInfo li{-1, -1, 0};
                    float d = inf;
                    bool found = compute_distance(S, g, pt, 1.f, &ms, &cp, &li, &d);
                    bool in = is_inside(S, g, pt, nullptr);
                    if (found || in) {
                        if (!in) d = -d;
                        float w = smoothstep(d);
                        if (w > 0) {
                            V4 ca = sample_paint(S, S.fill_type[g], S.fill_off[g], S.fill_ns[g], pt);
                            ca.w *= w;
                            if (nf < kMaxPFrags)
                                frags[nf++] = PFrag{V3{ca.x, ca.y, ca.z}, ca.w, g, false, ms, d, cp, li, found};
                            else
                                t_overflow = true;
                        }
                    }
                }
            } else {
                bump(kCntTopBox, 2);
                const Node &c0 = S.top_bvh[node.child0];
                const Node &c1 = S.top_bvh[node.child1];
                if (inside(c0.box, pt, c0.max_radius)) stack[sp++] = node.child0;
                if (inside(c1.box, pt, c1.max_radius)) stack[sp++] = node.child1;
                if (sp > 64) std::abort();
            }
        }
    }
    bump(kCntFragments, nf);
    if (nf <= 0) {
        if (bg) {
            if (d_bg) *d_bg = *d_color;
            return *bg;
        }
        return V4{0, 0, 0, 0};
    }
    for (int i = 1; i < nf; i++) {
        int j = i;
        PFrag temp = frags[j];
        while (j > 0 && frags[j - 1].group_id > temp.group_id) {
            frags[j] = frags[j - 1];
            j--;
        }
        frags[j] = temp;
    }
    V3 accum_color[kMaxPFrags];
    float accum_alpha[kMaxPFrags];
    float first_alpha = 0.f;
    V3 first_color{0, 0, 0};
    if (bg) {
        first_alpha = bg->w;
        first_color = V3{bg->x, bg->y, bg->z};
    }
    for (int i = 0; i < nf; i++) {
        float prev_alpha = i > 0 ? accum_alpha[i - 1] : first_alpha;
        V3 prev_color = i > 0 ? accum_color[i - 1] : first_color;
        accum_color[i] = prev_color * (1 - frags[i].alpha) + frags[i].alpha * frags[i].color;
        accum_alpha[i] = prev_alpha * (1 - frags[i].alpha) + frags[i].alpha;
    }
    V3 final_color = accum_color[nf - 1];
    float final_alpha = accum_alpha[nf - 1];
    if (final_alpha > 1e-6f) final_color = vdiv(final_color, final_alpha);
    if (d_color) {
        V3 d_final_color{d_color->x, d_color->y, d_color->z};
        V3 d_curr_color = d_final_color;
        float d_curr_alpha = d_color->w;
        if (final_alpha > 1e-6f) {
            d_curr_color = vdiv(d_final_color, final_alpha);
            d_curr_alpha -= vsum(d_final_color * final_color) / final_alpha;
        }
        for (int i = nf - 1; i >= 0; i--) {
            float prev_alpha = i > 0 ? accum_alpha[i - 1] : first_alpha;
            V3 prev_color = i > 0 ? accum_color[i - 1] : first_color;
            float d_prev_alpha = d_curr_alpha * (1.f - frags[i].alpha);
            float d_alpha_i = d_curr_alpha * (1.f - prev_alpha);
            d_alpha_i += vsum(d_curr_color * (frags[i].color - prev_color));
            V3 d_prev_color = d_curr_color * (1 - frags[i].alpha);
            V3 d_color_i = d_curr_color * frags[i].alpha;
            int g = frags[i].group_id;
            int sid = frags[i].shape_id;
            if (frags[i].is_stroke) {
                float d = frags[i].distance;
                float sw = S.stroke_width[sid];
                float ad = (float)std::fabs((double)d);
                float apw = ad + sw, amw = ad - sw;
                float w = smoothstep(apw) - smoothstep(amw);
                if (w != 0) {
                    float d_w = w > 0 ? (frags[i].alpha / w) * d_alpha_i : 0.f;
                    d_alpha_i *= w;
                    d_sample_paint(S, S.stroke_type[g], S.stroke_off[g], S.stroke_ns[g], pt,
                                   V4{d_color_i.x, d_color_i.y, d_color_i.z, d_alpha_i}, d_translation);
                    float d_apw = d_smoothstep(apw, d_w);
                    float d_amw = -d_smoothstep(amw, d_w);
                    float d_d = d_apw + d_amw;
                    if (d < 0) d_d = -d_d;
                    float d_sw = d_apw - d_amw;
                    if (std::fabs((double)d_d) > 1e-10f)
                        d_compute_distance(S, g, sid, pt, frags[i].closest_pt, frags[i].info, d_d, d_translation);
                    aadd(&S.d_stroke_width[sid], d_sw);
                }
            } else {
                float d = frags[i].distance;
                float w = smoothstep(d);
                if (w != 0) {
                    float d_w = w > 0 ? (frags[i].alpha / w) * d_alpha_i : 0.f;
                    d_alpha_i *= w;
                    d_sample_paint(S, S.fill_type[g], S.fill_off[g], S.fill_ns[g], pt,
                                   V4{d_color_i.x, d_color_i.y, d_color_i.z, d_alpha_i}, d_translation);
                    float d_d = d_smoothstep(d, d_w);
                    if (d < 0) d_d = -d_d;
                    if (std::fabs((double)d_d) > 1e-10f && frags[i].within)
                        d_compute_distance(S, g, sid, pt, frags[i].closest_pt, frags[i].info, d_d, d_translation);
                }
            }
            d_curr_color = d_prev_color;
            d_curr_alpha = d_prev_alpha;
        }
        if (d_bg) {
            d_bg->x += d_curr_color.x;
            d_bg->y += d_curr_color.y;
            d_bg->z += d_curr_color.z;
            d_bg->w += d_curr_alpha;
        }
    }
    return V4{final_color.x, final_color.y, final_color.z, final_alpha};
}

// ---------------------------------------------------------------- boundary sampling
struct BoundaryData {  // sample_boundary.h:9-18
    int base_point_id, point_id;
    float t;
    bool is_stroke;
};

// sample(cdf...) (cdf.h:5-29)
static int cdf_sample(const float *cdf, int n, float u, float *updated_u = nullptr) {
    int lb = 0;
    int len = n - 1 - lb;
    while (len > 0) {
        int half = len / 2;
        int mid = lb + half;
        if (u < cdf[mid]) {
            len = half;
        } else {
            lb = mid + 1;
            len = len - half - 1;
        }
    }
    lb = lb < 0 ? 0 : (lb > n - 1 ? n - 1 : lb);
    if (updated_u) {
        if (lb > 0)
            *updated_u = (u - cdf[lb - 1]) / (cdf[lb] - cdf[lb - 1]);
        else
            *updated_u = u / cdf[lb];
    }
    return lb;
}

static V2 perturb(V2 ret, float dir, float r, V2 &normal) {
    if (dir != 0.f) {
        ret += dir * r * normal;
        if (dir < 0) normal = -normal;
    }
    return ret;
}

// sample_boundary (sample_boundary.h:20-453)
static V2 sample_boundary(const Scene &S, int g, int s, float t, V2 &normal, float &pdf, BoundaryData &data) {
    pdf = 1;
    bool stroke_perturb = false;
    if (S.fill_type[g] != kNoColor && S.stroke_type[g] != kNoColor) {
        if (t < 0.5f) {
            stroke_perturb = false;
            t = 2 * t;
            pdf = 0.5f;
        } else {
            stroke_perturb = true;
            t = 2 * (t - 0.5f);
            pdf = 0.5f;
        }
    } else if (S.stroke_type[g] != kNoColor) {
        stroke_perturb = true;
    }
    data.is_stroke = stroke_perturb;
    float dir = 0.f;
    if (stroke_perturb) {
        if (t < 0.5f) {
            dir = -1.f;
            t = 2 * t;
            pdf *= 0.5f;
        } else {
            dir = 1.f;
            t = 2 * (t - 0.5f);
            pdf *= 0.5f;
        }
    }
    float sr = S.stroke_width[s];
    const float *P = S.param_off[s] >= 0 ? &S.params[S.param_off[s]] : nullptr;
    const float tp = 2 * (float)kPi;
    switch (S.type[s]) {
        case kCircle: {
            V2 offset{(float)((double)P[0] * std::cos((double)(tp * t))),
                      (float)((double)P[0] * std::sin((double)(tp * t)))};
            normal = normalize(offset);
            pdf /= (tp * P[0]);
            V2 ret = V2{P[1], P[2]} + offset;
            return perturb(ret, dir, sr, normal);
        }
        case kEllipse: {
            V2 offset{(float)((double)P[0] * std::cos((double)(tp * t))),
                      (float)((double)P[1] * std::sin((double)(tp * t)))};
            double dxdt = (double)-P[0] * std::sin((double)(tp * t)) * 2 * (double)(float)kPi;
            double dydt = (double)P[1] * std::cos((double)(tp * t)) * 2 * (double)(float)kPi;
            normal = normalize(V2{(float)dydt, (float)-dxdt});
            pdf = (float)((double)pdf / std::sqrt(dxdt * dxdt + dydt * dydt));
            V2 ret = V2{P[2], P[3]} + offset;
            return perturb(ret, dir, sr, normal);
        }
        case kRect: {
            float w = P[2] - P[0], h = P[3] - P[1];
            pdf /= (2 * (w + h));
            if (t <= w / (w + h)) {
                t *= (w + h) / w;
                if (t < 0.5f) {
                    normal = V2{0, -1};
                    V2 ret = V2{P[0], P[1]} + 2 * t * V2{P[2] - P[0], 0.f};
                    return perturb(ret, dir, sr, normal);
                } else {
                    normal = V2{0, 1};
                    V2 ret = V2{P[0], P[3]} + 2 * (t - 0.5f) * V2{P[2] - P[0], 0.f};
                    return perturb(ret, dir, sr, normal);
                }
            } else {
                t = (t - w / (w + h)) * (w + h) / h;
                if (t < 0.5f) {
                    normal = V2{-1, 0};
                    V2 ret = V2{P[0], P[1]} + 2 * t * V2{0.f, P[3] - P[1]};
                    return perturb(ret, dir, sr, normal);
                } else {
                    normal = V2{1, 0};
                    V2 ret = V2{P[2], P[1]} + 2 * (t - 0.5f) * V2{0.f, P[3] - P[1]};
                    return perturb(ret, dir, sr, normal);
                }
            }
        }
        default:
            break;
    }
    // path (sample_boundary.h:80-307)
    bool thick = S.has_thick[s];
    int np = S.npts[s];
    if (dir != 0.f && !S.closed[s]) {
        float cap_length;
        if (thick)
            cap_length = (float)kPi * (S.thick(s, 0) + S.thick(s, np - 1));
        else
            cap_length = 2 * (float)kPi * sr;
        float cap_prob = cap_length / (cap_length + S.shape_length[s]);
        if (t < cap_prob) {
            t = t / cap_prob;
            pdf *= cap_prob;
            float r0 = sr, r1 = sr;
            if (thick) {
                r0 = S.thick(s, 0);
                r1 = S.thick(s, np - 1);
            }
            if (dir < 0) {
                V2 p0 = S.pt(s, 0);
                V2 offset{(float)((double)r0 * std::cos((double)(tp * t))),
                          (float)((double)r0 * std::sin((double)(tp * t)))};
                normal = normalize(offset);
                pdf /= (tp * r0);
                data.base_point_id = 0;
                data.point_id = 0;
                data.t = 0;
                return p0 + offset;
            } else {
                V2 p0 = S.pt(s, np - 1);
                V2 offset{(float)((double)r1 * std::cos((double)(tp * t))),
                          (float)((double)r1 * std::sin((double)(tp * t)))};
                normal = normalize(offset);
                pdf /= (tp * r1);
                data.base_point_id = S.nbase[s] - 1;
                data.point_id = np - 2 - S.ncp[S.ctrl_off[s] + data.base_point_id];
                data.t = 1;
                return p0 + offset;
            }
        } else {
            t = (t - cap_prob) / (1 - cap_prob);
            pdf *= (1 - cap_prob);
        }
    }
    const float *cdf = &S.path_cdf[S.ctrl_off[s]];
    const float *pmf = &S.path_pmf[S.ctrl_off[s]];
    int sample_id = cdf_sample(cdf, S.nbase[s], t, &t);
    int point_id = S.path_point_id_map[S.ctrl_off[s] + sample_id];
    V2 p[4];
    int idx[4];
    int k = seg_fetch(S, s, sample_id, point_id, p, idx);
    data.base_point_id = sample_id;
    data.point_id = point_id;
    data.t = t;
    if (t < -1e-3f || t > 1 + 1e-3f) {
        pdf = 0;
        return V2{0, 0};
    }
    V2 tangent, ret;
    float rr;
    if (k == 0) {
        tangent = p[1] - p[0];
        ret = p[0] + t * (p[1] - p[0]);
        float r0 = thick ? S.thick(s, idx[0]) : sr, r1 = thick ? S.thick(s, idx[1]) : sr;
        rr = r0 + t * (r1 - r0);
    } else if (k == 1) {
        tangent = 2 * (1 - t) * (p[1] - p[0]) + 2 * t * (p[2] - p[1]);
        ret = qeval(p[0], p[1], p[2], t);
        float r0 = thick ? S.thick(s, idx[0]) : sr, r1 = thick ? S.thick(s, idx[1]) : sr,
              r2 = thick ? S.thick(s, idx[2]) : sr;
        float tt = 1 - t;
        rr = (tt * tt) * r0 + (2 * tt * t) * r1 + (t * t) * r2;
    } else {
        float omt = 1 - t;
        tangent = 3 * (omt * omt) * (p[1] - p[0]) + 6 * (1 - t) * t * (p[2] - p[1]) + 3 * t * t * (p[3] - p[2]);
        ret = ceval(p[0], p[1], p[2], p[3], t);
        float r0 = thick ? S.thick(s, idx[0]) : sr, r1 = thick ? S.thick(s, idx[1]) : sr,
              r2 = thick ? S.thick(s, idx[2]) : sr, r3 = thick ? S.thick(s, idx[3]) : sr;
        float tt = 1 - t;
        rr = (tt * tt * tt) * r0 + (3 * tt * tt * t) * r1 + (3 * tt * t * t) * r2 + (t * t * t) * r3;
    }
    float tan_len = len(tangent);
    if (tan_len == 0) {
        pdf = 0;
        return V2{0, 0};
    }
    normal = vdiv(V2{-tangent.y, tangent.x}, tan_len);
    pdf *= pmf[sample_id] / tan_len;
    if (dir != 0.f) {
        ret += dir * rr * normal;
        if (dir < 0) normal = -normal;
    }
    return ret;
}

struct BSample {  // BoundarySample (diffvg.cpp:1314-1323)
    V2 pt, local_pt, normal;
    int group_id, shape_id;
    float t;
    BoundaryData data;
    float pdf;
};

// accumulate_boundary_gradient (diffvg.cpp:89-274)
static void accumulate_boundary_gradient(Scene &S, int g, int s, float contrib, float t, V2 normal,
                                         const BoundaryData &bd, V2 local_boundary_pt) {
    if (bd.is_stroke) {
        bool has_thickness = S.type[s] == kPath && S.has_thick[s];
        if (has_thickness) {
            V2 p[4];
            int idx[4];
            int k = seg_fetch(S, s, bd.base_point_id, bd.point_id, p, idx);
            float tt = bd.t;
            float *dth = &S.d_thickness[S.pt_off[s]];
            if (k == 0) {
                aadd(dth + idx[0], (1 - tt) * contrib);
                aadd(dth + idx[1], (tt)*contrib);
            } else if (k == 1) {
                aadd(dth + idx[0], ((1 - tt) * (1 - tt)) * contrib);
                aadd(dth + idx[1], (2 * (1 - tt) * tt) * contrib);
                aadd(dth + idx[2], (tt * tt) * contrib);
            } else {
                float o = 1 - tt;
                aadd(dth + idx[0], (o * o * o) * contrib);
                aadd(dth + idx[1], 3 * (o * o) * tt * contrib);
                aadd(dth + idx[2], 3 * (1 - tt) * tt * tt * contrib);
                aadd(dth + idx[3], tt * tt * tt * contrib);
            }
        } else {
            aadd(&S.d_stroke_width[s], contrib);
        }
    }
    float *dp = S.param_off[s] >= 0 ? &S.d_params[S.param_off[s]] : nullptr;
    switch (S.type[s]) {
        case kCircle:
            aadd2(dp + 1, normal * contrib);
            aadd(dp, contrib);
            break;
        case kEllipse: {
            const float tp = 2 * (float)kPi;
            aadd2(dp + 2, normal * contrib);
            aadd(dp + 0, (float)(std::cos((double)(tp * t)) * normal.x * contrib));
            aadd(dp + 1, (float)(std::sin((double)(tp * t)) * normal.y * contrib));
            break;
        }
        case kPath: {
            V2 p[4];
            int idx[4];
            int k = seg_fetch(S, s, bd.base_point_id, bd.point_id, p, idx);
            float tt = bd.t;
            float *dpts = &S.d_points[2 * (size_t)S.pt_off[s]];
            float w[4];
            if (k == 0) {
                w[0] = 1 - tt;
                w[1] = tt;
            } else if (k == 1) {
                w[0] = (1 - tt) * (1 - tt);
                w[1] = 2 * (1 - tt) * tt;
                w[2] = tt * tt;
            } else {
                float o = 1 - tt;
                w[0] = o * o * o;
                w[1] = 3 * (o * o) * tt;
                w[2] = 3 * (1 - tt) * tt * tt;
                w[3] = tt * tt * tt;
            }
            for (int j = 0; j <= k + 1; j++) {
                aadd(dpts + 2 * idx[j] + 0, w[j] * normal.x * contrib);
                aadd(dpts + 2 * idx[j] + 1, w[j] * normal.y * contrib);
            }
            break;
        }
        case kRect:
            if (normal == V2{-1, 0})
                aadd(dp + 0, -contrib);
            else if (normal == V2{1, 0})
                aadd(dp + 2, contrib);
            else if (normal == V2{0, -1})
                aadd(dp + 1, -contrib);
            else if (normal == V2{0, 1})
                aadd(dp + 3, contrib);
            break;
    }
    M3 d_s2c = m3_zero();
    V2 d_local{0, 0};
    d_xform_pt(S.shape_to_canvas[g], local_boundary_pt, normal * contrib, d_s2c, d_local);
    aaddm(&S.d_xform[9 * (size_t)g], d_s2c);
}

// ---------------------------------------------------------------- driver
template <typename F>
static void parallel_for(F f, int64_t count) {
    const int64_t chunk = 256;
    int64_t ntasks = (count + chunk - 1) / chunk;
    int nthreads = g_num_threads > 0 ? g_num_threads : (int)std::thread::hardware_concurrency();
    if (nthreads <= 0) nthreads = 16;
    if (nthreads > ntasks) nthreads = (int)std::max<int64_t>(1, ntasks);
    std::atomic<int64_t> next(0);
    auto worker = [&]() {
        Counters local{};
        if (g_counting) g_tc = &local;
        for (;;) {
            int64_t task = next.fetch_add(1);
            if (task >= ntasks) break;
            int64_t beg = task * chunk, end = std::min(beg + chunk, count);
            for (int64_t i = beg; i < end; i++)
                if (g_stride <= 1 || i % g_stride == 0) f((int)i);
        }
        if (g_counting) {
            std::lock_guard<std::mutex> lk(g_counter_mu);
            for (int i = 0; i < 12; i++) g_counters[g_phase].c[i] += local.c[i];
            g_tc = nullptr;
        }
    };
    g_multi = nthreads > 1;
    if (nthreads == 1) {
        worker();
    } else {
        std::vector<std::thread> th;
        for (int i = 0; i < nthreads - 1; i++) th.emplace_back(worker);
        worker();
        for (auto &t : th) t.join();
    }
    g_multi = false;
}

static int render(Scene &S, const orc_render_args &a) {
    const int W = a.width, H = a.height, nsx = a.num_samples_x, nsy = a.num_samples_y;
    const bool eval_mode = a.eval_positions != nullptr && a.num_eval_positions > 0;
    const bool pre = a.use_prefiltering != 0;
    std::vector<float> weight;
    if (!eval_mode) weight.assign((size_t)W * H, 0.f);
    std::atomic<bool> overflow(false);
    if (S.ns == 0) {
        // An empty scene renders the background (or zeros) and has no gradients.
        if (a.render_image && a.background_image)
            std::memcpy(a.render_image, a.background_image, sizeof(float) * 4 * W * H);
        return 0;
    }
    g_phase = 0;
    if (a.render_image || a.d_render_image || a.render_sdf || a.d_render_sdf) {
        if (!eval_mode) {
            // weight_kernel (diffvg.cpp:1115-1158); use_prefiltering is never set there.
            parallel_for(
                [&](int idx) {
                    Pcg rng = pcg_init(idx, a.seed);
                    int sx = idx % nsx;
                    int sy = (idx / nsx) % nsy;
                    int x = (idx / (nsx * nsy)) % W;
                    int y = idx / (nsx * nsy * W);
                    float rx = pcg_float(rng);
                    float ry = pcg_float(rng);
                    V2 pt{x + ((float)sx + rx) / nsx, y + ((float)sy + ry) / nsy};
                    float radius = S.filter_radius;
                    int ri = (int)std::ceil((double)radius);
                    for (int dy = -ri; dy <= ri; dy++)
                        for (int dx = -ri; dx <= ri; dx++) {
                            int xx = x + dx, yy = y + dy;
                            if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
                                float xc = xx + 0.5f, yc = yy + 0.5f;
                                aadd(&weight[yy * W + xx], filter_weight(S.filter_type, radius, xc - pt.x, yc - pt.y));
                            }
                        }
                },
                (int64_t)W * H * nsx * nsy);
        }
        int64_t num_samples = eval_mode ? a.num_eval_positions : (int64_t)W * H * nsx * nsy;
        // render_kernel (diffvg.cpp:1160-1312)
        parallel_for(
            [&](int idx) {
                bump(kCntSamples);
                t_q8_slot = false;
                V2 pt{0, 0};
                int x = 0, y = 0;
                if (!eval_mode) {
                    Pcg rng = pcg_init(idx, a.seed);
                    int sx = idx % nsx;
                    int sy = (idx / nsx) % nsy;
                    x = (idx / (nsx * nsy)) % W;
                    y = idx / (nsx * nsy * W);
                    float rx = pcg_float(rng);
                    float ry = pcg_float(rng);
                    if (pre) rx = ry = 0.5f;
                    pt = V2{x + ((float)sx + rx) / nsx, y + ((float)sy + ry) / nsy};
                } else {
                    pt = V2{a.eval_positions[2 * idx], a.eval_positions[2 * idx + 1]};
                    x = (int)pt.x;
                    y = (int)pt.y;
                }
                V2 npt = pt;
                npt.x /= W;
                npt.y /= H;
                int num_samples_px = nsx * nsy;
                if (a.render_image || a.d_render_image) {
                    V4 d_color{0, 0, 0, 0};
                    if (a.d_render_image) d_color = gather_d_color(S, a.d_render_image, weight.data(), W, H, pt);
                    const V4 *bg = a.background_image ? (const V4 *)&a.background_image[4 * (y * W + x)] : nullptr;
                    V4 *d_bg = a.d_background_image ? (V4 *)&a.d_background_image[4 * (y * W + x)] : nullptr;
                    float *d_tr = a.d_translation ? &a.d_translation[2 * (y * W + x)] : nullptr;
                    V4 color = pre ? sample_color_prefiltered(S, bg, npt, a.d_render_image ? &d_color : nullptr, d_bg, d_tr)
                                   : sample_color_scene(S, bg, npt, a.d_render_image ? &d_color : nullptr, nullptr, d_bg,
                                                        d_tr);
                    float radius = S.filter_radius;
                    int ri = (int)std::ceil((double)radius);
                    for (int dy = -ri; dy <= ri; dy++)
                        for (int dx = -ri; dx <= ri; dx++) {
                            int xx = x + dx, yy = y + dy;
                            if (xx >= 0 && xx < W && yy >= 0 && yy < H && weight[yy * W + xx] > 0) {
                                float ws = weight[yy * W + xx];
                                float xc = xx + 0.5f, yc = yy + 0.5f;
                                float fw = filter_weight(S.filter_type, radius, xc - pt.x, yc - pt.y);
                                V4 wc = vdiv(fw * color, ws);
                                if (a.render_image) aadd4(&a.render_image[4 * (yy * W + xx)], wc);
                                if (a.d_render_image) {
                                    const float *dpx = &a.d_render_image[4 * (yy * W + xx)];
                                    V4 d_pixel{dpx[0], dpx[1], dpx[2], dpx[3]};
                                    float d_weight = (dot(d_pixel, color) * ws -
                                                      fw * dot(d_pixel, color) * (ws - fw)) /
                                                     (ws * ws);
                                    aadd(&S.d_filter_radius,
                                         d_filter_weight(S.filter_type, radius, xc - pt.x, yc - pt.y, d_weight));
                                }
                            }
                        }
                }
                if (a.render_sdf || a.d_render_sdf) {
                    float d_dist = 0.f;
                    if (a.d_render_sdf) d_dist = eval_mode ? a.d_render_sdf[idx] : a.d_render_sdf[y * W + x];
                    float wgt = eval_mode ? 1.f : 1.f / num_samples_px;
                    float dd = sample_distance(S, npt, wgt, a.d_render_sdf ? &d_dist : nullptr,
                                               a.d_translation ? &a.d_translation[2 * (y * W + x)] : nullptr);
                    if (a.render_sdf) {
                        if (eval_mode)
                            aadd(&a.render_sdf[idx], dd);
                        else
                            aadd(&a.render_sdf[y * W + x], dd);
                    }
                }
                tally_slot(0, -1);
                if (t_overflow) overflow = true;
            },
            num_samples);
    }
    if (!pre && a.d_render_image) {
        g_phase = 1;
        int64_t num_samples = (int64_t)W * H * nsx * nsy;
        std::vector<BSample> bs(num_samples);
        // sample_boundary_kernel (diffvg.cpp:1325-1386)
        parallel_for(
            [&](int idx) {
                bs[idx].pt = V2{0, 0};
                bs[idx].shape_id = -1;
                Pcg rng = pcg_init(idx, a.seed);
                float u = pcg_float(rng);
                int sample_id = cdf_sample(S.shapes_cdf.data(), S.num_total_shapes, u);
                int shape_id = S.sample_shape_id[sample_id];
                int g = S.sample_group_id[sample_id];
                float shape_pmf = S.shapes_pmf[shape_id];  // indexed by shape id, as the reference does
                if (shape_pmf <= 0) return;
                float boundary_pdf = 0.f;
                V2 normal{0, 0};
                float t = pcg_float(rng);
                BoundaryData bd{0, 0, 0.f, false};
                V2 local_pt = sample_boundary(S, g, shape_id, t, normal, boundary_pdf, bd);
                if (boundary_pdf <= 0) return;
                bump(kCntBoundary);
                V2 bpt = xform_pt(S.shape_to_canvas[g], local_pt);
                normal = xform_normal(S.canvas_to_shape[g], normal);
                bpt.x /= S.cw;
                bpt.y /= S.ch;
                bs[idx] = BSample{bpt, local_pt, normal, g, shape_id, t, bd, shape_pmf * boundary_pdf};
            },
            num_samples);
        // render_edge_kernel (diffvg.cpp:1388-1475); the CPU path does not sort.
        parallel_for(
            [&](int idx) {
                const BSample &b = bs[idx];
                if (b.shape_id == -1) return;
                int bx = (int)(b.pt.x * W);
                int by = (int)(b.pt.y * H);
                if (bx < 0 || bx >= W || by < 0 || by >= H) return;
                V2 normal = b.normal;
                EdgeQuery iq{b.group_id, b.shape_id, false};
                EdgeQuery oq{b.group_id, b.shape_id, false};
                t_q8_slot = false;
                const V4 *bg = a.background_image ? (const V4 *)&a.background_image[4 * (by * W + bx)] : nullptr;
                V4 ci = sample_color_scene(S, bg, b.pt - 1e-4f * normal, nullptr, &iq, nullptr, nullptr);
                V4 co = sample_color_scene(S, bg, b.pt + 1e-4f * normal, nullptr, &oq, nullptr, nullptr);
                if (!iq.hit && !oq.hit) return;
                if (!iq.hit) {
                    normal = -normal;
                    std::swap(iq, oq);
                    std::swap(ci, co);
                }
                V2 sbpt = b.pt;
                sbpt.x *= W;
                sbpt.y *= H;
                V4 d_color = gather_d_color(S, a.d_render_image, weight.data(), W, H, sbpt);
                d_color = vdiv(d_color, (float)(S.cw * S.ch));
                float contrib = dot(ci - co, d_color) / b.pdf;
                accumulate_boundary_gradient(S, b.group_id, b.shape_id, contrib, b.t, normal, b.data, b.local_pt);
                tally_slot(1, b.shape_id);
                if (a.d_translation) {
                    aadd(&a.d_translation[2 * (by * W + bx) + 0], normal.x * contrib);
                    aadd(&a.d_translation[2 * (by * W + bx) + 1], normal.y * contrib);
                }
                if (t_overflow) overflow = true;
            },
            num_samples);
    }
    g_phase = 0;
    if (overflow) {
        g_err = "fragment list overflow (more overlapping shapes than the reference supports)";
        return 1;
    }
    return 0;
}

}  // namespace orc

using namespace orc;

struct orc_scene {
    Scene s;
};

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_scene_create(const orc_scene_desc *desc, orc_scene **out) {
    if (!desc || !out) {
        g_err = "null argument";
        return 1;
    }
    auto *h = new orc_scene();
    std::string err;
    if (!build_scene(*desc, h->s, err)) {
        g_err = err;
        delete h;
        return 1;
    }
    *out = h;
    return 0;
}

int orc_scene_destroy(orc_scene *scene) {
    delete scene;
    return 0;
}

int orc_render(orc_scene *scene, const orc_render_args *args) {
    if (!scene || !args) {
        g_err = "null argument";
        return 1;
    }
    t_overflow = false;
    if (g_counting) {
        g_counters[0] = g_counters[1] = Counters{};
        g_slot_samples[0] = g_slot_samples[1] = 0;
        g_slot_edge_by_shape.assign(scene->s.ns, 0);
    }
    return render(scene->s, *args);
}

int orc_scene_zero_grads(orc_scene *scene) {
    Scene &S = scene->s;
    std::fill(S.d_params.begin(), S.d_params.end(), 0.f);
    std::fill(S.d_points.begin(), S.d_points.end(), 0.f);
    std::fill(S.d_thickness.begin(), S.d_thickness.end(), 0.f);
    std::fill(S.d_stroke_width.begin(), S.d_stroke_width.end(), 0.f);
    std::fill(S.d_color.begin(), S.d_color.end(), 0.f);
    std::fill(S.d_xform.begin(), S.d_xform.end(), 0.f);
    S.d_filter_radius = 0.f;
    return 0;
}

int orc_scene_get_grads(const orc_scene *scene, orc_grads *o) {
    const Scene &S = scene->s;
    auto cp = [](float *dst, const std::vector<float> &v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(float));
    };
    cp(o->d_shape_params, S.d_params);
    cp(o->d_points, S.d_points);
    cp(o->d_thickness, S.d_thickness);
    cp(o->d_stroke_width, S.d_stroke_width);
    cp(o->d_color_params, S.d_color);
    cp(o->d_shape_to_canvas, S.d_xform);
    if (o->d_filter_radius) *o->d_filter_radius = S.d_filter_radius;
    return 0;
}

void orc_set_num_threads(int n) { g_num_threads = n; }
int orc_get_num_threads(void) {
    return g_num_threads > 0 ? g_num_threads : (int)std::thread::hardware_concurrency();
}
void orc_set_counting(int on) { g_counting = on != 0; }
void orc_set_sample_stride(int stride) { g_stride = stride < 1 ? 1 : stride; }
int64_t orc_slot_samples(int phase) { return g_slot_samples[phase & 1]; }
int orc_slot_edge_samples_by_shape(int64_t *out, int n) {
    int m = std::min(n, (int)g_slot_edge_by_shape.size());
    for (int i = 0; i < m; i++) out[i] = g_slot_edge_by_shape[i];
    return m;
}
void orc_get_counters(int phase, orc_counters *out) {
    int64_t *o = (int64_t *)out;
    for (int i = 0; i < 12; i++) o[i] = g_counters[phase & 1].c[i];
}

int orc_scene_bvh(const orc_scene *scene, int level, int index, int32_t *children, float *boxes, float *max_radius,
                  int max_nodes) {
    const Scene &S = scene->s;
    const Node *nodes = nullptr;
    int n = 0;
    if (level == 0) {
        nodes = S.top_bvh.data();
        n = (int)S.top_bvh.size();
    } else if (level == 1) {
        if (index < 0 || index >= S.ng) return -1;
        nodes = S.gbvh(index);
        n = 2 * S.g_n[index] - 1;
    } else {
        if (index < 0 || index >= S.ns || S.type[index] != kPath) return -1;
        nodes = S.pbvh(index);
        n = 2 * S.nbase[index] - 1;
    }
    for (int i = 0; i < n && i < max_nodes; i++) {
        if (children) {
            children[2 * i] = nodes[i].child0;
            children[2 * i + 1] = nodes[i].child1;
        }
        if (boxes) {
            boxes[4 * i] = nodes[i].box.minx;
            boxes[4 * i + 1] = nodes[i].box.miny;
            boxes[4 * i + 2] = nodes[i].box.maxx;
            boxes[4 * i + 3] = nodes[i].box.maxy;
        }
        if (max_radius) max_radius[i] = nodes[i].max_radius;
    }
    return n;
}

int orc_scene_sampling_tables(const orc_scene *scene, float *shape_length, float *shapes_cdf, float *shapes_pmf,
                              int32_t *sample_shape_id, int32_t *sample_group_id) {
    const Scene &S = scene->s;
    for (int i = 0; i < S.ns && shape_length; i++) shape_length[i] = S.shape_length[i];
    for (int i = 0; i < S.num_total_shapes; i++) {
        if (shapes_cdf) shapes_cdf[i] = S.shapes_cdf[i];
        if (shapes_pmf) shapes_pmf[i] = S.shapes_pmf[i];
        if (sample_shape_id) sample_shape_id[i] = S.sample_shape_id[i];
        if (sample_group_id) sample_group_id[i] = S.sample_group_id[i];
    }
    return S.num_total_shapes;
}

uint32_t orc_pcg32_sample(int idx, uint64_t seed, float *out2) {
    Pcg r = pcg_init(idx, seed);
    Pcg r2 = r;
    uint32_t first = pcg_next(r2);
    out2[0] = pcg_float(r);
    out2[1] = pcg_float(r);
    return first;
}
int orc_solve_quadratic_f(float a, float b, float c, float *t0, float *t1) {
    return solve_quadratic_f(a, b, c, t0, t1) ? 1 : 0;
}
int orc_solve_cubic_f(float a, float b, float c, float d, float *t) { return solve_cubic_f(a, b, c, d, t); }
int orc_solve_cubic_d(double a, double b, double c, double d, double *t) { return solve_cubic_d(a, b, c, d, t); }
float orc_filter_weight(int type, float radius, float dx, float dy) { return filter_weight(type, radius, dx, dy); }
float orc_d_filter_weight(int type, float radius, float dx, float dy, float d_return) {
    return d_filter_weight(type, radius, dx, dy, d_return);
}
int orc_winding_number(const orc_scene *scene, int g, float x, float y) {
    const Scene &S = scene->s;
    V2 lp = xform_pt(S.canvas_to_shape[g], V2{x, y});
    int wn = 0;
    for (int i = 0; i < S.g_n[g]; i++) wn += winding_shape(S, S.g_ids[S.g_off[g] + i], lp);
    return wn;
}
int orc_is_inside(const orc_scene *scene, int g, float x, float y) {
    return is_inside(scene->s, g, V2{x, y}, nullptr) ? 1 : 0;
}
int orc_within_distance(const orc_scene *scene, int g, float x, float y) {
    return within_distance_group(scene->s, g, V2{x, y}, nullptr) ? 1 : 0;
}
int orc_compute_distance(const orc_scene *scene, int g, float x, float y, float max_radius, float *d, float *cxy,
                         int32_t *shape_id, int32_t *path_info, float *t_root) {
    int s = -1;
    V2 cp{0, 0};
    PathInfo info{-1, -1, 0.f};
    float dd = 0.f;
    bool found = compute_distance(scene->s, g, V2{x, y}, max_radius, &s, &cp, &info, &dd);
    if (d) *d = dd;
    if (cxy) {
        cxy[0] = cp.x;
        cxy[1] = cp.y;
    }
    if (shape_id) *shape_id = s;
    if (path_info) {
        path_info[0] = info.base_point_id;
        path_info[1] = info.point_id;
    }
    if (t_root) *t_root = info.t_root;
    return found ? 1 : 0;
}
void orc_sample_color(const orc_scene *scene, float sx, float sy, float *rgba) {
    V4 c = sample_color_scene(const_cast<Scene &>(scene->s), nullptr, V2{sx, sy}, nullptr, nullptr, nullptr, nullptr);
    rgba[0] = c.x;
    rgba[1] = c.y;
    rgba[2] = c.z;
    rgba[3] = c.w;
}
int orc_sample_boundary(const orc_scene *scene, int idx, uint64_t seed, float *out) {
    const Scene &S = scene->s;
    Pcg rng = pcg_init(idx, seed);
    float u = pcg_float(rng);
    int sample_id = cdf_sample(S.shapes_cdf.data(), S.num_total_shapes, u);
    int shape_id = S.sample_shape_id[sample_id];
    int g = S.sample_group_id[sample_id];
    float t = pcg_float(rng);
    V2 normal{0, 0};
    float pdf = 0.f;
    BoundaryData bd{0, 0, 0.f, false};
    V2 lp = sample_boundary(S, g, shape_id, t, normal, pdf, bd);
    out[0] = lp.x;
    out[1] = lp.y;
    out[2] = normal.x;
    out[3] = normal.y;
    out[4] = pdf * S.shapes_pmf[shape_id];
    out[5] = (float)g;
    out[6] = (float)shape_id;
    out[7] = bd.t;
    return pdf > 0 ? 1 : 0;
}
uint32_t orc_morton2d(float x, float y, int w, int h) { return morton2d(V2{x, y}, w, h); }

}  // extern "C"
