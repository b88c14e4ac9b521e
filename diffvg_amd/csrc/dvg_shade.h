// dvg_shade.h -- paints, fragment compositing and boundary sampling on the device.
//
//   sample_paint / d_sample_paint   diffvg.cpp:276-504
//   compositing and its adjoint     diffvg.cpp:525-707 (fragments composited in
//                                   increasing group id, stroke before fill of a group)
//   prefiltered compositing         diffvg.cpp:835-1113
//   sample_boundary                 sample_boundary.h:20-453
//   accumulate_boundary_gradient    diffvg.cpp:89-274
//   gather_d_color                  diffvg.cpp:777-815
// Candidates arrive in increasing group id, so the forward composites each fragment as
// soon as it is found (no fragment list, no sort); the adjoint keeps a per-lane list of
// the fragments and the running accumulation for the reverse sweep.
#pragma once

#include "dvg_query.h"

namespace dvg {

DVG_DEV V4 stop_color(const float *C, int i) { return V4{C[4 * i], C[4 * i + 1], C[4 * i + 2], C[4 * i + 3]}; }

DVG_DEV V4 sample_paint(const DevScene &S, int type, int off, int ns, V2 pt) {
    const float *c = S.color + off;
    if (type == 0) return V4{c[0], c[1], c[2], c[3]};
    const float *offsets = c + 4;
    const float *colors = c + 4 + ns;
    float t;
    if (type == 1) {
        V2 beg{c[0], c[1]}, end{c[2], c[3]};
        t = dot(pt - beg, end - beg) / fmax_r(dot(end - beg, end - beg), 1e-3f);
    } else {
        V2 center{c[0], c[1]}, radius{c[2], c[3]};
        t = len(vdivv(pt - center, radius));
    }
    if (t < offsets[0]) return stop_color(colors, 0);
    for (int i = 0; i < ns - 1; i++) {
        float oc = offsets[i], on = offsets[i + 1];
        if (t >= oc && t < on) {
            float tt = (t - oc) / (on - oc);
            return stop_color(colors, i) * (1 - tt) + stop_color(colors, i + 1) * tt;
        }
    }
    return stop_color(colors, ns - 1);
}

// d_sample_color; the radial branch keeps the reference's missing `return`.
DVG_DEV void d_sample_paint(const DevScene &S, int type, int off, int ns, V2 pt, V4 d_color, float *d_translation) {
    float *dc = S.d_color + off;
    if (type == 0) {
        aadd4(dc, d_color);
        return;
    }
    const float *c = S.color + off;
    const float *offsets = c + 4;
    const float *colors = c + 4 + ns;
    float *d_offsets = dc + 4;
    float *d_colors = dc + 4 + ns;
    float t;
    V2 a{c[0], c[1]}, b{c[2], c[3]};
    V2 normalized_offset{0, 0}, offset{0, 0};
    if (type == 1) {
        t = dot(pt - a, b - a) / fmax_r(dot(b - a, b - a), 1e-3f);
    } else {
        offset = pt - a;
        normalized_offset = vdivv(offset, b);
        t = len(normalized_offset);
    }
    if (t < offsets[0]) {
        aadd4(d_colors, d_color);
        return;
    }
    for (int i = 0; i < ns - 1; i++) {
        float oc = offsets[i], on = offsets[i + 1];
        if (t >= oc && t < on) {
            V4 cc = stop_color(colors, i), cn = stop_color(colors, i + 1);
            float tt = (t - oc) / (on - oc);
            float d_tt = vsum(d_color * (cn - cc));
            float d_on = -d_tt * tt / (on - oc);
            float d_oc = d_tt * ((tt - 1.f) / (on - oc));
            float d_t = d_tt / (on - oc);
            aadd4(d_colors + 4 * i, d_color * (1 - tt));
            aadd4(d_colors + 4 * (i + 1), d_color * tt);
            aadd(d_offsets + i, d_oc);
            aadd(d_offsets + i + 1, d_on);
            if (type == 1) {
                float l = fmax_r(dot(b - a, b - a), 1e-3f);
                V2 d_beg = vdiv(d_t * (-(pt - a) - (b - a)), l);
                V2 d_end = vdiv(d_t * (pt - a), l);
                float d_l = -d_t * t / l;
                if (dot(b - a, b - a) > 1e-3f) {
                    d_beg += 2 * d_l * (a - b);
                    d_end += 2 * d_l * (b - a);
                }
                aadd2(dc, d_beg);
                aadd2(dc + 2, d_end);
                if (d_translation) aadd2(d_translation, d_beg + d_end);
                return;
            }
            V2 d_no = d_length(normalized_offset, d_t);
            V2 d_offset = vdivv(d_no, b);
            V2 d_radius = vdivv((-d_no) * offset, b * b);
            V2 d_center = -d_offset;
            aadd2(dc, d_center);
            aadd2(dc + 2, d_radius);
            if (d_translation) aadd2(d_translation, d_center);
        }
    }
    aadd4(d_colors + 4 * (ns - 1), d_color);
}

// Running "over" compositor (diffvg.cpp:616-648) with the EdgeQuery bookkeeping
// (diffvg.cpp:624-632): `trav_hit` collects hits raised inside the coverage tests, and
// the per-fragment rules are folded so that the result equals the reference's
// sequential update over the sorted fragment list.
struct Compositor {
    V3 color;
    float alpha;
    int nf;
    bool seen_target, opaque_after_target, any_opaque;
    DVG_DEV void init(bool has_bg, V4 bg) {
        color = has_bg ? V3{bg.x, bg.y, bg.z} : V3{0, 0, 0};
        alpha = has_bg ? bg.w : 0.f;
        nf = 0;
        seen_target = opaque_after_target = any_opaque = false;
    }
    DVG_DEV void add(V4 c, int g, int target_group) {
        if (c.w >= 1.f) {
            any_opaque = true;
            if (seen_target) opaque_after_target = true;
        }
        if (g == target_group) {
            seen_target = true;
            opaque_after_target = false;
        }
        color = color * (1 - c.w) + c.w * V3{c.x, c.y, c.z};
        alpha = alpha * (1 - c.w) + c.w;
        nf++;
    }
    DVG_DEV bool edge_hit(bool trav_hit) const { return seen_target ? !opaque_after_target : (trav_hit && !any_opaque); }
    DVG_DEV V4 result(bool has_bg, V4 bg) const {
        if (nf == 0) return has_bg ? bg : V4{0, 0, 0, 0};
        V3 c = color;
        if (alpha > 1e-6f) c = vdiv(c, alpha);
        return V4{c.x, c.y, c.z, alpha};
    }
};

// Fragment record of the adjoint sweep.
struct FragRec {
    V3 color;
    float alpha;
    V3 prev_color;
    float prev_alpha;
    int gid;  // group * 2 + is_stroke
};
constexpr int kMaxFrags = 256;  // the reference's limit (diffvg.cpp:542)

// Reverse compositing sweep (diffvg.cpp:654-705).  Returns the adjoint that reaches the
// background colour.
DVG_DEV V4 composite_adjoint(const DevScene &S, const FragRec *frags, int nf, V3 final_color, float final_alpha,
                             V4 d_color, V2 pt, float *d_translation) {
    V3 d_final_color{d_color.x, d_color.y, d_color.z};
    V3 d_curr_color = d_final_color;
    float d_curr_alpha = d_color.w;
    if (final_alpha > 1e-6f) {
        d_curr_color = vdiv(d_final_color, final_alpha);
        d_curr_alpha -= vsum(d_final_color * final_color) / final_alpha;
    }
    for (int i = nf - 1; i >= 0; i--) {
        const FragRec &f = frags[i];
        float d_prev_alpha = d_curr_alpha * (1.f - f.alpha);
        float d_alpha_i = d_curr_alpha * (1.f - f.prev_alpha);
        d_alpha_i += vsum(d_curr_color * (f.color - f.prev_color));
        V3 d_prev_color = d_curr_color * (1 - f.alpha);
        V3 d_color_i = d_curr_color * f.alpha;
        int g = f.gid >> 1;
        V4 dci{d_color_i.x, d_color_i.y, d_color_i.z, d_alpha_i};
        if (f.gid & 1)
            d_sample_paint(S, S.stroke_type[g], S.stroke_off[g], S.stroke_ns[g], pt, dci, d_translation);
        else
            d_sample_paint(S, S.fill_type[g], S.fill_off[g], S.fill_ns[g], pt, dci, d_translation);
        d_curr_color = d_prev_color;
        d_curr_alpha = d_prev_alpha;
    }
    return V4{d_curr_color.x, d_curr_color.y, d_curr_color.z, d_curr_alpha};
}

// Prefiltered fragment record (diffvg.cpp:514-523)
struct PFragRec {
    V3 color;
    float alpha;
    V3 prev_color;
    float prev_alpha;
    int gid;
    int shape_id;
    float distance;
    V2 closest;
    PathInfo info;
    int within;
};
constexpr int kMaxPFrags = 64;  // diffvg.cpp:847

DVG_DEV V4 composite_adjoint_prefiltered(const DevScene &S, const PFragRec *frags, int nf, V3 final_color,
                                         float final_alpha, V4 d_color, V2 pt, float *d_translation) {
    V3 d_final_color{d_color.x, d_color.y, d_color.z};
    V3 d_curr_color = d_final_color;
    float d_curr_alpha = d_color.w;
    if (final_alpha > 1e-6f) {
        d_curr_color = vdiv(d_final_color, final_alpha);
        d_curr_alpha -= vsum(d_final_color * final_color) / final_alpha;
    }
    for (int i = nf - 1; i >= 0; i--) {
        const PFragRec &f = frags[i];
        float d_prev_alpha = d_curr_alpha * (1.f - f.alpha);
        float d_alpha_i = d_curr_alpha * (1.f - f.prev_alpha);
        d_alpha_i += vsum(d_curr_color * (f.color - f.prev_color));
        V3 d_prev_color = d_curr_color * (1 - f.alpha);
        V3 d_color_i = d_curr_color * f.alpha;
        int g = f.gid >> 1;
        int s = f.shape_id;
        float d = f.distance;
        if (f.gid & 1) {
            float sw = S.stroke_width[s];
            float ad = (float)fabs((double)d);
            float apw = ad + sw, amw = ad - sw;
            float w = smoothstep(apw) - smoothstep(amw);
            if (w != 0) {
                float d_w = w > 0 ? (f.alpha / w) * d_alpha_i : 0.f;
                d_alpha_i *= w;
                d_sample_paint(S, S.stroke_type[g], S.stroke_off[g], S.stroke_ns[g], pt,
                               V4{d_color_i.x, d_color_i.y, d_color_i.z, d_alpha_i}, d_translation);
                float d_apw = d_smoothstep(apw, d_w);
                float d_amw = -d_smoothstep(amw, d_w);
                float d_d = d_apw + d_amw;
                if (d < 0) d_d = -d_d;
                float d_sw = d_apw - d_amw;
                if (fabs((double)d_d) > 1e-10f) d_compute_distance(S, g, s, pt, f.closest, f.info, d_d, d_translation);
                aadd(S.d_stroke_width + s, d_sw);
            }
        } else {
            float w = smoothstep(d);
            if (w != 0) {
                float d_w = w > 0 ? (f.alpha / w) * d_alpha_i : 0.f;
                d_alpha_i *= w;
                d_sample_paint(S, S.fill_type[g], S.fill_off[g], S.fill_ns[g], pt,
                               V4{d_color_i.x, d_color_i.y, d_color_i.z, d_alpha_i}, d_translation);
                float d_d = d_smoothstep(d, d_w);
                if (d < 0) d_d = -d_d;
                if (fabs((double)d_d) > 1e-10f && f.within)
                    d_compute_distance(S, g, s, pt, f.closest, f.info, d_d, d_translation);
            }
        }
        d_curr_color = d_prev_color;
        d_curr_alpha = d_prev_alpha;
    }
    return V4{d_curr_color.x, d_curr_color.y, d_curr_color.z, d_curr_alpha};
}

// gather_d_color (diffvg.cpp:777-815)
DVG_DEV V4 gather_d_color(int ftype, float radius, const float *d_img, const float *w_img, int W, int H, V2 pt) {
    int x = (int)pt.x, y = (int)pt.y;
    int ri = (int)ceil((double)radius);
    V4 d_color{0, 0, 0, 0};
    for (int dy = -ri; dy <= ri; dy++) {
        for (int dx = -ri; dx <= ri; dx++) {
            int xx = x + dx, yy = y + dy;
            if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
                float xc = xx + 0.5f, yc = yy + 0.5f;
                float fw = filter_weight(ftype, radius, xc - pt.x, yc - pt.y);
                float ws = w_img[yy * W + xx];
                if (ws > 0) {
                    float4 d = reinterpret_cast<const float4 *>(d_img)[yy * W + xx];
                    d_color += (fw / ws) * V4{d.x, d.y, d.z, d.w};
                }
            }
        }
    }
    return d_color;
}

// ---------------------------------------------------------------- boundary sampling
struct BoundaryData {
    int base_point_id, point_id;
    float t;
    int is_stroke;
};

// Edge sample record, 64 bytes (BoundarySample, diffvg.cpp:1314-1323)
struct alignas(16) EdgeSample {
    float px, py;          // normalised canvas position
    float lx, ly;          // local boundary point
    float nx, ny;          // canvas normal
    int group, shape;      // shape < 0: invalid sample
    float t;
    float pdf;
    BoundaryData data;
};

DVG_DEV int cdf_sample(const float *cdf, int n, float u, float *updated_u) {
    int lb = 0;
    int len_ = n - 1 - lb;
    while (len_ > 0) {
        int half = len_ / 2;
        int mid = lb + half;
        if (u < cdf[mid]) {
            len_ = half;
        } else {
            lb = mid + 1;
            len_ = len_ - half - 1;
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

DVG_DEV V2 perturb(V2 ret, float dir, float r, V2 &normal) {
    if (dir != 0.f) {
        ret += dir * r * normal;
        if (dir < 0) normal = -normal;
    }
    return ret;
}

DVG_DEV V2 sample_boundary(const DevScene &S, int g, int s, float t, V2 &normal, float &pdf, BoundaryData &data) {
    pdf = 1;
    bool stroke_perturb = false;
    bool has_fill = S.fill_type[g] >= 0, has_stroke = S.stroke_type[g] >= 0;
    if (has_fill && has_stroke) {
        if (t < 0.5f) {
            t = 2 * t;
        } else {
            stroke_perturb = true;
            t = 2 * (t - 0.5f);
        }
        pdf = 0.5f;
    } else if (has_stroke) {
        stroke_perturb = true;
    }
    data.is_stroke = stroke_perturb;
    float dir = 0.f;
    if (stroke_perturb) {
        if (t < 0.5f) {
            dir = -1.f;
            t = 2 * t;
        } else {
            dir = 1.f;
            t = 2 * (t - 0.5f);
        }
        pdf *= 0.5f;
    }
    float sr = S.stroke_width[s];
    int type = S.shape_type[s];
    const float tp = 2 * (float)kPi;
    if (type != 2) {
        const float *P = S.params + S.param_off[s];
        if (type == 0) {
            V2 offset{(float)((double)P[0] * cos((double)(tp * t))), (float)((double)P[0] * sin((double)(tp * t)))};
            normal = normalize(offset);
            pdf /= (tp * P[0]);
            return perturb(V2{P[1], P[2]} + offset, dir, sr, normal);
        } else if (type == 1) {
            V2 offset{(float)((double)P[0] * cos((double)(tp * t))), (float)((double)P[1] * sin((double)(tp * t)))};
            double dxdt = (double)-P[0] * sin((double)(tp * t)) * 2 * (double)(float)kPi;
            double dydt = (double)P[1] * cos((double)(tp * t)) * 2 * (double)(float)kPi;
            normal = normalize(V2{(float)dydt, (float)-dxdt});
            pdf = (float)((double)pdf / sqrt(dxdt * dxdt + dydt * dydt));
            return perturb(V2{P[2], P[3]} + offset, dir, sr, normal);
        }
        float w = P[2] - P[0], h = P[3] - P[1];
        pdf /= (2 * (w + h));
        if (t <= w / (w + h)) {
            t *= (w + h) / w;
            if (t < 0.5f) {
                normal = V2{0, -1};
                return perturb(V2{P[0], P[1]} + 2 * t * V2{P[2] - P[0], 0.f}, dir, sr, normal);
            }
            normal = V2{0, 1};
            return perturb(V2{P[0], P[3]} + 2 * (t - 0.5f) * V2{P[2] - P[0], 0.f}, dir, sr, normal);
        }
        t = (t - w / (w + h)) * (w + h) / h;
        if (t < 0.5f) {
            normal = V2{-1, 0};
            return perturb(V2{P[0], P[1]} + 2 * t * V2{0.f, P[3] - P[1]}, dir, sr, normal);
        }
        normal = V2{1, 0};
        return perturb(V2{P[2], P[1]} + 2 * (t - 0.5f) * V2{0.f, P[3] - P[1]}, dir, sr, normal);
    }
    bool thick = S.has_thick[s] != 0;
    int np = S.npt[s], po = S.pt_off[s];
    int sb = S.seg_begin[s], nb = S.nsegs[s];
    if (dir != 0.f && !S.closed[s]) {
        float cap_length = thick ? (float)kPi * (S.thickness[po] + S.thickness[po + np - 1]) : 2 * (float)kPi * sr;
        float cap_prob = cap_length / (cap_length + S.shape_length[s]);
        if (t < cap_prob) {
            t = t / cap_prob;
            pdf *= cap_prob;
            float r0 = thick ? S.thickness[po] : sr;
            float r1 = thick ? S.thickness[po + np - 1] : sr;
            float rr = dir < 0 ? r0 : r1;
            V2 p0 = dir < 0 ? ld_pt(S, po) : ld_pt(S, po + np - 1);
            V2 offset{(float)((double)rr * cos((double)(tp * t))), (float)((double)rr * sin((double)(tp * t)))};
            normal = normalize(offset);
            pdf /= (tp * rr);
            if (dir < 0) {
                data.base_point_id = 0;
                data.point_id = 0;
                data.t = 0;
            } else {
                data.base_point_id = nb - 1;
                data.point_id = np - 2 - S.seg_kind[sb + nb - 1];
                data.t = 1;
            }
            return p0 + offset;
        }
        t = (t - cap_prob) / (1 - cap_prob);
        pdf *= (1 - cap_prob);
    }
    int sample_id = cdf_sample(S.seg_cdf + sb, nb, t, &t);
    int seg = sb + sample_id;
    V2 p[4];
    int4 idx;
    int k = seg_load(S, seg, p, idx);
    data.base_point_id = sample_id;
    data.point_id = S.seg_pid[seg];
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
        float r0 = thick ? S.thickness[idx.x] : sr, r1 = thick ? S.thickness[idx.y] : sr;
        rr = r0 + t * (r1 - r0);
    } else if (k == 1) {
        tangent = 2 * (1 - t) * (p[1] - p[0]) + 2 * t * (p[2] - p[1]);
        ret = qeval(p[0], p[1], p[2], t);
        float r0 = thick ? S.thickness[idx.x] : sr, r1 = thick ? S.thickness[idx.y] : sr,
              r2 = thick ? S.thickness[idx.z] : sr;
        float tt = 1 - t;
        rr = (tt * tt) * r0 + (2 * tt * t) * r1 + (t * t) * r2;
    } else {
        float omt = 1 - t;
        tangent = 3 * (omt * omt) * (p[1] - p[0]) + 6 * (1 - t) * t * (p[2] - p[1]) + 3 * t * t * (p[3] - p[2]);
        ret = ceval(p[0], p[1], p[2], p[3], t);
        float r0 = thick ? S.thickness[idx.x] : sr, r1 = thick ? S.thickness[idx.y] : sr,
              r2 = thick ? S.thickness[idx.z] : sr, r3 = thick ? S.thickness[idx.w] : sr;
        float tt = 1 - t;
        rr = (tt * tt * tt) * r0 + (3 * tt * tt * t) * r1 + (3 * tt * t * t) * r2 + (t * t * t) * r3;
    }
    float tan_len = len(tangent);
    if (tan_len == 0) {
        pdf = 0;
        return V2{0, 0};
    }
    normal = vdiv(V2{-tangent.y, tangent.x}, tan_len);
    pdf *= S.seg_pmf[seg] / tan_len;
    if (dir != 0.f) {
        ret += dir * rr * normal;
        if (dir < 0) normal = -normal;
    }
    return ret;
}

// sample_boundary_kernel body (diffvg.cpp:1325-1386): fills `es`, returns validity.
DVG_DEV bool make_edge_sample(const DevScene &S, int idx, uint64_t seed, EdgeSample &es) {
    es.shape = -1;
    Pcg rng = pcg_init(idx, seed);
    float u = pcg_float(rng);
    int sample_id = cdf_sample(S.shapes_cdf, S.num_total_shapes, u, nullptr);
    int s = S.sample_shape_id[sample_id];
    int g = S.sample_group_id[sample_id];
    float shape_pmf = S.shapes_pmf[s];  // indexed by shape id, as the reference does
    if (shape_pmf <= 0) return false;
    float t = pcg_float(rng);
    V2 normal{0, 0};
    float pdf = 0.f;
    BoundaryData bd{0, 0, 0.f, 0};
    V2 lp = sample_boundary(S, g, s, t, normal, pdf, bd);
    if (pdf <= 0) return false;
    M3 s2c = m3_load(S.s2c + 9 * g);
    M3 c2s = m3_load(S.c2s + 9 * g);
    V2 bpt = xform_pt(s2c, lp);
    normal = xform_normal(c2s, normal);
    bpt.x /= S.cw;
    bpt.y /= S.ch;
    es.px = bpt.x;
    es.py = bpt.y;
    es.lx = lp.x;
    es.ly = lp.y;
    es.nx = normal.x;
    es.ny = normal.y;
    es.group = g;
    es.shape = s;
    es.t = t;
    es.pdf = shape_pmf * pdf;
    es.data = bd;
    return true;
}

// accumulate_boundary_gradient (diffvg.cpp:89-274)
DVG_DEV void accumulate_boundary_gradient(const DevScene &S, int g, int s, float contrib, float t, V2 normal,
                                          const BoundaryData &bd, V2 local_pt) {
    int type = S.shape_type[s];
    if (bd.is_stroke) {
        if (type == 2 && S.has_thick[s]) {
            int seg = S.seg_begin[s] + bd.base_point_id;
            int4 idx = S.seg_idx[seg];
            int k = S.seg_kind[seg];
            float tt = bd.t;
            if (k == 0) {
                aadd(S.d_thickness + idx.x, (1 - tt) * contrib);
                aadd(S.d_thickness + idx.y, (tt)*contrib);
            } else if (k == 1) {
                aadd(S.d_thickness + idx.x, ((1 - tt) * (1 - tt)) * contrib);
                aadd(S.d_thickness + idx.y, (2 * (1 - tt) * tt) * contrib);
                aadd(S.d_thickness + idx.z, (tt * tt) * contrib);
            } else {
                float o = 1 - tt;
                aadd(S.d_thickness + idx.x, (o * o * o) * contrib);
                aadd(S.d_thickness + idx.y, 3 * (o * o) * tt * contrib);
                aadd(S.d_thickness + idx.z, 3 * (1 - tt) * tt * tt * contrib);
                aadd(S.d_thickness + idx.w, tt * tt * tt * contrib);
            }
        } else {
            aadd(S.d_stroke_width + s, contrib);
        }
    }
    if (type == 0) {
        float *dp = S.d_params + S.param_off[s];
        aadd2(dp + 1, normal * contrib);
        aadd(dp, contrib);
    } else if (type == 1) {
        float *dp = S.d_params + S.param_off[s];
        const float tp = 2 * (float)kPi;
        aadd2(dp + 2, normal * contrib);
        aadd(dp + 0, (float)(cos((double)(tp * t)) * normal.x * contrib));
        aadd(dp + 1, (float)(sin((double)(tp * t)) * normal.y * contrib));
    } else if (type == 2) {
        int seg = S.seg_begin[s] + bd.base_point_id;
        int4 idx = S.seg_idx[seg];
        int k = S.seg_kind[seg];
        float tt = bd.t;
        float w[4];
        int ii[4] = {idx.x, idx.y, idx.z, idx.w};
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
            aadd(S.d_points + 2 * ii[j] + 0, w[j] * normal.x * contrib);
            aadd(S.d_points + 2 * ii[j] + 1, w[j] * normal.y * contrib);
        }
    } else {
        float *dp = S.d_params + S.param_off[s];
        if (normal == V2{-1, 0})
            aadd(dp + 0, -contrib);
        else if (normal == V2{1, 0})
            aadd(dp + 2, contrib);
        else if (normal == V2{0, -1})
            aadd(dp + 1, -contrib);
        else if (normal == V2{0, 1})
            aadd(dp + 3, contrib);
    }
    M3 s2c = m3_load(S.s2c + 9 * g);
    M3 d_s2c;
#pragma unroll
    for (int i = 0; i < 9; i++) d_s2c.m[i] = 0.f;
    V2 d_local{0, 0};
    d_xform_pt(s2c, local_pt, normal * contrib, d_s2c, d_local);
    float *dx = S.d_xform + 9 * g;
#pragma unroll
    for (int i = 0; i < 9; i++) aadd(dx + i, d_s2c.m[i]);
}

// ---- wave-combined gradient scatter for boundary samples ----
// Device-scope float atomics to ordinary device memory are resolved past the XCD's L2,
// so the ~20 atomics per boundary sample dominate the edge kernel's memory traffic.  A
// wave's boundary samples come from one pixel tile and mostly share a few shapes and
// segments: lanes with equal keys add their values in registers first and one lane
// issues the atomics.  Must be called by every lane of the wave (inactive lanes pass
// active = false).
template <int N>
DVG_DEV void wave_combine_add(bool active, int key, const float (&v)[N], float *dst_of_key_leader,
                              int n_used = N) {
    const int lane = threadIdx.x & 63;
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lk = __shfl(key, leader, 64);
        const bool mine = active && key == lk && ((pending >> lane) & 1ull);
        const uint64_t m = __ballot(mine);
        float sum[N];
#pragma unroll
        for (int i = 0; i < N; i++) {
            float x = mine ? v[i] : 0.f;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
            sum[i] = x;
        }
        // the destination of this key, as the leader computed it
        const unsigned long long addr =
            ((unsigned long long)(unsigned)__shfl((int)((unsigned long long)dst_of_key_leader >> 32), leader, 64)
             << 32) |
            (unsigned)__shfl((int)((unsigned long long)dst_of_key_leader & 0xffffffffull), leader, 64);
        const int nu = __shfl(n_used, leader, 64);
        if (lane == leader) {
            float *d = (float *)addr;
#pragma unroll
            for (int i = 0; i < N; i++)
                if (i < nu) aadd(d + i, sum[i]);
        }
        pending &= ~m;
    }
}

// accumulate_boundary_gradient (diffvg.cpp:89-274) for a whole wave: the same values as
// the per-sample version, with the stroke-width, path-point and transform adjoints of
// lanes that share a shape, segment or group added together before the atomics.
DVG_DEV void accumulate_boundary_gradient_wave(const DevScene &S, bool active, int g, int s, float contrib, float t,
                                               V2 normal, const BoundaryData &bd, V2 local_pt) {
    int type = active ? S.shape_type[s] : -1;
    // strokes: per-point thickness goes direct (rare); stroke width combined by shape
    bool sw_active = active && bd.is_stroke && !(type == 2 && S.has_thick[s]);
    if (active && bd.is_stroke && type == 2 && S.has_thick[s]) {
        int seg = S.seg_begin[s] + bd.base_point_id;
        int4 idx = S.seg_idx[seg];
        int k = S.seg_kind[seg];
        float tt = bd.t;
        if (k == 0) {
            aadd(S.d_thickness + idx.x, (1 - tt) * contrib);
            aadd(S.d_thickness + idx.y, (tt)*contrib);
        } else if (k == 1) {
            aadd(S.d_thickness + idx.x, ((1 - tt) * (1 - tt)) * contrib);
            aadd(S.d_thickness + idx.y, (2 * (1 - tt) * tt) * contrib);
            aadd(S.d_thickness + idx.z, (tt * tt) * contrib);
        } else {
            float o = 1 - tt;
            aadd(S.d_thickness + idx.x, (o * o * o) * contrib);
            aadd(S.d_thickness + idx.y, 3 * (o * o) * tt * contrib);
            aadd(S.d_thickness + idx.z, 3 * (1 - tt) * tt * tt * contrib);
            aadd(S.d_thickness + idx.w, tt * tt * tt * contrib);
        }
    }
    {
        float v[1] = {contrib};
        wave_combine_add<1>(sw_active, s, v, S.d_stroke_width + (sw_active ? s : 0));
    }
    // shape parameters of circles, ellipses and rects go direct (few per scene)
    if (active && type == 0) {
        float *dp = S.d_params + S.param_off[s];
        aadd2(dp + 1, normal * contrib);
        aadd(dp, contrib);
    } else if (active && type == 1) {
        float *dp = S.d_params + S.param_off[s];
        const float tp = 2 * (float)kPi;
        aadd2(dp + 2, normal * contrib);
        aadd(dp + 0, (float)(cos((double)(tp * t)) * normal.x * contrib));
        aadd(dp + 1, (float)(sin((double)(tp * t)) * normal.y * contrib));
    } else if (active && type == 3) {
        float *dp = S.d_params + S.param_off[s];
        if (normal == V2{-1, 0})
            aadd(dp + 0, -contrib);
        else if (normal == V2{1, 0})
            aadd(dp + 2, contrib);
        else if (normal == V2{0, -1})
            aadd(dp + 1, -contrib);
        else if (normal == V2{0, 1})
            aadd(dp + 3, contrib);
    }
    // path control points, combined by segment.  A segment's points are not contiguous
    // (the last point of a closed path's last segment wraps to its first), so each of the
    // up to four points is combined and added separately, keyed by segment.
    {
        bool pa = active && type == 2;
        int seg = pa ? S.seg_begin[s] + bd.base_point_id : 0;
        int4 idx = pa ? S.seg_idx[seg] : make_int4(0, 0, 0, 0);
        int k = pa ? S.seg_kind[seg] : 0;
        float tt = bd.t;
        float w[4] = {0.f, 0.f, 0.f, 0.f};
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
        int ii[4] = {idx.x, idx.y, idx.z, idx.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            bool on = pa && j <= k + 1;
            float v[2] = {w[j] * normal.x * contrib, w[j] * normal.y * contrib};
            wave_combine_add<2>(on, seg, v, S.d_points + (on ? 2 * ii[j] : 0));
        }
    }
    // shape_to_canvas adjoint, combined by group
    {
        float m9[9];
        if (active) {
            M3 s2c = m3_load(S.s2c + 9 * g);
            M3 d_s2c;
#pragma unroll
            for (int i = 0; i < 9; i++) d_s2c.m[i] = 0.f;
            V2 d_local{0, 0};
            d_xform_pt(s2c, local_pt, normal * contrib, d_s2c, d_local);
#pragma unroll
            for (int i = 0; i < 9; i++) m9[i] = d_s2c.m[i];
        } else {
#pragma unroll
            for (int i = 0; i < 9; i++) m9[i] = 0.f;
        }
        wave_combine_add<9>(active, g, m9, S.d_xform + (active ? 9 * g : 0));
    }
}

// d_compute_distance for a whole wave (every lane calls it; lanes without a distance
// adjoint pass active = false): the values of the per-lane version, with the point,
// shape-parameter and transform adjoints of lanes that share a segment, shape or group
// added together before the atomics.  A distance field or a prefiltered image of a few
// shapes sends nearly every sample's adjoint to the same dozen addresses.
DVG_DEV void d_compute_distance_wave(const DevScene &S, bool active, int g, int s, V2 pt, V2 closest,
                                     const PathInfo &info, float d_dist, float *d_translation) {
    const bool act = active && !(dist_sq(pt, closest) < 1e-10f);
    DistAdj a;
    if (act) {
        d_compute_distance_vals(S, g, s, pt, closest, info, d_dist, a);
    } else {
        a.seg = -1;
        a.npar = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            a.pts.use[j] = false;
            a.pts.idx[j] = 0;
            a.pts.d[j] = V2{0, 0};
            a.par[j] = 0.f;
        }
#pragma unroll
        for (int i = 0; i < 9; i++) a.xform[i] = 0.f;
        a.d_pt = V2{0, 0};
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const bool on = act && a.pts.use[j];
        float v[2] = {a.pts.d[j].x, a.pts.d[j].y};
        wave_combine_add<2>(on, a.seg, v, S.d_points + (on ? 2 * a.pts.idx[j] : 0));
    }
    {
        // circle: radius, centre; rect: p_min, p_max -- contiguous at the shape's offset
        const bool on = act && a.npar > 0;
        wave_combine_add<4>(on, s, a.par, S.d_params + (on ? S.param_off[s] : 0), a.npar);
    }
    wave_combine_add<9>(act, g, a.xform, S.d_xform + (act ? 9 * g : 0));
    if (act && d_translation) aadd2(d_translation, -a.d_pt);
}

// composite_adjoint_prefiltered for a whole wave, lanes in lockstep by fragment position
// (as composite_adjoint_wave): constant paints, stroke widths and distance adjoints are
// combined across lanes.  Every lane must call it; lanes without fragments pass nf = 0.
DVG_DEV V4 composite_adjoint_prefiltered_wave(const DevScene &S, const PFragRec *frags, int nf, V3 final_color,
                                              float final_alpha, V4 d_color, V2 pt, float *d_translation) {
    V3 d_final_color{d_color.x, d_color.y, d_color.z};
    V3 d_curr_color = d_final_color;
    float d_curr_alpha = d_color.w;
    if (nf > 0 && final_alpha > 1e-6f) {
        d_curr_color = vdiv(d_final_color, final_alpha);
        d_curr_alpha -= vsum(d_final_color * final_color) / final_alpha;
    }
    int mx = nf;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    for (int i = mx - 1; i >= 0; i--) {
        const bool on = i < nf;
        PFragRec f{};
        if (on) f = frags[i];
        int g = 0, s = 0, type = -1, off = 0;
        bool paint_on = false, dist_on = false, sw_on = false;
        float d_d = 0.f, d_sw = 0.f;
        V4 dci{0, 0, 0, 0};
        if (on) {
            float d_prev_alpha = d_curr_alpha * (1.f - f.alpha);
            float d_alpha_i = d_curr_alpha * (1.f - f.prev_alpha);
            d_alpha_i += vsum(d_curr_color * (f.color - f.prev_color));
            V3 d_prev_color = d_curr_color * (1 - f.alpha);
            V3 d_color_i = d_curr_color * f.alpha;
            g = f.gid >> 1;
            s = f.shape_id;
            float d = f.distance;
            int ns = 0;
            if (f.gid & 1) {
                type = S.stroke_type[g];
                off = S.stroke_off[g];
                ns = S.stroke_ns[g];
                float sw = S.stroke_width[s];
                float ad = (float)fabs((double)d);
                float apw = ad + sw, amw = ad - sw;
                float w = smoothstep(apw) - smoothstep(amw);
                if (w != 0) {
                    float d_w = w > 0 ? (f.alpha / w) * d_alpha_i : 0.f;
                    d_alpha_i *= w;
                    paint_on = true;
                    dci = V4{d_color_i.x, d_color_i.y, d_color_i.z, d_alpha_i};
                    float d_apw = d_smoothstep(apw, d_w);
                    float d_amw = -d_smoothstep(amw, d_w);
                    d_d = d_apw + d_amw;
                    if (d < 0) d_d = -d_d;
                    d_sw = d_apw - d_amw;
                    dist_on = fabs((double)d_d) > 1e-10f;
                    sw_on = true;
                }
            } else {
                type = S.fill_type[g];
                off = S.fill_off[g];
                ns = S.fill_ns[g];
                float w = smoothstep(d);
                if (w != 0) {
                    float d_w = w > 0 ? (f.alpha / w) * d_alpha_i : 0.f;
                    d_alpha_i *= w;
                    paint_on = true;
                    dci = V4{d_color_i.x, d_color_i.y, d_color_i.z, d_alpha_i};
                    d_d = d_smoothstep(d, d_w);
                    if (d < 0) d_d = -d_d;
                    dist_on = fabs((double)d_d) > 1e-10f && f.within;
                }
            }
            if (paint_on && type != 0) d_sample_paint(S, type, off, ns, pt, dci, d_translation);
            d_curr_color = d_prev_color;
            d_curr_alpha = d_prev_alpha;
        }
        {
            float v[4] = {dci.x, dci.y, dci.z, dci.w};
            const bool konst = paint_on && type == 0;
            wave_combine_add<4>(konst, on ? f.gid : 0, v, S.d_color + (konst ? off : 0));
        }
        d_compute_distance_wave(S, dist_on, g, s, pt, on ? f.closest : pt, f.info, d_d, d_translation);
        {
            float v[1] = {d_sw};
            wave_combine_add<1>(sw_on, s, v, S.d_stroke_width + (sw_on ? s : 0));
        }
    }
    return V4{d_curr_color.x, d_curr_color.y, d_curr_color.z, d_curr_alpha};
}

// composite_adjoint for a whole wave: lanes walk their fragment lists in lockstep by
// position (each lane still goes from its last fragment to its first), so that lanes of
// a tile that see the same paint at the same depth add their constant-paint adjoints
// together before the atomics.  Every lane of the wave must call it; lanes without
// fragments pass nf = 0 and get d_color back.
DVG_DEV V4 composite_adjoint_wave(const DevScene &S, const FragRec *frags, int nf, V3 final_color, float final_alpha,
                                  V4 d_color, V2 pt, float *d_translation) {
    V3 d_final_color{d_color.x, d_color.y, d_color.z};
    V3 d_curr_color = d_final_color;
    float d_curr_alpha = d_color.w;
    if (nf > 0 && final_alpha > 1e-6f) {
        d_curr_color = vdiv(d_final_color, final_alpha);
        d_curr_alpha -= vsum(d_final_color * final_color) / final_alpha;
    }
    int mx = nf;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    for (int i = mx - 1; i >= 0; i--) {
        const bool on = i < nf;
        FragRec f = on ? frags[i] : FragRec{};
        V4 dci{0, 0, 0, 0};
        int g = 0, type = -1, off = 0, ns = 0;
        if (on) {
            float d_prev_alpha = d_curr_alpha * (1.f - f.alpha);
            float d_alpha_i = d_curr_alpha * (1.f - f.prev_alpha);
            d_alpha_i += vsum(d_curr_color * (f.color - f.prev_color));
            V3 d_prev_color = d_curr_color * (1 - f.alpha);
            V3 d_color_i = d_curr_color * f.alpha;
            g = f.gid >> 1;
            dci = V4{d_color_i.x, d_color_i.y, d_color_i.z, d_alpha_i};
            if (f.gid & 1) {
                type = S.stroke_type[g];
                off = S.stroke_off[g];
                ns = S.stroke_ns[g];
            } else {
                type = S.fill_type[g];
                off = S.fill_off[g];
                ns = S.fill_ns[g];
            }
            if (type != 0) d_sample_paint(S, type, off, ns, pt, dci, d_translation);
            d_curr_color = d_prev_color;
            d_curr_alpha = d_prev_alpha;
        }
        float v[4] = {dci.x, dci.y, dci.z, dci.w};
        const bool konst = on && type == 0;
        wave_combine_add<4>(konst, on ? f.gid : 0, v, S.d_color + (konst ? off : 0));
    }
    return V4{d_curr_color.x, d_curr_color.y, d_curr_color.z, d_curr_alpha};
}

}  // namespace dvg
