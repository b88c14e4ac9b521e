// Fuzz check of the renderer's fast cubic winding path (diffvg_amd/csrc/dvg_winding_fast.h)
// against the oracle's restatement of the reference's cubic winding (oracle/orc_query.h,
// winding_number.h:119-157): whenever the fast path decides, it must return the
// reference's value.  Test infrastructure only.
#include <cstdint>
#include <cstring>
#include <random>

#include "../../diffvg_amd/csrc/dvg_winding_fast.h"
#include "../../oracle/orc_query.h"

namespace orc {
thread_local Counters *g_tc = nullptr;
}

extern "C" {
// mode 0: random curves and points; 1: points on control-point y values (ray through a
// joint); 2: near-tangent rays; 3: large coordinates; 4: nearly flat curves; 5: points in
// the control box (what reaches the test after the ray-box culling); 6: both sides of a
// boundary sample; 7: straight segments written as cubics.
// out[0] = decided, out[1] = mismatches, out[2] = total; first mismatch in bad[0..9]
void wf_fuzz(int mode, int64_t n, uint64_t seed, int64_t *out, float *bad) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<float> U(0.f, 1.f);
    std::normal_distribution<float> N(0.f, 1.f);
    int64_t decided = 0, mism = 0;
    for (int64_t i = 0; i < n; i++) {
        float scale = mode == 3 ? 900.f : 60.f;
        float cx = U(g) * scale, cy = U(g) * scale;
        float sp = mode == 4 ? 0.01f : (0.5f + 20.f * U(g));
        orc::V2 p[4];
        for (int k = 0; k < 4; k++) p[k] = orc::V2{cx + sp * N(g), cy + (mode == 4 ? 1e-3f : sp) * N(g)};
        orc::V2 pt{cx + 2.f * sp * N(g), cy + 1.5f * sp * N(g)};
        if (mode == 1) pt.y = p[(int)(U(g) * 4) & 3].y;
        if (mode == 5) {  // inside the control box, as after the renderer's ray-box test
            float x0 = std::fmin(std::fmin(p[0].x, p[1].x), std::fmin(p[2].x, p[3].x));
            float x1 = std::fmax(std::fmax(p[0].x, p[1].x), std::fmax(p[2].x, p[3].x));
            float y0 = std::fmin(std::fmin(p[0].y, p[1].y), std::fmin(p[2].y, p[3].y));
            float y1 = std::fmax(std::fmax(p[0].y, p[1].y), std::fmax(p[2].y, p[3].y));
            pt = orc::V2{x0 - 0.3f * (x1 - x0) + 1.3f * (x1 - x0) * U(g), y0 + (y1 - y0) * U(g)};
        }
        if (mode == 6) {
            // both sides of a boundary sample, as the edge kernel shades them: a curve
            // point moved 1e-4 of a 512-unit canvas along the normal
            for (int k = 0; k < 4; k++) p[k] = orc::V2{p[k].x * 8.f, p[k].y * 8.f};
            float t = U(g);
            float s = 1.f - t;
            float x = s * s * s * p[0].x + 3 * s * s * t * p[1].x + 3 * s * t * t * p[2].x + t * t * t * p[3].x;
            float y = s * s * s * p[0].y + 3 * s * s * t * p[1].y + 3 * s * t * t * p[2].y + t * t * t * p[3].y;
            float dx = 3 * s * s * (p[1].x - p[0].x) + 6 * s * t * (p[2].x - p[1].x) + 3 * t * t * (p[3].x - p[2].x);
            float dy = 3 * s * s * (p[1].y - p[0].y) + 6 * s * t * (p[2].y - p[1].y) + 3 * t * t * (p[3].y - p[2].y);
            float l = std::sqrt(dx * dx + dy * dy) + 1e-30f;
            float side = U(g) < 0.5f ? 1.f : -1.f;
            pt = orc::V2{x - side * 0.0512f * dy / l, y + side * 0.0512f * dx / l};
        }
        if (mode == 7) {
            // a straight segment written as a cubic (control points at thirds), as SVG
            // writers emit lines inside curve paths; the point anywhere near its box
            for (int k = 1; k < 3; k++)
                p[k] = orc::V2{p[0].x + (p[3].x - p[0].x) * (float)k / 3.f, p[0].y + (p[3].y - p[0].y) * (float)k / 3.f};
            pt = orc::V2{std::fmin(p[0].x, p[3].x) - sp + (std::fabs(p[3].x - p[0].x) + 2.f * sp) * U(g),
                         std::fmin(p[0].y, p[3].y) + std::fabs(p[3].y - p[0].y) * U(g)};
        }
        if (mode == 2) {
            // the y of an interior extremum of the curve, nudged by a few ulps
            float d0 = p[1].y - p[0].y, d1 = p[2].y - p[1].y, d2 = p[3].y - p[2].y;
            float a = d0 - 2 * d1 + d2, b = 2 * (d1 - d0), c = d0;
            float disc = b * b - 4 * a * c;
            if (a != 0.f && disc >= 0.f) {
                float t = (-b + std::sqrt(disc)) / (2 * a);
                if (t > 0.f && t < 1.f) {
                    float y = dvg::wf::bern3(p[0].y, p[1].y, p[2].y, p[3].y, t);
                    pt.y = y + (float)((int)(U(g) * 9) - 4) * 1e-6f * (std::fabs(y) + 1.f);
                }
            }
        }
        int wf = 0;
        if (!dvg::wf::cubic_winding_fast(p[0].x, p[0].y, p[1].x, p[1].y, p[2].x, p[2].y, p[3].x, p[3].y, pt.x, pt.y,
                                         wf))
            continue;
        decided++;
        int wr = orc::segment_winding(2, p, pt);
        if (wr != wf) {
            if (mism == 0) {
                for (int k = 0; k < 4; k++) {
                    bad[2 * k] = p[k].x;
                    bad[2 * k + 1] = p[k].y;
                }
                bad[8] = pt.x;
                bad[9] = pt.y;
            }
            mism++;
        }
    }
    out[0] = decided;
    out[1] = mism;
    out[2] = n;
}
}

#include "../../oracle/oracle.h"

extern "C" {
// Stroke coverage: whenever the fast rejection fires, the oracle's within-distance test
// of a one-segment open path with that stroke width must reject.  Mode 3 checks the
// closest-point use of the same test: whenever it fires for radius r, the oracle's
// closest point of the segment (compute_distance with max_radius r) must find nothing
// closer than r, so a search that skips the segment keeps its result.  out as wf_fuzz.
void wf_within_fuzz(int mode, int64_t n, uint64_t seed, int64_t *out, float *bad) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<float> U(0.f, 1.f);
    std::normal_distribution<float> N(0.f, 1.f);
    int64_t decided = 0, mism = 0;
    for (int64_t i = 0; i < n; i++) {
        float cx = 100.f + U(g) * 300.f, cy = 100.f + U(g) * 300.f;
        float sp = 2.f + 20.f * U(g);
        float pts[8];
        for (int k = 0; k < 4; k++) {
            pts[2 * k] = cx + sp * N(g);
            pts[2 * k + 1] = cy + sp * N(g);
        }
        float r = mode == 1 ? 0.05f + 0.2f * U(g) : 0.5f + 3.5f * U(g);
        // mode 2: per-point thickness of either sign (the radius the test sees is its magnitude)
        float th[4];
        float rmax = fabsf(r);
        for (int k = 0; k < 4; k++) {
            th[k] = mode == 2 ? (U(g) * 8.f - 4.f) : r;
            rmax = mode == 2 ? (k == 0 ? fabsf(th[k]) : fmaxf(rmax, fabsf(th[k]))) : rmax;
        }
        float ptx = cx + 1.5f * sp * N(g), pty = cy + 1.5f * sp * N(g);
        if (!dvg::wf::cubic_certainly_farther(pts[0], pts[1], pts[2], pts[3], pts[4], pts[5], pts[6], pts[7], ptx,
                                              pty, rmax))
            continue;
        decided++;
        int32_t st = 2, poff = -1, ctrl_off = 0, nb = 1, pto = 0, npt = 4, ncp = 2, goff = 0, gn = 1, gid = 0;
        uint8_t no = 0;
        int32_t ft = -1, fo = 0, fns = 0, sty = 0, so = 0, sns = 0;
        float col[4] = {1, 1, 1, 1}, xf[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, sw = r, zero = 0.f;
        orc_scene_desc d;
        std::memset(&d, 0, sizeof(d));
        d.canvas_width = d.canvas_height = 512;
        d.num_shapes = 1;
        d.shape_type = &st;
        d.stroke_width = &sw;
        d.shape_param_offset = &poff;
        d.shape_params = &zero;
        d.num_shape_params = 0;
        d.path_ctrl_offset = &ctrl_off;
        d.path_num_base_points = &nb;
        d.path_point_offset = &pto;
        d.path_num_points = &npt;
        uint8_t has_th = mode == 2 ? 1 : 0;
        d.path_has_thickness = &has_th;
        d.path_is_closed = &no;
        d.path_use_distance_approx = &no;
        d.num_control_points = &ncp;
        d.num_ctrl_total = 1;
        d.points = pts;
        d.thickness = mode == 2 ? th : nullptr;
        d.num_points_total = 4;
        d.num_groups = 1;
        d.group_shape_offset = &goff;
        d.group_num_shapes = &gn;
        d.group_shape_ids = &gid;
        d.num_group_shape_ids = 1;
        d.fill_type = &ft;
        d.fill_offset = &fo;
        d.fill_num_stops = &fns;
        d.stroke_type = &sty;
        d.stroke_offset = &so;
        d.stroke_num_stops = &sns;
        d.color_params = col;
        d.num_color_params = 4;
        d.use_even_odd_rule = &no;
        d.shape_to_canvas = xf;
        d.filter_type = 0;
        d.filter_radius = 0.5f;
        orc_scene *sc = nullptr;
        if (orc_scene_create(&d, &sc) != 0) continue;
        int within;
        if (mode == 3) {
            float dd = 0.f, cxy[2];
            int32_t sid, pinfo[2];
            float tr;
            within = orc_compute_distance(sc, 0, ptx, pty, r, &dd, cxy, &sid, pinfo, &tr);
        } else {
            within = orc_within_distance(sc, 0, ptx, pty);
        }
        orc_scene_destroy(sc);
        if (within) {
            if (mism == 0) {
                for (int k = 0; k < 8; k++) bad[k] = pts[k];
                bad[8] = ptx;
                bad[9] = pty;
            }
            mism++;
        }
    }
    out[0] = decided;
    out[1] = mism;
    out[2] = n;
}
}
