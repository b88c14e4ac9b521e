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
// the control box (what reaches the test after the ray-box culling).
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
