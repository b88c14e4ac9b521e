// dvg_build.h -- the per-shape and per-group arithmetic of scene construction, shared by
// the host builder (dvg_host.cpp, dvg_scene_create) and the device rebuild
// (dvg_refit.hip, dvg_scene_update), so that both derive bit-identical tables from the
// same parameters.  Every formula follows scene.cpp; with floating-point contraction off
// (the library is built with -ffp-contract=off) host and device round each operation
// the same way.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DVG_BLD __host__ __device__ inline
#else
#define DVG_BLD inline
#endif

namespace dvg {
namespace bld {

struct P2 {
    float x, y;
};
DVG_BLD P2 add(P2 a, P2 b) { return {a.x + b.x, a.y + b.y}; }
DVG_BLD P2 sub(P2 a, P2 b) { return {a.x - b.x, a.y - b.y}; }
DVG_BLD P2 scale(float s, P2 a) { return {s * a.x, s * a.y}; }
DVG_BLD float length(P2 v) { return (float)sqrt((double)(v.x * v.x + v.y * v.y)); }
DVG_BLD float distance(P2 a, P2 b) { return length(sub(b, a)); }

struct Box {
    float minx, miny, maxx, maxy;
};
DVG_BLD Box box_empty() { return {INFINITY, INFINITY, -INFINITY, -INFINITY}; }
DVG_BLD float mn(float a, float b) { return a < b ? a : b; }
DVG_BLD float mx(float a, float b) { return a > b ? a : b; }
DVG_BLD Box merge(const Box &b, P2 p) { return {mn(p.x, b.minx), mn(p.y, b.miny), mx(p.x, b.maxx), mx(p.y, b.maxy)}; }
DVG_BLD Box merge(const Box &a, const Box &b) {
    return {mn(a.minx, b.minx), mn(a.miny, b.miny), mx(a.maxx, b.maxx), mx(a.maxy, b.maxy)};
}

DVG_BLD P2 qeval(P2 p0, P2 p1, P2 p2, float t) {
    float tt = 1 - t;
    return add(add(scale(tt * tt, p0), scale(2 * tt * t, p1)), scale(t * t, p2));
}
DVG_BLD P2 ceval(P2 p0, P2 p1, P2 p2, P2 p3, float t) {
    float tt = 1 - t;
    return add(add(add(scale(tt * tt * tt, p0), scale(3 * tt * tt * t, p1)), scale(3 * tt * t * t, p2)),
               scale(t * t * t, p3));
}

// compute_shape_length of one path segment (scene.cpp:147-184): the chord length of a
// line, two chords through t = 1/2 of a quadratic, three chords through t = 1/3, 2/3 of
// a cubic
DVG_BLD float segment_length(int kind, const P2 *p) {
    if (kind == 0) return distance(p[1], p[0]);
    if (kind == 1) {
        P2 v1 = qeval(p[0], p[1], p[2], 0.5f);
        return distance(v1, p[0]) + distance(v1, p[2]);
    }
    P2 v1 = ceval(p[0], p[1], p[2], p[3], 1.f / 3.f);
    P2 v2 = ceval(p[0], p[1], p[2], p[3], 2.f / 3.f);
    return distance(v1, p[0]) + distance(v1, v2) + distance(v2, p[3]);
}

// boundary length and local box of a circle (0), ellipse (1) or rect (3) from its
// parameters (scene.cpp:113-146, 496-570)
DVG_BLD float basic_shape(int type, const float *P, Box &box) {
    if (type == 0) {
        box = Box{P[1] - P[0], P[2] - P[0], P[1] + P[0], P[2] + P[0]};
        return 0.f + (float)(2.f * 3.14159265358979323846) * P[0];
    }
    if (type == 1) {
        box = Box{P[2] - P[0], P[3] - P[1], P[2] + P[0], P[3] + P[1]};
        float a = P[0], b = P[1];
        return (float)(0.0 + (double)(float)3.14159265358979323846 *
                                 ((double)(3 * (a + b)) - sqrt((double)((3 * a + b) * (a + 3 * b)))));
    }
    box = Box{P[0], P[1], P[2], P[3]};
    return 0.f + 2 * (P[2] - P[0] + P[3] - P[1]);
}

// 3x3 inverse (matrix.h:211-231)
DVG_BLD void inverse3(const float *m, float *r) {
    float det = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                m[2] * (m[3] * m[7] - m[4] * m[6]);
    float invdet = 1 / det;
    r[0] = (m[4] * m[8] - m[7] * m[5]) * invdet;
    r[1] = (m[2] * m[7] - m[1] * m[8]) * invdet;
    r[2] = (m[1] * m[5] - m[2] * m[4]) * invdet;
    r[3] = (m[5] * m[6] - m[3] * m[8]) * invdet;
    r[4] = (m[0] * m[8] - m[2] * m[6]) * invdet;
    r[5] = (m[3] * m[2] - m[0] * m[5]) * invdet;
    r[6] = (m[3] * m[7] - m[6] * m[4]) * invdet;
    r[7] = (m[6] * m[1] - m[0] * m[7]) * invdet;
    r[8] = (m[0] * m[4] - m[3] * m[1]) * invdet;
}
DVG_BLD P2 xform(const float *m, P2 p) {
    float t0 = m[0] * p.x + m[1] * p.y + m[2];
    float t1 = m[3] * p.x + m[4] * p.y + m[5];
    float t2 = m[6] * p.x + m[7] * p.y + m[8];
    return {t0 / t2, t1 / t2};
}

// morton2D of a box centre (diffvg.h:111-126, scene.cpp:431-437): 10 bits per axis,
// the float-to-unsigned conversion as an x86-64 compiler lowers it (64-bit truncation,
// low word kept)
DVG_BLD uint32_t expand10(uint32_t x) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < 10; i++) r |= (x & (1u << i)) << i;
    return r;
}
DVG_BLD uint32_t f2u32(float f) {
    if (!(f > -9.2233720368547758e18f && f < 9.2233720368547758e18f)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)f;
}
DVG_BLD uint32_t morton_of_box(const Box &b, int w, int h) {
    P2 c = scale(0.5f, add(P2{b.minx, b.miny}, P2{b.maxx, b.maxy}));
    float px = c.x / (float)w, py = c.y / (float)h;
    return (expand10(f2u32(px * 1023)) << 1u) | expand10(f2u32(py * 1023));
}
DVG_BLD float box_ycentre(const Box &b) { return 0.5f * (b.miny + b.maxy); }

// canvas-space box of a local box under shape_to_canvas (scene.cpp:676)
DVG_BLD Box canvas_box(const float *m, const Box &rb) {
    Box tb = box_empty();
    tb = merge(tb, xform(m, P2{rb.minx, rb.miny}));
    tb = merge(tb, xform(m, P2{rb.minx, rb.maxy}));
    tb = merge(tb, xform(m, P2{rb.maxx, rb.miny}));
    tb = merge(tb, xform(m, P2{rb.maxx, rb.maxy}));
    return tb;
}

// shape_to_canvas preserves distance ratios (the SDF culling bound needs it)
DVG_BLD bool similarity(const float *m) {
    return m[6] == 0.f && m[7] == 0.f && m[8] == 1.f &&
           ((m[0] == m[4] && m[1] == -m[3]) || (m[0] == -m[4] && m[1] == m[3]));
}

// group flags for the SDF bound: which member supplies the anchor point is topology
enum { kSdfFindable = 1, kSdfApproxQuad = 2 };

}  // namespace bld
}  // namespace dvg
