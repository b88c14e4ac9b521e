// dvg_math.h -- device arithmetic that reproduces the reference's float/double mix.
//
// The reference's CPU build evaluates its unqualified libm calls on float arguments
// in double (the C overloads), so sqrt/cos/sin/acos/pow/fabs promote the enclosing
// expression to double before it is stored back into a float.  These helpers make that
// promotion explicit so that, compiled without FMA contraction, every coverage decision
// (winding numbers, stroke tests, sample positions) is taken on the same bits as the
// reference.  Vector division by a scalar multiplies by the float reciprocal
// (reference vector.h:385-407); xform_pt divides (matrix.h:506-512).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dvg_crmath.h"

namespace dvg {

#define DVG_DEV __device__ __forceinline__

constexpr double kPi = 3.14159265358979323846;

struct V2 {
    float x, y;
};
struct V3 {
    float x, y, z;
};
struct V4 {
    float x, y, z, w;
};

DVG_DEV V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
DVG_DEV V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
DVG_DEV V2 operator-(V2 a) { return {-a.x, -a.y}; }
DVG_DEV V2 operator*(float s, V2 a) { return {s * a.x, s * a.y}; }
DVG_DEV V2 operator*(V2 a, float s) { return {a.x * s, a.y * s}; }
DVG_DEV V2 operator*(V2 a, V2 b) { return {a.x * b.x, a.y * b.y}; }
DVG_DEV V2 &operator+=(V2 &a, V2 b) {
    a.x += b.x;
    a.y += b.y;
    return a;
}
DVG_DEV V2 &operator-=(V2 &a, V2 b) {
    a.x -= b.x;
    a.y -= b.y;
    return a;
}
DVG_DEV bool operator==(V2 a, V2 b) { return a.x == b.x && a.y == b.y; }
DVG_DEV V2 vdiv(V2 a, float s) {
    float inv = 1.f / s;
    return a * inv;
}
DVG_DEV V2 vdivv(V2 a, V2 b) { return {a.x / b.x, a.y / b.y}; }
DVG_DEV float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
DVG_DEV float vsum(V2 a) { return a.x + a.y; }
DVG_DEV float len_sq(V2 v) { return v.x * v.x + v.y * v.y; }
DVG_DEV float len(V2 v) { return __builtin_sqrtf(len_sq(v)); }  // == (float)sqrt((double)x)
DVG_DEV float dist(V2 a, V2 b) { return len(b - a); }
DVG_DEV float dist_sq(V2 a, V2 b) { return len_sq(b - a); }
DVG_DEV V2 normalize(V2 v) { return vdiv(v, len(v)); }

DVG_DEV V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
DVG_DEV V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
DVG_DEV V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
DVG_DEV V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
DVG_DEV V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
DVG_DEV float vsum(V3 a) { return a.x + a.y + a.z; }
DVG_DEV V3 vdiv(V3 a, float s) {
    float inv = 1.f / s;
    return a * inv;
}

DVG_DEV V4 operator+(V4 a, V4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
DVG_DEV V4 operator-(V4 a, V4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
DVG_DEV V4 operator*(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
DVG_DEV V4 operator*(float s, V4 a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
DVG_DEV V4 operator*(V4 a, V4 b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
DVG_DEV V4 &operator+=(V4 &a, V4 b) {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    a.w += b.w;
    return a;
}
DVG_DEV float vsum(V4 a) { return a.x + a.y + a.z + a.w; }
DVG_DEV float dot(V4 a, V4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
DVG_DEV V4 vdiv(V4 a, float s) {
    float inv = 1.f / s;
    return a * inv;
}

struct V2d {
    double x, y;
};
DVG_DEV V2d dmul(double s, V2 v) { return {s * v.x, s * v.y}; }
DVG_DEV V2d operator+(V2d a, V2d b) { return {a.x + b.x, a.y + b.y}; }
DVG_DEV void operator+=(V2 &a, V2d b) {
    a.x = (float)((double)a.x + b.x);
    a.y = (float)((double)a.y + b.y);
}

// d_length with the reference's double intermediate (vector.h:~470)
DVG_DEV V2 d_length(V2 v, float d_l) {
    float l_sq = len_sq(v);
    double l = __builtin_sqrt((double)l_sq);
    float d_l_sq = (float)((double)(0.5f * d_l) / l);
    return (2 * d_l_sq) * v;
}
DVG_DEV void d_distance(V2 v0, V2 v1, float d_out, V2 &d_v0, V2 &d_v1) {
    V2 d = d_length(v1 - v0, d_out);
    d_v0 -= d;
    d_v1 += d;
}
DVG_DEV V2 d_normalize(V2 v0, V2 d_n) {
    float l = len(v0);
    V2 n = vdiv(v0, l);
    V2 d_v0 = vdiv(d_n, l);
    float d_l = -dot(d_n, n) / l;
    d_v0 += d_length(v0, d_l);
    return d_v0;
}

// 3x3 row-major matrices stored as 9 floats
struct M3 {
    float m[9];
};
DVG_DEV M3 m3_load(const float *p) {
    M3 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.m[i] = p[i];
    return r;
}
DVG_DEV V2 xform_pt(const M3 &m, V2 pt) {
    float t0 = m.m[0] * pt.x + m.m[1] * pt.y + m.m[2];
    float t1 = m.m[3] * pt.x + m.m[4] * pt.y + m.m[5];
    float t2 = m.m[6] * pt.x + m.m[7] * pt.y + m.m[8];
    return {t0 / t2, t1 / t2};
}
DVG_DEV void d_xform_pt(const M3 &m, V2 pt, V2 d_out, M3 &d_m, V2 &d_pt) {
    float t0 = m.m[0] * pt.x + m.m[1] * pt.y + m.m[2];
    float t1 = m.m[3] * pt.x + m.m[4] * pt.y + m.m[5];
    float t2 = m.m[6] * pt.x + m.m[7] * pt.y + m.m[8];
    V2 out = {t0 / t2, t1 / t2};
    float dt0 = d_out.x / t2;
    float dt1 = d_out.y / t2;
    float dt2 = -(d_out.x * out.x + d_out.y * out.y) / t2;
    d_m.m[0] += dt0 * pt.x;
    d_m.m[1] += dt0 * pt.y;
    d_m.m[2] += dt0;
    d_m.m[3] += dt1 * pt.x;
    d_m.m[4] += dt1 * pt.y;
    d_m.m[5] += dt1;
    d_m.m[6] += dt2 * pt.x;
    d_m.m[7] += dt2 * pt.y;
    d_m.m[8] += dt2;
    d_pt.x += dt0 * m.m[0] + dt1 * m.m[3] + dt2 * m.m[6];
    d_pt.y += dt0 * m.m[1] + dt1 * m.m[4] + dt2 * m.m[7];
}
DVG_DEV V2 xform_normal(const M3 &mi, V2 n) {
    return normalize(V2{mi.m[0] * n.x + mi.m[3] * n.y, mi.m[1] * n.x + mi.m[4] * n.y});
}

DVG_DEV float fmin_r(float a, float b) { return a < b ? a : b; }
DVG_DEV float fmax_r(float a, float b) { return a > b ? a : b; }
DVG_DEV bool in_box(float4 b, V2 p) { return p.x >= b.x && p.x <= b.z && p.y >= b.y && p.y <= b.w; }
DVG_DEV bool in_box(float4 b, V2 p, float r) {
    return p.x >= b.x - r && p.x <= b.z + r && p.y >= b.y - r && p.y <= b.w + r;
}
DVG_DEV bool ray_box(float4 b, V2 p) {
    if (p.y < b.y || p.y > b.w) return false;
    if (p.x > b.z) return false;
    return true;
}

// ---- PCG32 XSH-RR (reference pcg.h:11-40) ----
struct Pcg {
    uint64_t state, inc;
};
DVG_DEV uint32_t pcg_next(Pcg &r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + (r.inc | 1);
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
DVG_DEV float pcg_float(Pcg &r) { return __uint_as_float((pcg_next(r) >> 9) | 0x3f800000u) - 1.0f; }
DVG_DEV Pcg pcg_init(int idx, uint64_t seed) {
    Pcg s;
    s.state = 0U;
    s.inc = (((uint64_t)idx + 1) << 1u) | 1u;
    pcg_next(s);
    s.state += (0x853c49e6748fea9bULL + seed);
    pcg_next(s);
    return s;
}

// ---- root solvers (reference solve.h:5-59) ----
DVG_DEV bool solve_quadratic_f(float a, float b, float c, float *t0, float *t1) {
    float discrim = b * b - 4 * a * c;
    if (discrim < 0) return false;
    float root_discrim = __builtin_sqrtf(discrim);
    float q = b < 0 ? -0.5f * (b - root_discrim) : -0.5f * (b + root_discrim);
    float r0 = q / a, r1 = c / q;
    if (r0 > r1) {
        float tmp = r0;
        r0 = r1;
        r1 = tmp;
    }
    *t0 = r0;
    *t1 = r1;
    return true;
}
DVG_DEV bool solve_quadratic_d(double a, double b, double c, double *t0, double *t1) {
    double discrim = b * b - 4 * a * c;
    if (discrim < 0) return false;
    double root_discrim = __builtin_sqrt(discrim);
    double q = b < 0 ? -0.5f * (b - root_discrim) : -0.5f * (b + root_discrim);
    double r0 = q / a, r1 = c / q;
    if (r0 > r1) {
        double tmp = r0;
        r0 = r1;
        r1 = tmp;
    }
    *t0 = r0;
    *t1 = r1;
    return true;
}
// solve_cubic<float>: sqrt/acos/cos/pow are evaluated in double, as in the reference.
DVG_DEV int solve_cubic_f(float a, float b, float c, float d, float t[3]) {
    if (fabs((double)a) < 1e-6f) {
        if (solve_quadratic_f(b, c, d, &t[0], &t[1])) return 2;
        return 0;
    }
    b /= a;
    c /= a;
    d /= a;
    float Q = (b * b - 3 * c) / 9.f;
    float R = (2 * b * b * b - 9 * b * c + 27 * d) / 54.f;
    if (R * R < Q * Q * Q) {
        float theta = (float)acos((double)R / sqrt((double)(Q * Q * Q)));
        double sq = sqrt((double)Q);
        double b3 = (double)(b / 3.f);
        t[0] = (float)((double)-2.f * sq * cos((double)(theta / 3.f)) - b3);
        t[1] = (float)((double)-2.f * sq * cos((double)((theta + 2.f * (float)kPi) / 3.f)) - b3);
        t[2] = (float)((double)-2.f * sq * cos((double)((theta - 2.f * (float)kPi) / 3.f)) - b3);
        return 3;
    } else {
        double disc = sqrt((double)(R * R - Q * Q * Q));
        double third = (double)(float)(1. / 3.);
        float A = R > 0 ? (float)-pow((double)R + disc, third) : (float)pow((double)-R + disc, third);
        float B = fabs((double)A) > 1e-6f ? Q / A : 0.f;
        t[0] = (A + B) - b / 3.f;
        return 1;
    }
}
// solve_cubic<double> (solve.h:29-64).  CR selects the correctly rounded acos, cos and
// pow; the default uses the device libm.
template <bool CR = false>
DVG_DEV int solve_cubic_d(double a, double b, double c, double d, double t[3]) {
    if (fabs(a) < 1e-6f) {
        if (solve_quadratic_d(b, c, d, &t[0], &t[1])) return 2;
        return 0;
    }
    b /= a;
    c /= a;
    d /= a;
    double Q = (b * b - 3 * c) / 9.f;
    double R = (2 * b * b * b - 9 * b * c + 27 * d) / 54.f;
    if (R * R < Q * Q * Q) {
        double x = R / sqrt(Q * Q * Q);
        double theta = CR ? cr::acos_cr(x) : acos(x);
        double sq = sqrt(Q);
        double a0 = theta / 3.f, a1 = (theta + 2.f * kPi) / 3.f, a2 = (theta - 2.f * kPi) / 3.f;
        t[0] = -2.f * sq * (CR ? cr::cos_cr(a0) : cos(a0)) - b / 3.f;
        t[1] = -2.f * sq * (CR ? cr::cos_cr(a1) : cos(a1)) - b / 3.f;
        t[2] = -2.f * sq * (CR ? cr::cos_cr(a2) : cos(a2)) - b / 3.f;
        return 3;
    } else {
        double base = R > 0 ? R + sqrt(R * R - Q * Q * Q) : -R + sqrt(R * R - Q * Q * Q);
        double root = CR ? cr::pow_cr(base, 1. / 3.) : pow(base, 1. / 3.);
        double A = R > 0 ? -root : root;
        double B = fabs(A) > 1e-6f ? Q / A : 0.0;
        t[0] = (A + B) - b / 3.0;
        return 1;
    }
}

// Out of line, so that the rarely taken correctly rounded solve does not add to the
// register budget of the kernels that inline the fast one.
static __device__ inline int solve_cubic_d_cr(double a, double b, double c, double d, double *t) {
    return solve_cubic_d<true>(a, b, c, d, t);
}

// ---- pixel filters (reference filter.h:22-106) and the Gaussian extension ----
enum { kBox = 0, kTent = 1, kParabolic = 2, kHann = 3, kGaussian = 4 };

DVG_DEV float filter_weight(int type, float radius, float dx, float dy) {
    if (fabs((double)dx) > radius || fabs((double)dy) > radius) return 0;
    if (type == kBox) {
        float t = 2 * radius;
        return 1.f / (t * t);
    } else if (type == kTent) {
        double fx = (double)radius - fabs((double)dx);
        double fy = (double)radius - fabs((double)dy);
        float r2 = radius * radius;
        return (float)(fx * fy / (double)(r2 * r2));
    } else if (type == kParabolic) {
        float qx = dx / radius, qy = dy / radius;
        return (4.f / 3.f) * (1 - qx * qx) * (4.f / 3.f) * (1 - qy * qy);
    } else if (type == kHann) {
        float ndx = (dx / (2 * radius)) + 0.5f;
        float ndy = (dy / (2 * radius)) + 0.5f;
        float tp = (float)(2 * kPi);
        double v = 0.5f * (1.0 - cos((double)(tp * ndx))) * 0.5f * (1.0 - cos((double)(tp * ndy))) /
                   (double)(radius * radius);
        return (float)v;
    } else {
        // sigma = radius / 3, truncated at |dx|,|dy| <= radius
        double s = (double)radius / 3.0;
        double d2 = (double)dx * dx + (double)dy * dy;
        return (float)(exp(-d2 / (2.0 * s * s)) / (2.0 * kPi * s * s));
    }
}
DVG_DEV float d_filter_weight(int type, float radius, float dx, float dy, float d_return) {
    if (type == kBox) {
        float t = 2 * radius;
        return d_return * (-2) * 2 * radius / (t * t * t);
    } else if (type == kTent) {
        double fx = (double)radius - fabs((double)dx);
        double fy = (double)radius - fabs((double)dy);
        float norm = 1 / (radius * radius);
        double d_fx = (double)d_return * fy * norm;
        double d_fy = (double)d_return * fx * norm;
        double d_norm = (double)d_return * fx * fy;
        return (float)(d_fx + d_fy + (-4) * d_norm / pow((double)radius, 5.0));
    } else if (type == kParabolic) {
        float r3 = radius * radius * radius;
        return -(2 * (dx * dx) + 2 * (dy * dy)) / r3;  // the reference ignores d_return here
    } else if (type == kHann) {
        float ndx = (dx / (2 * radius)) + 0.5f;
        float ndy = (dy / (2 * radius)) + 0.5f;
        float tp = (float)(2 * kPi);
        double fx = 0.5f * (1.0 - cos((double)(tp * ndx)));
        double fy = 0.5f * (1.0 - cos((double)(tp * ndy)));
        float norm = 1 / (radius * radius);
        double d_fx = (double)d_return * fy * norm;
        double d_fy = (double)d_return * fx * norm;
        double d_norm = (double)d_return * fx * fy;
        double d_ndx = d_fx * 0.5f * sin((double)(tp * ndx)) * (double)tp;
        double d_ndy = d_fy * 0.5f * sin((double)(tp * ndy)) * (double)tp;
        float tr = 2 * radius;
        return (float)(d_ndx * (double)(-2 * dx / (tr * tr)) + d_ndy * (double)(-2 * dy / (tr * tr)) +
                       (-2) * d_norm / (double)(radius * radius * radius));
    } else {
        double s = (double)radius / 3.0;
        double d2 = (double)dx * dx + (double)dy * dy;
        double w = exp(-d2 / (2.0 * s * s)) / (2.0 * kPi * s * s);
        double dw_ds = w * (d2 / (s * s * s) - 2.0 / s);
        return (float)((double)d_return * dw_ds / 3.0);
    }
}

DVG_DEV float smoothstep(float d) {
    float t = (d + 1.f) / 2.f;
    t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
    return t * t * (3 - 2 * t);
}
DVG_DEV float d_smoothstep(float d, float d_ret) {
    if (d < -1.f || d > 1.f) return 0.f;
    float t = (d + 1.f) / 2.f;
    float d_t = d_ret * (6 * t - 6 * t * t);
    return d_t / 2.f;
}

// float -> uint32 as the x86-64 reference build converts it (64-bit truncation, low word)
DVG_DEV uint32_t f2u32_x86(float f) {
    if (!(f > -9.2233720368547758e18f && f < 9.2233720368547758e18f)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)f;
}
DVG_DEV uint32_t expand_bits(uint32_t x) {
    uint32_t r = 0;
#pragma unroll
    for (uint32_t i = 0; i < 10; i++) r |= (x & (1u << i)) << i;
    return r;
}

// Hardware float atomics on coarse-grained device memory.
DVG_DEV void aadd(float *p, float v) { unsafeAtomicAdd(p, v); }
DVG_DEV void aadd2(float *p, V2 v) {
    unsafeAtomicAdd(p, v.x);
    unsafeAtomicAdd(p + 1, v.y);
}
DVG_DEV void aadd4(float *p, V4 v) {
    unsafeAtomicAdd(p, v.x);
    unsafeAtomicAdd(p + 1, v.y);
    unsafeAtomicAdd(p + 2, v.z);
    unsafeAtomicAdd(p + 3, v.w);
}

}  // namespace dvg
