// orc_math.h -- scalar/vector/matrix primitives of the oracle (test infrastructure only).
//
// Every primitive reproduces the reference arithmetic bit for bit on a CPU without
// FMA: the reference's unqualified libm calls on float arguments bind to the double
// overloads, so sqrt/cos/sin/acos/pow/fabs results are double and the surrounding
// expression is evaluated in double before it is stored into a float.  The helpers
// below make that promotion explicit.  Vector division by a scalar multiplies by the
// float reciprocal (vector.h:385-407); matrix xform_pt divides (matrix.h:506-512).
#pragma once

#include <cmath>
#include <cstdint>
#include <limits>

namespace orc {

constexpr double kPi = 3.14159265358979323846;  // M_PI as used by the reference (diffvg.h:23)

struct V2 {
    float x, y;
};
struct V3 {
    float x, y, z;
};
struct V4 {
    float x, y, z, w;
};

inline V2 operator+(V2 a, V2 b) { return {a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline V2 operator-(V2 a) { return {-a.x, -a.y}; }
inline V2 operator*(float s, V2 a) { return {s * a.x, s * a.y}; }
inline V2 operator*(V2 a, float s) { return {a.x * s, a.y * s}; }
inline V2 operator*(V2 a, V2 b) { return {a.x * b.x, a.y * b.y}; }
inline V2 &operator+=(V2 &a, V2 b) { a.x += b.x; a.y += b.y; return a; }
inline V2 &operator-=(V2 &a, V2 b) { a.x -= b.x; a.y -= b.y; return a; }
inline bool operator==(V2 a, V2 b) { return a.x == b.x && a.y == b.y; }
// v / s == v * (1.f / s) (vector.h:385-392)
inline V2 vdiv(V2 a, float s) { float inv = 1.f / s; return a * inv; }
inline V2 vdivv(V2 a, V2 b) { return {a.x / b.x, a.y / b.y}; }   // vector.h:425-430
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline float vsum(V2 a) { return a.x + a.y; }
inline float len_sq(V2 v) { return v.x * v.x + v.y * v.y; }
// length(): sqrt(double) rounded back to float == sqrtf (vector.h:~466)
inline float len(V2 v) { return (float)std::sqrt((double)len_sq(v)); }
inline float dist(V2 a, V2 b) { return len(b - a); }        // distance(v0,v1) = length(v1 - v0)
inline float dist_sq(V2 a, V2 b) { return len_sq(b - a); }
inline V2 normalize(V2 v) { return vdiv(v, len(v)); }

inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float vsum(V3 a) { return a.x + a.y + a.z; }
inline V3 vdiv(V3 a, float s) { float inv = 1.f / s; return a * inv; }

inline V4 operator+(V4 a, V4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline V4 operator-(V4 a, V4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
inline V4 operator*(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline V4 operator*(float s, V4 a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
inline V4 operator*(V4 a, V4 b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
inline V4 &operator+=(V4 &a, V4 b) { a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w; return a; }
inline float vsum(V4 a) { return a.x + a.y + a.z + a.w; }
inline float dot(V4 a, V4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
inline V4 vdiv(V4 a, float s) { float inv = 1.f / s; return a * inv; }

// d_length (vector.h:~470): l is a double (auto l = sqrt(l_sq)), d_l_sq is double and
// is cast to float before d_length_squared = 2 * d_l_sq * v0.
inline V2 d_length(V2 v, float d_l) {
    float l_sq = len_sq(v);
    double l = std::sqrt((double)l_sq);
    float d_l_sq = (float)((double)(0.5f * d_l) / l);
    return (2 * d_l_sq) * v;
}
// d_distance (vector.h:~520)
inline void d_distance(V2 v0, V2 v1, float d_out, V2 &d_v0, V2 &d_v1) {
    V2 d = d_length(v1 - v0, d_out);
    d_v0 -= d;
    d_v1 += d;
}
// d_normalize (vector.h:~555)
inline V2 d_normalize(V2 v0, V2 d_n) {
    float l = len(v0);
    V2 n = vdiv(v0, l);
    V2 d_v0 = vdiv(d_n, l);
    float d_l = -dot(d_n, n) / l;
    d_v0 += d_length(v0, d_l);
    return d_v0;
}

struct M3 {
    float m[3][3];
    float &operator()(int i, int j) { return m[i][j]; }
    float operator()(int i, int j) const { return m[i][j]; }
};
inline M3 m3_zero() {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = 0.f;
    return r;
}
inline M3 m3_from(const float *a) {
    M3 r;
    for (int i = 0; i < 9; i++) r.m[i / 3][i % 3] = a[i];
    return r;
}
// inverse (matrix.h:211-231)
inline M3 m3_inverse(const M3 &m) {
    float det = m(0, 0) * (m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)) -
                m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
                m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
    float invdet = 1 / det;
    M3 r = m3_zero();
    r(0, 0) = (m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2)) * invdet;
    r(0, 1) = (m(0, 2) * m(2, 1) - m(0, 1) * m(2, 2)) * invdet;
    r(0, 2) = (m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1)) * invdet;
    r(1, 0) = (m(1, 2) * m(2, 0) - m(1, 0) * m(2, 2)) * invdet;
    r(1, 1) = (m(0, 0) * m(2, 2) - m(0, 2) * m(2, 0)) * invdet;
    r(1, 2) = (m(1, 0) * m(0, 2) - m(0, 0) * m(1, 2)) * invdet;
    r(2, 0) = (m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1)) * invdet;
    r(2, 1) = (m(2, 0) * m(0, 1) - m(0, 0) * m(2, 1)) * invdet;
    r(2, 2) = (m(0, 0) * m(1, 1) - m(1, 0) * m(0, 1)) * invdet;
    return r;
}
inline M3 m3_transpose(const M3 &m) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = m.m[j][i];
    return r;
}
inline M3 m3_mul(const M3 &a, const M3 &b) {  // matrix.h:170-181
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            r.m[i][j] = 0.f;
            for (int k = 0; k < 3; k++) r.m[i][j] += a.m[i][k] * b.m[k][j];
        }
    return r;
}
inline M3 m3_neg(const M3 &a) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = -a.m[i][j];
    return r;
}
// xform_pt (matrix.h:506-512)
inline V2 xform_pt(const M3 &m, V2 pt) {
    float t0 = m(0, 0) * pt.x + m(0, 1) * pt.y + m(0, 2);
    float t1 = m(1, 0) * pt.x + m(1, 1) * pt.y + m(1, 2);
    float t2 = m(2, 0) * pt.x + m(2, 1) * pt.y + m(2, 2);
    return {t0 / t2, t1 / t2};
}
// d_xform_pt (matrix.h:515-539)
inline void d_xform_pt(const M3 &m, V2 pt, V2 d_out, M3 &d_m, V2 &d_pt) {
    float t0 = m(0, 0) * pt.x + m(0, 1) * pt.y + m(0, 2);
    float t1 = m(1, 0) * pt.x + m(1, 1) * pt.y + m(1, 2);
    float t2 = m(2, 0) * pt.x + m(2, 1) * pt.y + m(2, 2);
    V2 out = {t0 / t2, t1 / t2};
    float dt0 = d_out.x / t2;
    float dt1 = d_out.y / t2;
    float dt2 = -(d_out.x * out.x + d_out.y * out.y) / t2;
    d_m(0, 0) += dt0 * pt.x;
    d_m(0, 1) += dt0 * pt.y;
    d_m(0, 2) += dt0;
    d_m(1, 0) += dt1 * pt.x;
    d_m(1, 1) += dt1 * pt.y;
    d_m(1, 2) += dt1;
    d_m(2, 0) += dt2 * pt.x;
    d_m(2, 1) += dt2 * pt.y;
    d_m(2, 2) += dt2;
    d_pt.x += dt0 * m(0, 0) + dt1 * m(1, 0) + dt2 * m(2, 0);
    d_pt.y += dt0 * m(0, 1) + dt1 * m(1, 1) + dt2 * m(2, 1);
}
// xform_normal (matrix.h:541-545)
inline V2 xform_normal(const M3 &m_inv, V2 n) {
    return normalize(V2{m_inv(0, 0) * n.x + m_inv(1, 0) * n.y,
                        m_inv(0, 1) * n.x + m_inv(1, 1) * n.y});
}

struct Box {
    float minx, miny, maxx, maxy;
};
inline Box box_empty() {
    float inf = std::numeric_limits<float>::infinity();
    return {inf, inf, -inf, -inf};
}
// The reference's min/max templates: a < b ? a : b, a > b ? a : b (diffvg.h:63-73)
inline float fmin_r(float a, float b) { return a < b ? a : b; }
inline float fmax_r(float a, float b) { return a > b ? a : b; }
inline Box merge(const Box &b, V2 p) {  // aabb.h:16-20
    return {fmin_r(p.x, b.minx), fmin_r(p.y, b.miny), fmax_r(p.x, b.maxx), fmax_r(p.y, b.maxy)};
}
inline Box merge(const Box &a, const Box &b) {  // aabb.h:22-27
    return {fmin_r(a.minx, b.minx), fmin_r(a.miny, b.miny), fmax_r(a.maxx, b.maxx),
            fmax_r(a.maxy, b.maxy)};
}
inline bool inside(const Box &b, V2 p) {  // aabb.h:29-33
    return p.x >= b.minx && p.x <= b.maxx && p.y >= b.miny && p.y <= b.maxy;
}
inline bool inside(const Box &b, V2 p, float r) {  // aabb.h:35-40, 61-66
    return p.x >= b.minx - r && p.x <= b.maxx + r && p.y >= b.miny - r && p.y <= b.maxy + r;
}
inline Box transform_box(const M3 &m, const Box &b) {  // aabb.h:49-57
    Box r = box_empty();
    r = merge(r, xform_pt(m, V2{b.minx, b.miny}));
    r = merge(r, xform_pt(m, V2{b.minx, b.maxy}));
    r = merge(r, xform_pt(m, V2{b.maxx, b.miny}));
    r = merge(r, xform_pt(m, V2{b.maxx, b.maxy}));
    return r;
}

// ---- PCG32 (pcg.h:11-40) ----
struct Pcg {
    uint64_t state, inc;
};
inline uint32_t pcg_next(Pcg &r) {
    uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + (r.inc | 1);
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
}
inline float pcg_float(Pcg &r) {
    uint32_t u = (pcg_next(r) >> 9) | 0x3f800000u;
    float f;
    static_assert(sizeof(f) == sizeof(u), "");
    __builtin_memcpy(&f, &u, 4);
    return f - 1.0f;
}
inline Pcg pcg_init(int idx, uint64_t seed) {
    Pcg s;
    s.state = 0U;
    s.inc = (((uint64_t)idx + 1) << 1u) | 1u;
    pcg_next(s);
    s.state += (0x853c49e6748fea9bULL + seed);
    pcg_next(s);
    return s;
}

// ---- Morton code (diffvg.h:109-126, scene.cpp:431-437) ----
inline uint32_t expand_bits(uint32_t x) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < 10; i++) r |= (x & (1u << i)) << i;
    return r;
}
// float -> uint32_t as the x86-64 compilers lower it: a 64-bit truncating conversion
// whose low 32 bits are kept; out-of-range and NaN inputs give 0x8000000000000000.
inline uint32_t f2u32_x86(float f) {
    if (!(f > -9.2233720368547758e18f && f < 9.2233720368547758e18f)) return 0u;
    return (uint32_t)(uint64_t)(int64_t)f;
}
inline uint32_t morton2d(V2 p, int w, int h) {
    V2 pp = vdivv(p, V2{(float)w, (float)h});
    uint32_t xi = f2u32_x86(pp.x * 1023);
    uint32_t yi = f2u32_x86(pp.y * 1023);
    return (expand_bits(xi) << 1u) | (expand_bits(yi) << 0u);
}

// ---- Root solvers (solve.h) ----
// solve_quadratic<float> (solve.h:5-27); T root_discrim = sqrt(discrim) rounds to float.
inline bool solve_quadratic_f(float a, float b, float c, float *t0, float *t1) {
    float discrim = b * b - 4 * a * c;
    if (discrim < 0) return false;
    float root_discrim = (float)std::sqrt((double)discrim);
    float q;
    if (b < 0)
        q = -0.5f * (b - root_discrim);
    else
        q = -0.5f * (b + root_discrim);
    *t0 = q / a;
    *t1 = c / q;
    if (*t0 > *t1) {
        float tmp = *t0;
        *t0 = *t1;
        *t1 = tmp;
    }
    return true;
}
inline bool solve_quadratic_d(double a, double b, double c, double *t0, double *t1) {
    double discrim = b * b - 4 * a * c;
    if (discrim < 0) return false;
    double root_discrim = std::sqrt(discrim);
    double q;
    if (b < 0)
        q = -0.5f * (b - root_discrim);
    else
        q = -0.5f * (b + root_discrim);
    *t0 = q / a;
    *t1 = c / q;
    if (*t0 > *t1) {
        double tmp = *t0;
        *t0 = *t1;
        *t1 = tmp;
    }
    return true;
}
// solve_cubic<float> (solve.h:29-59) with the double promotion of sqrt/acos/cos/pow.
inline int solve_cubic_f(float a, float b, float c, float d, float t[3]) {
    if (std::fabs((double)a) < 1e-6f) {
        if (solve_quadratic_f(b, c, d, &t[0], &t[1])) return 2;
        return 0;
    }
    b /= a;
    c /= a;
    d /= a;
    float Q = (b * b - 3 * c) / 9.f;
    float R = (2 * b * b * b - 9 * b * c + 27 * d) / 54.f;
    if (R * R < Q * Q * Q) {
        float theta = (float)std::acos((double)R / std::sqrt((double)(Q * Q * Q)));
        double sq = std::sqrt((double)Q);
        double b3 = (double)(b / 3.f);
        t[0] = (float)((double)-2.f * sq * std::cos((double)(theta / 3.f)) - b3);
        t[1] = (float)((double)-2.f * sq *
                       std::cos((double)((theta + 2.f * (float)kPi) / 3.f)) - b3);
        t[2] = (float)((double)-2.f * sq *
                       std::cos((double)((theta - 2.f * (float)kPi) / 3.f)) - b3);
        return 3;
    } else {
        double disc = std::sqrt((double)(R * R - Q * Q * Q));
        double third = (double)(float)(1. / 3.);
        float A = R > 0 ? (float)-std::pow((double)R + disc, third)
                        : (float)std::pow((double)-R + disc, third);
        float B = std::fabs((double)A) > 1e-6f ? Q / A : 0.f;
        t[0] = (A + B) - b / 3.f;
        return 1;
    }
}
// solve_cubic<double> (solve.h:29-59)
inline int solve_cubic_d(double a, double b, double c, double d, double t[3]) {
    if (std::fabs(a) < 1e-6f) {
        if (solve_quadratic_d(b, c, d, &t[0], &t[1])) return 2;
        return 0;
    }
    b /= a;
    c /= a;
    d /= a;
    double Q = (b * b - 3 * c) / 9.f;
    double R = (2 * b * b * b - 9 * b * c + 27 * d) / 54.f;
    if (R * R < Q * Q * Q) {
        double theta = std::acos(R / std::sqrt(Q * Q * Q));
        t[0] = -2.f * std::sqrt(Q) * std::cos(theta / 3.f) - b / 3.f;
        t[1] = -2.f * std::sqrt(Q) * std::cos((theta + 2.f * kPi) / 3.f) - b / 3.f;
        t[2] = -2.f * std::sqrt(Q) * std::cos((theta - 2.f * kPi) / 3.f) - b / 3.f;
        return 3;
    } else {
        double A = R > 0 ? -std::pow(R + std::sqrt(R * R - Q * Q * Q), 1. / 3.)
                         : std::pow(-R + std::sqrt(R * R - Q * Q * Q), 1. / 3.);
        double B = std::fabs(A) > 1e-6f ? Q / A : 0.0;
        t[0] = (A + B) - b / 3.0;
        return 1;
    }
}

// ---- Pixel filters (filter.h:22-106) + the Gaussian extension ----
enum FilterType { kBox = 0, kTent = 1, kParabolic = 2, kHann = 3, kGaussian = 4 };

// Gaussian extension (no reference counterpart): sigma = radius / 3, truncated at
// |dx|,|dy| <= radius, weight exp(-(dx^2+dy^2)/(2 sigma^2)) / (2 pi sigma^2).
inline float gaussian_weight(float r, float dx, float dy) {
    double s = (double)r / 3.0;
    double d2 = (double)dx * dx + (double)dy * dy;
    return (float)(std::exp(-d2 / (2.0 * s * s)) / (2.0 * kPi * s * s));
}
inline float filter_weight(int type, float radius, float dx, float dy) {
    if (std::fabs((double)dx) > radius || std::fabs((double)dy) > radius) return 0;
    if (type == kBox) {
        float t = 2 * radius;
        return 1.f / (t * t);
    } else if (type == kTent) {
        double fx = (double)radius - std::fabs((double)dx);
        double fy = (double)radius - std::fabs((double)dy);
        float r2 = radius * radius;
        return (float)(fx * fy / (double)(r2 * r2));
    } else if (type == kParabolic) {
        float qx = dx / radius, qy = dy / radius;
        return (4.f / 3.f) * (1 - qx * qx) * (4.f / 3.f) * (1 - qy * qy);
    } else if (type == kHann) {
        float ndx = (dx / (2 * radius)) + 0.5f;
        float ndy = (dy / (2 * radius)) + 0.5f;
        float tp = (float)(2 * kPi);
        double v = 0.5f * (1.0 - std::cos((double)(tp * ndx))) * 0.5f *
                   (1.0 - std::cos((double)(tp * ndy))) / (double)(radius * radius);
        return (float)v;
    } else {
        return gaussian_weight(radius, dx, dy);
    }
}
// d_compute_filter_weight (filter.h:50-106): returns the radius adjoint that the
// reference atomically adds into d_filter->radius.
inline float d_filter_weight(int type, float radius, float dx, float dy, float d_return) {
    if (type == kBox) {
        float t = 2 * radius;
        return d_return * (-2) * 2 * radius / (t * t * t);
    } else if (type == kTent) {
        double fx = (double)radius - std::fabs((double)dx);
        double fy = (double)radius - std::fabs((double)dy);
        float norm = 1 / (radius * radius);
        double d_fx = (double)d_return * fy * norm;
        double d_fy = (double)d_return * fx * norm;
        double d_norm = (double)d_return * fx * fy;
        return (float)(d_fx + d_fy + (-4) * d_norm / std::pow((double)radius, 5.0));
    } else if (type == kParabolic) {
        float r3 = radius * radius * radius;
        return -(2 * (dx * dx) + 2 * (dy * dy)) / r3;
    } else if (type == kHann) {
        float ndx = (dx / (2 * radius)) + 0.5f;
        float ndy = (dy / (2 * radius)) + 0.5f;
        float tp = (float)(2 * kPi);
        double fx = 0.5f * (1.0 - std::cos((double)(tp * ndx)));
        double fy = 0.5f * (1.0 - std::cos((double)(tp * ndy)));
        float norm = 1 / (radius * radius);
        double d_fx = (double)d_return * fy * norm;
        double d_fy = (double)d_return * fx * norm;
        double d_norm = (double)d_return * fx * fy;
        double d_ndx = d_fx * 0.5f * std::sin((double)(tp * ndx)) * (double)tp;
        double d_ndy = d_fy * 0.5f * std::sin((double)(tp * ndy)) * (double)tp;
        float tr = 2 * radius;
        return (float)(d_ndx * (double)(-2 * dx / (tr * tr)) +
                       d_ndy * (double)(-2 * dy / (tr * tr)) +
                       (-2) * d_norm / (double)(radius * radius * radius));
    } else {
        // d/dr of the truncated Gaussian with sigma = r / 3 (truncation edge ignored).
        double s = (double)radius / 3.0;
        double d2 = (double)dx * dx + (double)dy * dy;
        double w = std::exp(-d2 / (2.0 * s * s)) / (2.0 * kPi * s * s);
        double dw_ds = w * (d2 / (s * s * s) - 2.0 / s);
        return (float)((double)d_return * dw_ds / 3.0);
    }
}

inline float smoothstep(float d) {  // diffvg.cpp:817-821
    float t = (d + 1.f) / 2.f;
    t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
    return t * t * (3 - 2 * t);
}
inline float d_smoothstep(float d, float d_ret) {  // diffvg.cpp:823-833
    if (d < -1.f || d > 1.f) return 0.f;
    float t = (d + 1.f) / 2.f;
    float d_t = d_ret * (6 * t - 6 * t * t);
    return d_t / 2.f;
}

}  // namespace orc
