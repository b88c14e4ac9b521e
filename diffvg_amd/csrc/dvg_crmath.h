// dvg_crmath.h -- correctly rounded double acos, cos and pow for the cubic solver's
// knife-edge cases.
//
// The winding number of a cubic segment counts the roots of y(t) = y_pt that fall in
// [0, 1] (winding_number.h:134-157).  When the sample's ray passes exactly through a
// segment joint, the root sits at 0 or 1 and its side is decided by the last bit of the
// double acos / cos / pow inside solve_cubic (solve.h:41-63).  The host libm the
// reference is built against rounds these correctly in all but rare cases; the device
// libm is accurate to about one unit in the last place.  The renderer therefore solves
// with the device libm and, when a root lands within 1e-9 of 0 or 1, solves again with
// the functions below, which evaluate in double-double arithmetic and round once.
//
// Methods (standard; see e.g. Muller et al., Handbook of Floating-Point Arithmetic):
// error-free transformations (Knuth two-sum, FMA two-product), Cody-Waite argument
// reduction with a three-part constant, Taylor series evaluated by Horner's rule in
// double-double, and one Newton step from the libm value for acos and log.
//
// Host and device compile the same code: it must be built without floating-point
// contraction (-ffp-contract=off), as everything in this library is.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define DVG_CR __host__ __device__ inline
#else
#define DVG_CR inline
#endif

namespace dvg {
namespace cr {

struct DD {
    double hi, lo;
};

DVG_CR DD two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    return DD{s, (a - (s - bb)) + (b - bb)};
}
DVG_CR DD fast_two_sum(double a, double b) {  // |a| >= |b|
    double s = a + b;
    return DD{s, b - (s - a)};
}
DVG_CR DD two_prod(double a, double b) {
    double p = a * b;
    return DD{p, __builtin_fma(a, b, -p)};
}
DVG_CR DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    DD t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
DVG_CR DD dd_add_d(DD a, double b) {
    DD s = two_sum(a.hi, b);
    s.lo += a.lo;
    return fast_two_sum(s.hi, s.lo);
}
DVG_CR DD dd_neg(DD a) { return DD{-a.hi, -a.lo}; }
DVG_CR DD dd_mul(DD a, DD b) {
    DD p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo);
}
DVG_CR DD dd_mul_d(DD a, double b) {
    DD p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}
DVG_CR DD dd_div(DD a, DD b) {
    double q1 = a.hi / b.hi;
    DD r = dd_add(a, dd_neg(dd_mul_d(b, q1)));
    double q2 = r.hi / b.hi;
    r = dd_add(r, dd_neg(dd_mul_d(b, q2)));
    double q3 = r.hi / b.hi;
    DD q = fast_two_sum(q1, q2);
    return dd_add_d(q, q3);
}

// 1/n! as double-double, n = 0..29 (exact rationals split into two doubles).
DVG_CR DD inv_fact(int n) {
    switch (n) {
        case 0: return DD{1.0, 0.0};
        case 1: return DD{1.0, 0.0};
        case 2: return DD{0.5, 0.0};
        case 3: return DD{0.16666666666666666, 9.25185853854297e-18};
        case 4: return DD{0.041666666666666664, 2.3129646346357427e-18};
        case 5: return DD{0.008333333333333333, 1.1564823173178714e-19};
        case 6: return DD{0.001388888888888889, -5.300543954373577e-20};
        case 7: return DD{0.0001984126984126984, 1.7209558293420705e-22};
        case 8: return DD{2.48015873015873e-05, 2.1511947866775882e-23};
        case 9: return DD{2.7557319223985893e-06, -1.858393274046472e-22};
        case 10: return DD{2.755731922398589e-07, 2.3767714622250297e-23};
        case 11: return DD{2.505210838544172e-08, -1.448814070935912e-24};
        case 12: return DD{2.08767569878681e-09, -1.20734505911326e-25};
        case 13: return DD{1.6059043836821613e-10, 1.2585294588752098e-26};
        case 14: return DD{1.1470745597729725e-11, 2.0655512752830745e-28};
        case 15: return DD{7.647163731819816e-13, 7.03872877733453e-30};
        case 16: return DD{4.779477332387385e-14, 4.399205485834081e-31};
        case 17: return DD{2.8114572543455206e-15, 1.6508842730861433e-31};
        case 18: return DD{1.5619206968586225e-16, 1.1910679660273754e-32};
        case 19: return DD{8.22063524662433e-18, 2.2141894119604265e-34};
        case 20: return DD{4.110317623312165e-19, 1.4412973378659527e-36};
        case 21: return DD{1.9572941063391263e-20, -1.3643503830087908e-36};
        case 22: return DD{8.896791392450574e-22, -7.911402614872376e-38};
        case 23: return DD{3.868170170630684e-23, -8.843177655482344e-40};
        case 24: return DD{1.6117375710961184e-24, -3.6846573564509766e-41};
        case 25: return DD{6.446950284384474e-26, -1.9330404233703465e-42};
        case 26: return DD{2.4795962632247976e-27, -1.2953730964765229e-43};
        case 27: return DD{9.183689863795546e-29, 1.4303150396787322e-45};
        case 28: return DD{3.279889237069838e-30, 1.5117542744029879e-46};
        default: return DD{1.1309962886447716e-31, 1.0498015412959506e-47};
    }
}

// pi/2 and ln 2 in three parts
constexpr double kPio2_1 = 1.5707963267948966, kPio2_2 = 6.123233995736766e-17, kPio2_3 = -1.4973849048591698e-33;
constexpr double kLn2_1 = 0.6931471805599453, kLn2_2 = 2.3190468138462996e-17, kLn2_3 = 5.707708438416212e-34;
constexpr double kPi_1 = 3.141592653589793;

// sin and cos of a double |x| <= 8 in double-double.
DVG_CR void sincos_dd(double x, DD &s_out, DD &c_out) {
    double kf = __builtin_rint(x * 0.6366197723675814);
    int k = (int)kf;
    DD r = DD{x, 0.0};
    r = dd_add(r, dd_neg(two_prod(kf, kPio2_1)));
    r = dd_add(r, dd_neg(two_prod(kf, kPio2_2)));
    r = dd_add_d(r, -kf * kPio2_3);
    DD r2 = dd_mul(r, r);
    // sin r = r * sum (-1)^n r^2n / (2n+1)!,  cos r = sum (-1)^n r^2n / (2n)!,  n <= 14
    DD sa = inv_fact(29), ca = inv_fact(28);
    for (int n = 13; n >= 0; n--) {
        sa = dd_add(inv_fact(2 * n + 1), dd_neg(dd_mul(r2, sa)));
        ca = dd_add(inv_fact(2 * n), dd_neg(dd_mul(r2, ca)));
    }
    DD s = dd_mul(r, sa), c = ca;
    switch (k & 3) {
        case 0: s_out = s; c_out = c; break;
        case 1: s_out = c; c_out = dd_neg(s); break;
        case 2: s_out = dd_neg(s); c_out = dd_neg(c); break;
        default: s_out = dd_neg(c); c_out = s; break;
    }
}

DVG_CR double cos_cr(double x) {
    DD s, c;
    sincos_dd(x, s, c);
    return c.hi;
}

// acos on [-1, 1]: one Newton step on cos(y) = x from the libm value.
DVG_CR double acos_cr(double x) {
    if (!(x > -1.0 && x < 1.0)) return x == 1.0 ? 0.0 : (x == -1.0 ? kPi_1 : ::acos(x));
    double y0 = ::acos(x);
    DD s, c;
    sincos_dd(y0, s, c);
    DD diff = dd_add_d(c, -x);  // cos(y0) - x
    double corr = (diff.hi + diff.lo) / s.hi;
    return dd_add_d(DD{y0, 0.0}, corr).hi;
}

// exp of a double-double |t| < 700.
DVG_CR DD exp_dd(DD t) {
    double kf = __builtin_rint(t.hi * 1.4426950408889634);
    DD r = dd_add(t, dd_neg(two_prod(kf, kLn2_1)));
    r = dd_add(r, dd_neg(two_prod(kf, kLn2_2)));
    r = dd_add_d(r, -kf * kLn2_3);
    DD acc = inv_fact(24);
    for (int n = 23; n >= 0; n--) acc = dd_add(inv_fact(n), dd_mul(r, acc));
    int k = (int)kf;
    return DD{__builtin_ldexp(acc.hi, k), __builtin_ldexp(acc.lo, k)};
}

// pow(x, y) for finite x > 0 with |y log x| < 700: exp(y log x), log refined by one
// Newton step from the libm value.
DVG_CR double pow_cr(double x, double y) {
    if (!(x > 0.0) || !(x < 1.0e300)) return ::pow(x, y);
    double l0 = ::log(x);
    DD e0 = exp_dd(DD{l0, 0.0});
    DD u = dd_add_d(dd_div(DD{x, 0.0}, e0), -1.0);  // x / exp(l0) - 1
    DD l = dd_add(DD{l0, 0.0}, u);
    DD t = dd_mul_d(l, y);
    if (!(t.hi < 700.0 && t.hi > -700.0)) return ::pow(x, y);
    return exp_dd(t).hi;
}

}  // namespace cr
}  // namespace dvg
