// crmath.hpp -- correctly rounded exp(x), x^(-1/2) and x^(-3/2) in binary64.
//
// The disk (disk.rs:155-268) and evolving-dissipation (constant_time_lag.rs:
// 20-165) terms call f64::exp and f64::powf, which the reference takes from
// the platform libm (results within about 0.52 ulp, so not always the
// nearest double).  The device and the oracle both return the correctly
// rounded value instead, so that they agree bit for bit with each other and
// differ from a libm result only where that result is not the nearest double
// (about 1 evaluation in 1,200 for glibc 2.35 on these functions).
//
// Method: every function evaluates its result as an unevaluated sum of two
// doubles (double-double: Dekker 1971, Knuth TAOCP vol. 2 4.2.2; the
// two_sum / two_prod error-free transformations) with a relative error near
// 2^-103, then rounds that sum once.  The result is the correctly rounded
// value unless the exact one lies within about 2^-103 (relative) of the
// midpoint between two doubles; the tests compare with binary128 evaluations
// on a million arguments.
//
//   exp:   x = k ln2 + r (ln2 in three parts, k * ln2 formed exactly),
//          r' = r / 2^8, expm1(r') by its Taylor series to r'^10 (Horner in
//          double-double), then eight squarings in the expm1 form
//          m <- 2m + m^2, and 2^k (1 + m).  Subnormal results are rounded
//          once from the double-double value.
//   x^(-1/2), x^(-3/2): sqrt(x) as s + s_lo (one FMA residual), then the
//          reciprocal of s (or of x s) refined by one Newton step with its
//          quadratic term.  For 2^-600 < x < 2^600.
//
// All arithmetic must be compiled without contraction (-ffp-contract=off):
// the error-free transformations rely on every + and * rounding once.
#pragma once

#include <math.h>

#ifndef POSCR_HD
#define POSCR_HD __host__ __device__
#endif

namespace poscr {

struct DD {
    double hi, lo;
};

POSCR_HD inline DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    const double e = (a - (s - bb)) + (b - bb);
    return DD{s, e};
}
POSCR_HD inline DD fast_two_sum(double a, double b) {   // |a| >= |b|
    const double s = a + b;
    return DD{s, b - (s - a)};
}
POSCR_HD inline DD two_prod(double a, double b) {
    const double p = a * b;
    return DD{p, __builtin_fma(a, b, -p)};
}
POSCR_HD inline DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    const DD t = two_sum(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return fast_two_sum(s.hi, s.lo);
}
POSCR_HD inline DD dd_mul(DD a, DD b) {
    DD p = two_prod(a.hi, b.hi);
    p.lo = p.lo + (a.hi * b.lo + a.lo * b.hi);
    return fast_two_sum(p.hi, p.lo);
}

// 1/n! as double-double, n = 2..10
POSCR_HD inline DD inv_factorial(int n) {
    switch (n) {
        case 2: return DD{0.5, 0.0};
        case 3: return DD{0.16666666666666666, 9.25185853854297e-18};
        case 4: return DD{0.041666666666666664, 2.3129646346357427e-18};
        case 5: return DD{0.008333333333333333, 1.1564823173178714e-19};
        case 6: return DD{0.001388888888888889, -5.300543954373577e-20};
        case 7: return DD{0.0001984126984126984, 1.7209558293420705e-22};
        case 8: return DD{2.48015873015873e-05, 2.1511947866775882e-23};
        case 9: return DD{2.7557319223985893e-06, -1.858393274046472e-22};
        default: return DD{2.755731922398589e-07, 2.3767714622250297e-23};
    }
}

// Round the double-double value v * 2^k to the nearest double (ties to even),
// for results in the subnormal range.
POSCR_HD inline double round_scaled_subnormal(DD v, int k) {
    // v * 2^(k + 1074) counts units of the smallest subnormal; it is below 2^53
    // (both scalings are exact: the scaled values are normal doubles)
    const double H = ldexp(v.hi, k + 1074), L = ldexp(v.lo, k + 1074);
    double q = rint(H);
    const DD d = two_sum(H - q, L);   // exact remainder (H + L) - q, |H - q| <= 1/2
    if (d.hi > 0.5 || (d.hi == 0.5 && d.lo > 0.)) {
        q = q + 1.;
    } else if (d.hi < -0.5 || (d.hi == -0.5 && d.lo < 0.)) {
        q = q - 1.;
    } else if ((d.hi == 0.5 || d.hi == -0.5) && d.lo == 0.) {   // exact tie: to even
        const double other = q + (d.hi > 0. ? 1. : -1.);
        if (fmod(other, 2.) == 0.) q = other;
    }
    return ldexp(q, -1074);
}

POSCR_HD inline double cr_exp(double x) {
    if (x != x) return x;
    if (x > 709.79) return HUGE_VAL;   // exp(709.79) > DBL_MAX + ulp/2
    if (x < -746.) return 0.;           // exp(-746) < 2^-1075
    constexpr double INV_LN2 = 1.4426950408889634;
    constexpr double LN2_0 = 0.6931471805599453, LN2_1 = 2.3190468138462996e-17, LN2_2 = 5.707708438416212e-34;
    const double k = rint(x * INV_LN2);
    // r = x - k ln2 as a double-double; k ln2 = p0 + p1 + k*LN2_2 with p0, p1 exact products
    const DD p0 = two_prod(k, LN2_0);
    const DD p1 = two_prod(k, LN2_1);
    DD r = two_sum(x, -p0.hi);
    r = dd_add(r, DD{-p0.lo, 0.});
    r = dd_add(r, DD{-p1.hi, -p1.lo});
    r = dd_add(r, DD{-(k * LN2_2), 0.});
    // r' = r / 256; expm1(r') = r' (1 + r'/2! + ... + r'^9/10!)
    const DD rs{r.hi * (1. / 256.), r.lo * (1. / 256.)};
    DD q = inv_factorial(10);
    for (int n = 9; n >= 2; --n) q = dd_add(dd_mul(q, rs), inv_factorial(n));
    q = dd_add(dd_mul(q, rs), DD{1., 0.});
    DD m = dd_mul(q, rs);
    // (1 + m)^2 = 1 + (2m + m^2), eight times
    for (int i = 0; i < 8; ++i) m = dd_add(DD{2. * m.hi, 2. * m.lo}, dd_mul(m, m));
    const DD s = two_sum(1., m.hi);
    const double t = s.lo + m.lo;
    const int ki = (int)k;
    if (ki < -1021) {   // subnormal (or near-subnormal) result: round once at the final scale
        const DD v = fast_two_sum(s.hi, t);
        return round_scaled_subnormal(v, ki);
    }
    return ldexp(s.hi + t, ki);
}

// sqrt(x) as a double-double
POSCR_HD inline DD dd_sqrt(double x) {
    const double s = sqrt(x);
    const double res = __builtin_fma(-s, s, x);   // exact: x - s^2
    return DD{s, res / (2. * s)};
}

// 1 / (d.hi + d.lo), correctly rounded
POSCR_HD inline double cr_recip_dd(DD d) {
    const double y = 1. / d.hi;
    const double e = __builtin_fma(-d.hi, y, 1.) - d.lo * y;   // 1 - d y
    const double ylo = y * (e + e * e);
    return y + ylo;
}

POSCR_HD inline double cr_rsqrt(double x) {   // x^(-1/2)
    return cr_recip_dd(dd_sqrt(x));
}

POSCR_HD inline double cr_rsqrt3(double x) {   // x^(-3/2)
    const DD s = dd_sqrt(x);
    DD p = two_prod(x, s.hi);
    p.lo = p.lo + x * s.lo;
    return cr_recip_dd(fast_two_sum(p.hi, p.lo));
}

}  // namespace poscr
