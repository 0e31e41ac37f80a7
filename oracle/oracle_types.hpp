// oracle_types.hpp -- scalar types for instantiating the oracle templates.
//
// TEST INFRASTRUCTURE ONLY (see posidonius_oracle.hpp).
//
// FlopCount: a binary64 value that counts every +, -, *, / and sqrt applied
// to it (one each; the reference fuses nothing, so there are no FMAs to count
// twice).  Running the restated step with it yields the algorithmic work per
// system-step for the iteration counts that the data actually take, which is
// the numerator of the roofline figure.  Comparisons, negation, fabs and
// floor are not counted.
#pragma once

#include <cmath>
#include <cstdint>
#include <quadmath.h>

namespace oracle {

// exp and x^y correctly rounded: the binary128 evaluation (113-bit
// significand, error below one binary128 ulp) rounded to the nearest double.
// The reference calls the platform libm here (f64::exp, f64::powf), whose
// results are within 0.52 ulp; the device computes the correctly rounded
// value in double-double arithmetic, and both sides round the same number.
inline double exp_r(double x) { return (double)expq((__float128)x); }
inline double pow_r(double x, double y) { return (double)powq((__float128)x, (__float128)y); }

struct FlopCount {
    double v = 0.;
    static thread_local uint64_t count;
    FlopCount() = default;
    FlopCount(double x) : v(x) {}
};
inline thread_local uint64_t FlopCount::count = 0;

inline FlopCount operator+(FlopCount a, FlopCount b) { ++FlopCount::count; return FlopCount(a.v + b.v); }
inline FlopCount operator-(FlopCount a, FlopCount b) { ++FlopCount::count; return FlopCount(a.v - b.v); }
inline FlopCount operator*(FlopCount a, FlopCount b) { ++FlopCount::count; return FlopCount(a.v * b.v); }
inline FlopCount operator/(FlopCount a, FlopCount b) { ++FlopCount::count; return FlopCount(a.v / b.v); }
inline FlopCount operator-(FlopCount a) { return FlopCount(-a.v); }
inline double val(const FlopCount &x) { return x.v; }
inline FlopCount sqrt_r(const FlopCount &x) { ++FlopCount::count; return FlopCount(std::sqrt(x.v)); }
inline FlopCount abs_r(const FlopCount &x) { return FlopCount(std::fabs(x.v)); }
inline FlopCount exp_r(const FlopCount &x) { ++FlopCount::count; return FlopCount(exp_r(x.v)); }
inline FlopCount pow_r(const FlopCount &x, double y) { ++FlopCount::count; return FlopCount(pow_r(x.v, y)); }

}  // namespace oracle

namespace oracle {

// Dual<K>: a binary64 value with K tangent directions (forward-mode
// differentiation).  Running the restated step with it is the tangent-linear
// (variational) integration of the discrete step: the derivative of exactly
// the operations executed, including the iterations of the Kepler solver and
// of the implicit midpoint.  Comparisons look at the value only.
template <int K> struct Dual {
    double v = 0.;
    double d[K] = {};
    Dual() = default;
    Dual(double x) : v(x) {}
};
template <int K> inline Dual<K> operator+(const Dual<K> &a, const Dual<K> &b) {
    Dual<K> r(a.v + b.v);
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int K> inline Dual<K> operator-(const Dual<K> &a, const Dual<K> &b) {
    Dual<K> r(a.v - b.v);
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int K> inline Dual<K> operator*(const Dual<K> &a, const Dual<K> &b) {
    Dual<K> r(a.v * b.v);
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int K> inline Dual<K> operator/(const Dual<K> &a, const Dual<K> &b) {
    Dual<K> r(a.v / b.v);
    const double ib = 1. / b.v;
    for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int K> inline Dual<K> operator-(const Dual<K> &a) {
    Dual<K> r(-a.v);
    for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int K> inline bool operator==(const Dual<K> &a, const Dual<K> &b) { return a.v == b.v; }
template <int K> inline bool operator!=(const Dual<K> &a, const Dual<K> &b) { return a.v != b.v; }
template <int K> inline double val(const Dual<K> &x) { return x.v; }
template <int K> inline Dual<K> sqrt_r(const Dual<K> &x) {
    Dual<K> r(std::sqrt(x.v));
    const double h = 0.5 / r.v;
    for (int k = 0; k < K; ++k) r.d[k] = x.d[k] * h;
    return r;
}
template <int K> inline Dual<K> abs_r(const Dual<K> &x) { return x.v < 0. ? -x : x; }
template <int K> inline Dual<K> exp_r(const Dual<K> &x) {
    Dual<K> r(exp_r(x.v));
    for (int k = 0; k < K; ++k) r.d[k] = x.d[k] * r.v;
    return r;
}
template <int K> inline Dual<K> pow_r(const Dual<K> &x, double y) {
    Dual<K> r(pow_r(x.v, y));
    const double h = y * r.v / x.v;
    for (int k = 0; k < K; ++k) r.d[k] = x.d[k] * h;
    return r;
}

}  // namespace oracle
