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

namespace oracle {

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

}  // namespace oracle
