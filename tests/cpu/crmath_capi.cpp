// Host build of the correctly rounded exp / x^(-1/2) / x^(-3/2) of
// posidonius_amd/csrc/crmath.hpp, next to binary128 (libquadmath) evaluations
// rounded to double, for tests/test_crmath.py.  Test infrastructure only.
#include <cmath>
#include <quadmath.h>
using std::ldexp; using std::rint; using std::sqrt; using std::fmod;
#define POSCR_HD
#include "../../posidonius_amd/csrc/crmath.hpp"

extern "C" {
void cr_exp_n(const double *x, long n, double *out) { for (long i = 0; i < n; ++i) out[i] = poscr::cr_exp(x[i]); }
void cr_rsqrt_n(const double *x, long n, double *out) { for (long i = 0; i < n; ++i) out[i] = poscr::cr_rsqrt(x[i]); }
void cr_rsqrt3_n(const double *x, long n, double *out) { for (long i = 0; i < n; ++i) out[i] = poscr::cr_rsqrt3(x[i]); }
void q_exp_n(const double *x, long n, double *out) { for (long i = 0; i < n; ++i) out[i] = (double)expq((__float128)x[i]); }
void q_pow_n(const double *x, long n, double y, double *out) {
    for (long i = 0; i < n; ++i) out[i] = (double)powq((__float128)x[i], (__float128)y);
}
void libm_exp_n(const double *x, long n, double *out) { for (long i = 0; i < n; ++i) out[i] = std::exp(x[i]); }
void libm_pow_n(const double *x, long n, double y, double *out) { for (long i = 0; i < n; ++i) out[i] = std::pow(x[i], y); }
}
