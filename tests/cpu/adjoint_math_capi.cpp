// Host build of the per-lane adjoint math (posidonius_amd/csrc/adjoint_math.hpp)
// for the derivative checks in tests/test_adjoint_math.py.  Test
// infrastructure only: the HIP library compiles the same header for the GPU.
#include <cmath>
using std::sqrt; using std::fabs; using std::cos; using std::sin; using std::cosh; using std::sinh;
#define POS_HD inline
#include "../../posidonius_amd/csrc/adjoint_math.hpp"

using namespace posadj;

namespace {
V3 ld(const double *p) { return V3{p[0], p[1], p[2]}; }
void st(double *p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
PairConst mkc(const double *c, int flags) {
    return PairConst{c[0], c[1], c[2], c[3], c[4], c[5], (flags & 1) != 0, (flags & 2) != 0, (flags & 4) != 0};
}
PairPar mkp(const double *p) { return PairPar{p[0], p[1], p[2], p[3], p[4], p[5]}; }
PairIn mki(const double *i) { return PairIn{ld(i), ld(i + 3), ld(i + 6), ld(i + 9)}; }
}  // namespace

extern "C" {
// in12 = r, u, Wh, Wp; cst6 = mh, mp, Rh, Rp, Ih, Ip; par6 = sh, sp, k2h, k2p, kfh, kfp;
// flags: 1 tides, 2 flattening, 4 GR; out12 = F, g, Np, Nh
void adj_pair_forward(const double *in12, const double *cst6, const double *par6, int flags, double *out12) {
    PairOut o = pair_forward(mkc(cst6, flags), mkp(par6), mki(in12));
    st(out12, o.F); st(out12 + 3, o.g); st(out12 + 6, o.Np); st(out12 + 9, o.Nh);
}
void adj_pair_vjp(const double *in12, const double *cst6, const double *par6, int flags, const double *ob12,
                  double *ib12, double *pb6) {
    PairOut ob{ld(ob12), ld(ob12 + 3), ld(ob12 + 6), ld(ob12 + 9)};
    PairIn ib{zero3(), zero3(), zero3(), zero3()};
    PairParBar pb{0., 0., 0., 0., 0., 0.};
    pair_vjp(mkc(cst6, flags), mkp(par6), mki(in12), ob, ib, pb);
    st(ib12, ib.r); st(ib12 + 3, ib.u); st(ib12 + 6, ib.Wh); st(ib12 + 9, ib.Wp);
    pb6[0] = pb.sh; pb6[1] = pb.sp; pb6[2] = pb.k2h; pb6[3] = pb.k2p; pb6[4] = pb.kfh; pb6[5] = pb.kfp;
}
void adj_stiefel(double beta, double X, double *g6) { stiefel_g05(beta, X, g6); }
double adj_kepler_residual(double mu, double dt, double X, const double *x0, const double *v0) {
    return kepler_residual(mu, dt, X, ld(x0), ld(v0));
}
void adj_kepler_forward(double mu, double dt, double X, const double *x0, const double *v0, double *x1, double *v1) {
    V3 a, b;
    kepler_forward_x(mu, dt, X, ld(x0), ld(v0), a, b);
    st(x1, a); st(v1, b);
}
void adj_kepler_vjp(double mu, double dt, double X, const double *x0, const double *v0, const double *x1b,
                    const double *v1b, double *x0b, double *v0b) {
    V3 a = zero3(), b = zero3();
    kepler_vjp(mu, dt, X, ld(x0), ld(v0), ld(x1b), ld(v1b), a, b);
    st(x0b, a); st(v0b, b);
}
void adj_gravity_pair_vjp(double mj, const double *d, const double *ab, double *db) {
    st(db, gravity_pair_vjp(mj, ld(d), ld(ab)));
}
}
