// adjoint_kernels.hpp -- reverse sweep of the discrete adjoint of the WHFast
// democratic-heliocentric step (tides, rotational flattening, Kidder 1995 GR, wind,
// spin), one segment of recorded steps at a time.
//
// Mapping: the forward step's (one lane per body, one lane group per system,
// several systems per wavefront, of one size or of mixed sizes, one
// wavefront per workgroup).  Each lane
// holds the cotangents of its body's position, velocity and spin angular
// momentum and accumulates the gradients of its body's parameters.  The step
// being reversed (src/integrator/whfast.rs:235-305) is
//   K(dt/2) . [T^-1 . Kepler(dt/2) . jump(dt/2)] . interaction(dt) . [jump(dt/2) . Kepler(dt/2) . T] . K(dt/2)
// with T the democratic-heliocentric transform (whfast.rs:973-1023,
// 1090-1155) and K the implicit-midpoint kick (whfast.rs:322-466).
//
// Discretisation choices of the adjoint:
//   * the implicit midpoint y1 = y0 + h F(x, (y0 + y1) / 2) is differentiated
//     as the implicit equation it solves, linearised at (y0 + y1) / 2: the
//     adjoint equation mu = lambda + (h/2) F_y^T mu is solved by fixed-point
//     passes to convergence (ADJ_FIXED_POINT_TOL);
//   * the Kepler drift is differentiated through the universal Kepler
//     equation at the recorded universal anomaly (implicit function theorem);
//   * the compensated-summation residues (whfast.rs:117-119) carry no
//     derivative (they are rounding errors);
//   * the spins entering the scalar products are L / I at the linearisation
//     point (the forward uses the spin of the previous midpoint iteration,
//     which differs from it by the convergence tolerance).
#pragma once

#include "whfast_step.hpp"

// The adjoint is compared with finite differences and the tangent-linear run,
// not bit for bit with anything, so its arithmetic (and only its: the forward
// step above keeps every multiply and add separately rounded) may fuse
// multiplies and adds.  The pragma covers every function defined from here to
// the matching contract(off) at the end of this header.
#pragma clang fp contract(fast)

#define POS_HD __device__ inline
#include "adjoint_math.hpp"

namespace posdev {

constexpr int ADJ_ROWS = 28;
// The adjoint midpoint equation mu = lambda + (h/2) F_y^T mu is iterated
// until mu changes by less than this (relative, per lane and block) or for
// at most ADJ_MAX_PASSES passes.  The contraction factor c is h |F_y| / 2:
// below 1e-7 for the forces, up to ~1e-4 for the tidal spin-down of close
// planets.  After a pass that changed mu by d the remaining error is about
// c d / (1 - c), so stopping at a change of 1e-10 leaves mu accurate to
// about 1e-14 relative at c = 1e-4 (two or three passes; a 1e-14 stop took
// one more pass and gave the same gradients to their last digits).
constexpr double ADJ_FIXED_POINT_TOL = 1e-10;
constexpr int ADJ_MAX_PASSES = 12;

struct AdjLds {
    double m[WAVE], I[WAVE], R[WAVE], sg[WAVE], k2[WAVE], kf[WAVE];
    double eta[WAVE];   // Jacobi: the mass of the host and the bodies up to this one in chain order
    // the last kick's fixed-point correction mu - lambda of each lane (v and
    // L parts), the next kick's starting guess; each lane uses its own column
    double wv[3][WAVE], wL[3][WAVE];
    double xb[ADJ_ROWS][WAVE];
};

struct AdjArgs {
    DevEnsemble E;                  // constants and launch table (state arrays unused)
    TapeRef T;                      // the segment's tape
    long long seg_start;            // index of the segment's first step in the recorded run
    int seg_len;                    // steps in the segment
    const long long *steps_done;    // [n_systems] steps each system executed in the recorded run
    double *lx, *lv, *lL;           // [3][n_bodies] cotangents: in = after the segment, out = before it
    double *gs, *gk2, *gkf;         // [n_bodies] accumulated parameter gradients
};

struct AdjLane {
    int lane, gbase, k, n, host, sys, body;
    bool valid, is_host, tides, flat, gr, spin, wind, jacobi;
    bool evolving;        // evolution tables: radius and moment of inertia follow time
    bool whds;            // WHDS coordinates
    double fv, wj;        // WHDS: velocity factor (m_h + m) / m_h and jump weight m / (m_h + m)
    int chain;            // position in the chain of the Jacobi recurrences: host 0, then the others in index order
    double eta_le, eta_lt;   // mass of the chain up to this body, with and without it
    double dt, m, I, R, sg, k2, kf, mh, Ih, M;
    double wk, wsat, wshape;   // this body's wind k factor, saturation and shape factor
    long long done;
    double lx[3], lv[3], lL[3];
    double gs, gk2, gkf;
    __device__ __forceinline__ bool orbiting() const { return valid && !is_host; }
    __device__ __forceinline__ int hl() const { return gbase + host; }
};

using posadj::V3;

__device__ __forceinline__ V3 ld3(const double a[3]) { return V3{a[0], a[1], a[2]}; }
__device__ __forceinline__ void st3(double a[3], V3 v) { a[0] = v.x; a[1] = v.y; a[2] = v.z; }

// Sums over the bodies of the group (order is irrelevant for the adjoint).
__device__ __forceinline__ V3 gsum3(const AdjLds &L, const AdjLane &A, int row, bool skip_host) {
    V3 s{0., 0., 0.};
    for (int j = 0; j < A.n; ++j) {
        if (skip_host && j == A.host) continue;
        const int l = A.gbase + j;
        s.x += L.xb[row][l]; s.y += L.xb[row + 1][l]; s.z += L.xb[row + 2][l];
    }
    return s;
}
__device__ __forceinline__ V3 gsum3m(const AdjLds &L, const AdjLane &A, int row, bool skip_host) {
    V3 s{0., 0., 0.};
    for (int j = 0; j < A.n; ++j) {
        if (skip_host && j == A.host) continue;
        const int l = A.gbase + j;
        const double mj = L.m[l];
        s.x += mj * L.xb[row][l]; s.y += mj * L.xb[row + 1][l]; s.z += mj * L.xb[row + 2][l];
    }
    return s;
}
__device__ __forceinline__ void put3(AdjLds &L, int row, int lane, V3 v) {
    L.xb[row][lane] = v.x; L.xb[row + 1][lane] = v.y; L.xb[row + 2][lane] = v.z;
}
__device__ __forceinline__ V3 get3(const AdjLds &L, int row, int lane) {
    return V3{L.xb[row][lane], L.xb[row + 1][lane], L.xb[row + 2][lane]};
}
// Position of body j in the chain of the Jacobi recurrences (whfast.rs:889-933):
// the host first, then the other bodies in index order.
__device__ __forceinline__ int chain_of(const AdjLane &A, int j) {
    return j == A.host ? 0 : (j < A.host ? j + 1 : j);
}

// One pass of the effects' vector-Jacobian product at the linearisation
// point published in rows 0-8 (x, v, W = L / I of every body): given the
// cotangents (mu_a, mu_L) of every body's additional acceleration and dL/dt,
// return this body's cotangents of v and L and, with FULL, of x and of its
// parameters (without FULL those are zero and the work only they need is
// dropped by the compiler).
template <bool FULL>
__device__ inline void effects_vjp_pass(AdjLds &L, const AdjLane &A, const posadj::PairConst &C,
                                        const posadj::PairPar &P, const posadj::PairIn &in, V3 mu_a, V3 mu_L,
                                        V3 &gx, V3 &gv, V3 &gL, double pg[3]) {
    const int lane = A.lane;
    if (A.valid) { put3(L, 9, lane, mu_a); put3(L, 12, lane, mu_L); }
    wsync();
    gx = V3{0., 0., 0.}; gv = gx; gL = gx;
    pg[0] = pg[1] = pg[2] = 0.;
    if (A.orbiting()) {
        const int hl = A.hl();
        const V3 mah = get3(L, 9, hl), mLh = get3(L, 12, hl);
        posadj::PairOut ob;
        ob.F = (1. / A.m) * mu_a - (1. / A.mh) * mah;
        ob.g = mu_a - (A.m / A.mh) * mah;
        ob.Np = mu_L;
        ob.Nh = mLh;
        posadj::PairIn Ib{V3{0., 0., 0.}, V3{0., 0., 0.}, V3{0., 0., 0.}, V3{0., 0., 0.}};
        posadj::PairParBar Pb{0., 0., 0., 0., 0., 0.};
        posadj::pair_vjp(C, P, in, ob, Ib, Pb);
        gv = Ib.u;
        gL = (1. / A.I) * Ib.Wp;
        put3(L, 18, lane, -Ib.u);
        put3(L, 21, lane, (1. / A.Ih) * Ib.Wh);
        if (FULL) {
            gx = Ib.r;
            pg[0] = Pb.sp; pg[1] = Pb.k2p; pg[2] = Pb.kfp;
            put3(L, 15, lane, -Ib.r);
            put3(L, 24, lane, V3{Pb.sh, Pb.k2h, Pb.kfh});
        }
    }
    wsync();
    if (A.valid && A.is_host) {
        gv = gsum3(L, A, 18, true);
        gL = gsum3(L, A, 21, true);
        if (FULL) {
            gx = gsum3(L, A, 15, true);
            const V3 p = gsum3(L, A, 24, true);
            pg[0] = p.x; pg[1] = p.y; pg[2] = p.z;
        }
    }
    if (A.wind) {
        // this body's own wind torque (wind.rs:61-91), -k W f shape with
        // f = wsat^2 once |W| reaches the saturation and |W|^2 below it:
        // its transpose Jacobian with respect to W, then W = L / I
        const V3 W = in.Wp;
        const double s2 = posadj::dot(W, W);
        const V3 dW = sqrt(s2) >= A.wsat ? (-A.wk * A.wsat * A.wsat * A.wshape) * mu_L
                                         : (-A.wk * A.wshape) * (s2 * mu_L + (2. * posadj::dot(W, mu_L)) * W);
        gL = gL + (1. / A.I) * dW;
    }
    wsync();
}

// True when every lane of the group passes `flag` (row 27 is the scratch row).
__device__ inline bool group_all(AdjLds &L, const AdjLane &A, bool flag) {
    L.xb[27][A.lane] = flag ? 1. : 0.;
    wsync();
    bool all = true;
    for (int j = 0; j < A.n; ++j) all = all && L.xb[27][A.gbase + j] != 0.;
    wsync();
    return all;
}

__device__ __forceinline__ double maxabs3(V3 a) { return fmax(fabs(a.x), fmax(fabs(a.y), fabs(a.z))); }

// |a - b| small against |a| (both zero counts as equal)
__device__ __forceinline__ bool settled(V3 a, V3 b) {
    return maxabs3(a - b) <= ADJ_FIXED_POINT_TOL * maxabs3(a);
}

// Adjoint of one implicit-midpoint kick (x fixed, y = (v, L), half step h):
// on entry A.lv / A.lL are the cotangents of y1 and A.lx that of x; on exit
// A.lv / A.lL are those of y0, A.lx has the kick's contribution added and the
// parameter gradients are accumulated.
__device__ inline void kick_adjoint(AdjLds &L, AdjLane &A, V3 x, V3 v0, V3 L0, V3 v1, V3 L1, double h) {
    const int lane = A.lane;
    const V3 vm = 0.5 * (v0 + v1);
    const V3 Lm = A.spin ? 0.5 * (L0 + L1) : L0;
    const V3 W = (1. / A.I) * Lm;
    if (A.valid) { put3(L, 0, lane, x); put3(L, 3, lane, vm); put3(L, 6, lane, W); }
    wsync();
    const int hl = A.hl();
    const posadj::PairConst C{A.mh, A.m, L.R[hl], A.R, A.Ih, A.I, A.tides, A.flat, A.gr};
    const posadj::PairPar P{L.sg[hl], A.sg, L.k2[hl], A.k2, L.kf[hl], A.kf};
    posadj::PairIn in{V3{0., 0., 0.}, V3{0., 0., 0.}, V3{0., 0., 0.}, W};
    if (A.orbiting()) {
        in.r = x - get3(L, 0, hl);
        in.u = vm - get3(L, 3, hl);
        in.Wh = get3(L, 6, hl);
    }
    const V3 lam_v = ld3(A.lv), lam_L = ld3(A.lL);
    const V3 z{0., 0., 0.};
    V3 gx, gv, gL;
    double pg[3];
    // fixed point of mu = lambda + (h/2) F_y^T mu, starting from lambda plus
    // the previous kick's correction (in the reverse sweep the kick before is
    // the other kick at the same positions or the next one along the orbit,
    // so the correction changes little and the start is closer than lambda by
    // about the contraction factor); the passes need only the v and L parts
    // of the product
    const V3 wv{L.wv[0][lane], L.wv[1][lane], L.wv[2][lane]};
    const V3 wL{L.wL[0][lane], L.wL[1][lane], L.wL[2][lane]};
    V3 mu_v = lam_v + wv, mu_L = A.spin ? lam_L + wL : z;
    for (int pass = 0;; ++pass) {
        effects_vjp_pass<false>(L, A, C, P, in, mu_v, mu_L, gx, gv, gL, pg);
        const V3 nv = lam_v + (0.5 * h) * gv;
        const V3 nL = A.spin ? lam_L + (0.5 * h) * gL : z;
        const bool conv = !A.valid || (settled(nv, mu_v) && settled(nL, mu_L));
        mu_v = nv;
        mu_L = nL;
        if (pass + 1 >= ADJ_MAX_PASSES || group_all(L, A, conv)) break;
    }
    {
        const V3 cv = mu_v - lam_v, cL = A.spin ? mu_L - lam_L : z;
        L.wv[0][lane] = cv.x; L.wv[1][lane] = cv.y; L.wv[2][lane] = cv.z;
        L.wL[0][lane] = cL.x; L.wL[1][lane] = cL.y; L.wL[2][lane] = cL.z;
    }
    // the x and parameter parts (and, without spin, the L part) at the fixed point
    effects_vjp_pass<true>(L, A, C, P, in, mu_v, mu_L, gx, gv, gL, pg);
    // y0's cotangent is mu + (h/2) F_y^T mu = 2 mu - lambda
    st3(A.lv, 2. * mu_v - lam_v);
    st3(A.lL, A.spin ? 2. * mu_L - lam_L : lam_L + h * gL);
    st3(A.lx, ld3(A.lx) + h * gx);
    A.gs += h * pg[0];
    A.gk2 += h * pg[1];
    A.gkf += h * pg[2];
}

__device__ __forceinline__ V3 tape3(const TapeRef &T, int row, int field, int body) {
    return V3{*T.at(row, field, body), *T.at(row, field + 1, body), *T.at(row, field + 2, body)};
}

// This step's radius and moment of inertia.  Under evolution tables they
// follow time and depend on no state, so they carry no derivative and come
// from the tape; without them they keep the values loaded at the segment's
// start.  The host's moment of inertia is read back after the next wsync.
// EVO: the launch holds systems with evolution tables (without them the
// reads are compiled out).
template <bool EVO>
__device__ __forceinline__ void step_radii(const TapeRef &T, AdjLds &L, AdjLane &A, int row) {
    if (!(EVO && A.valid && A.evolving)) return;
    A.R = *T.at(row, TP_R, A.body);
    A.I = *T.at(row, TP_I, A.body);
    L.R[A.lane] = A.R;
    L.I[A.lane] = A.I;
    if (A.wind) A.wshape = sqrt(A.R / R_SUN * 1. / A.m);
}

// Reverse of step `row` of the segment for democratic heliocentric
// coordinates or, with WHDS, for WHDS coordinates (whfast.rs:973-1023,
// 1090-1155, 495-556, 558-625): the same step with the velocities scaled by
// f = (m_h + m) / m_h, Kepler drifts about m_h + m, a jump weighted by
// w = m / (m_h + m) without the body's own term, and the interaction kick
// scaled by f.  A.l* hold the cotangents of the
// state after the step on entry and before it on exit.
template <bool EVO, bool WHDS>
__device__ inline void adjoint_step(const AdjArgs &Ar, AdjLds &L, AdjLane &A, int row) {
    const TapeRef &T = Ar.T;
    const int lane = A.lane;
    const int b = A.body;
    const V3 z{0., 0., 0.};
    V3 x0 = z, v0 = z, L0 = z, v1 = z, L1 = z, x2 = z, v2 = z, L2 = z;
    double X1 = 0., X2 = 0.;
    if (A.valid) {
        x0 = tape3(T, row, TP_X, b); v0 = tape3(T, row, TP_V, b); L0 = tape3(T, row, TP_L, b);
        v1 = tape3(T, row, TP_V1, b); L1 = tape3(T, row, TP_L1, b);
        X1 = *T.at(row, TP_XK1, b); X2 = *T.at(row, TP_XK2, b);
        x2 = tape3(T, row + 1, TP_X, b); v2 = tape3(T, row + 1, TP_V, b); L2 = tape3(T, row + 1, TP_L, b);
    }
    step_radii<EVO>(T, L, A, row);
    const double h = 0.5 * A.dt, dt = A.dt;
    const double mu = posadj::G_K2 * (WHDS ? A.mh + A.m : A.mh);
    const bool orb = A.orbiting();
    const int hl = A.hl();

    // ---------------- recompute the drift from (x0, v1) ----------------
    if (A.valid) { put3(L, 0, lane, x0); put3(L, 3, lane, v1); }
    wsync();
    if (EVO && A.evolving) A.Ih = L.I[hl];   // the host's moment of inertia of this step
    const V3 Xc = (1. / A.M) * gsum3m(L, A, 0, false);
    const V3 V = (1. / A.M) * gsum3m(L, A, 3, false);
    const V3 Q0 = orb ? x0 - get3(L, 0, hl) : z;
    const V3 P0 = orb ? (WHDS ? A.fv * (v1 - V) : v1 - V) : z;
    wsync();
    V3 Q1 = z, P1 = z;
    if (orb) posadj::kepler_forward_x(mu, h, X1, Q0, P0, Q1, P1);
    const V3 Xc1 = Xc + h * V;
    if (orb) put3(L, 0, lane, WHDS ? A.wj * P1 : P1);
    wsync();
    const V3 Q1p = !orb ? z
                 : WHDS ? Q1 + h * (gsum3(L, A, 0, true) - A.wj * P1)
                        : Q1 + (h / A.mh) * gsum3m(L, A, 0, true);
    wsync();
    if (orb) put3(L, 0, lane, Q1p);
    wsync();
    const V3 xh = Xc1 - (1. / A.M) * gsum3m(L, A, 0, true);
    const V3 xg = orb ? Q1p + xh : xh;
    wsync();
    if (A.valid) put3(L, 0, lane, xg);
    wsync();
    V3 acc = z;
    if (orb) {
        for (int j = 0; j < A.n; ++j) {
            if (j == A.host || j == A.k) continue;
            const int l = A.gbase + j;
            const V3 d = xg - get3(L, 0, l);
            const double d2 = posadj::dot(d, d);
            const double id = 1. / sqrt(d2);
            acc += (-posadj::G_K2 * L.m[l] * id * id * id) * d;
        }
    }
    const V3 P1pp = P1 + (WHDS ? dt * A.fv : dt) * acc;
    wsync();
    if (orb) put3(L, 0, lane, WHDS ? A.wj * P1pp : P1pp);
    wsync();
    const V3 Q2 = !orb ? z
                : WHDS ? Q1p + h * (gsum3(L, A, 0, true) - A.wj * P1pp)
                       : Q1p + (h / A.mh) * gsum3m(L, A, 0, true);
    V3 Q3 = z, P3 = z;
    if (orb) posadj::kepler_forward_x(mu, h, X2, Q2, P1pp, Q3, P3);
    wsync();
    if (orb) put3(L, 0, lane, WHDS ? A.wj * P3 : P3);
    wsync();
    const V3 v2p = WHDS ? (orb ? (1. / A.fv) * P3 + V : V - gsum3(L, A, 0, true))
                        : (orb ? P3 + V : V - (1. / A.mh) * gsum3m(L, A, 0, true));
    wsync();

    // ---------------- reverse ----------------
    // second kick (at x2, between (v2', L1) and (v2, L2))
    kick_adjoint(L, A, x2, v2p, L1, v2, L2, h);
    // T^-1 of the second drift
    if (A.valid) { put3(L, 0, lane, ld3(A.lx)); put3(L, 3, lane, ld3(A.lv)); }
    wsync();
    const V3 xh_tot = gsum3(L, A, 0, false);
    const V3 lvh = get3(L, 3, hl);
    const V3 Vb0 = lvh + gsum3(L, A, 3, true);
    const V3 Q3b = orb ? ld3(A.lx) - (A.m / A.M) * xh_tot : z;
    const V3 P3b = !orb ? z : WHDS ? (1. / A.fv) * ld3(A.lv) - A.wj * lvh : ld3(A.lv) - (A.m / A.mh) * lvh;
    wsync();
    // second Kepler drift and the centre-of-mass drift
    V3 Q2b = z, P1ppb = z;
    if (orb) posadj::kepler_vjp(mu, h, X2, Q2, P1pp, Q3b, P3b, Q2b, P1ppb);
    V3 Xc1b = xh_tot;
    V3 Vb = Vb0 + h * xh_tot;
    // second jump
    if (orb) put3(L, 0, lane, Q2b);
    wsync();
    if (orb) P1ppb += WHDS ? (h * A.wj) * (gsum3(L, A, 0, true) - Q2b) : (h * A.m / A.mh) * gsum3(L, A, 0, true);
    V3 Q1pb = Q2b;
    wsync();
    // interaction and planet-planet gravity
    V3 P1b = P1ppb;
    const V3 ab = (WHDS ? dt * A.fv : dt) * P1ppb;
    if (A.valid) { put3(L, 0, lane, orb ? ab : z); put3(L, 3, lane, xg); }
    wsync();
    V3 xgb = z;
    if (orb) {
        for (int j = 0; j < A.n; ++j) {
            if (j == A.host || j == A.k) continue;
            const int l = A.gbase + j;
            const V3 d = xg - get3(L, 3, l);
            xgb += posadj::gravity_pair_vjp(L.m[l], d, ab) - posadj::gravity_pair_vjp(A.m, d, get3(L, 0, l));
        }
    }
    wsync();
    // positions for gravity: x_p = Q1p_p + x_h, x_h = Xc1 - sum m Q1p / M
    if (A.valid) put3(L, 0, lane, xgb);
    wsync();
    const V3 xgh_tot = gsum3(L, A, 0, true);
    if (orb) Q1pb += xgb - (A.m / A.M) * xgh_tot;
    Xc1b += xgh_tot;
    wsync();
    // first jump
    if (orb) put3(L, 0, lane, Q1pb);
    wsync();
    if (orb) P1b += WHDS ? (h * A.wj) * (gsum3(L, A, 0, true) - Q1pb) : (h * A.m / A.mh) * gsum3(L, A, 0, true);
    wsync();
    // first Kepler drift and the centre-of-mass drift
    V3 Q0b = z, P0b = z;
    if (orb) posadj::kepler_vjp(mu, h, X1, Q0, P0, Q1pb, P1b, Q0b, P0b);
    const V3 Xcb = Xc1b;
    Vb += h * Xc1b;
    // T of the first drift
    const V3 v1b = WHDS ? A.fv * P0b : P0b;   // P0 = f (v1 - V), f = 1 in heliocentric coordinates
    if (orb) { put3(L, 0, lane, Q0b); put3(L, 3, lane, v1b); }
    wsync();
    const V3 SQ = gsum3(L, A, 0, true), SP = gsum3(L, A, 3, true);
    const V3 Vtot = Vb - SP;
    if (A.valid) {
        st3(A.lx, (orb ? Q0b : -SQ) + (A.m / A.M) * Xcb);
        st3(A.lv, (orb ? v1b : z) + (A.m / A.M) * Vtot);
    }
    wsync();
    // first kick (at x0, between (v0, L0) and (v1, L1))
    kick_adjoint(L, A, x0, v0, L0, v1, L1, h);
}

// ---- sums over the chain of the Jacobi recurrences (rows of 3 values) ----
// sum over the bodies before this one in the chain of m_j * row_j
__device__ __forceinline__ V3 chain_before_m(const AdjLds &L, const AdjLane &A, int row) {
    V3 s{0., 0., 0.};
    for (int j = 0; j < A.n; ++j) {
        if (chain_of(A, j) >= A.chain) continue;
        const int l = A.gbase + j;
        s += L.m[l] * get3(L, row, l);
    }
    return s;
}
// sum over the bodies up to this one in the chain (itself included) of row_j
__device__ __forceinline__ V3 chain_upto(const AdjLds &L, const AdjLane &A, int row) {
    V3 s{0., 0., 0.};
    for (int j = 0; j < A.n; ++j) {
        if (chain_of(A, j) > A.chain) continue;
        s += get3(L, row, A.gbase + j);
    }
    return s;
}
// sum over the non-host bodies from chain position c0 on of (m_j / eta_le_j) * row_j
__device__ __forceinline__ V3 chain_from_w(const AdjLds &L, const AdjLane &A, int row, int c0) {
    V3 s{0., 0., 0.};
    for (int j = 0; j < A.n; ++j) {
        if (j == A.host || chain_of(A, j) < c0) continue;
        const int l = A.gbase + j;
        s += (L.m[l] / L.eta[l]) * get3(L, row, l);
    }
    return s;
}
// sum over the non-host bodies after this one in the chain of row_j / eta_lt_j
__device__ __forceinline__ V3 chain_after_inv(const AdjLds &L, const AdjLane &A, int row) {
    V3 s{0., 0., 0.};
    for (int j = 0; j < A.n; ++j) {
        if (j == A.host || chain_of(A, j) <= A.chain) continue;
        const int l = A.gbase + j;
        s += (1. / (L.eta[l] - L.m[l])) * get3(L, row, l);
    }
    return s;
}

// Reverse of step `row` of the segment for a system in Jacobi coordinates
// (whfast.rs:235-305 with the transforms of 889-963, 1034-1080: no jump; the
// Kepler drifts take the chain's mass up to the body; the interaction adds
// the Jacobi transform of the accelerations, which leave out the pair of the
// host and the first body of the chain, and the indirect term of the bodies
// after the first).  The transforms are linear in the positions and
// velocities for fixed masses: with eta the chain's mass up to a body,
//   ax_j = x_j - sum_{i before j} m_i x_i / eta_lt_j,   com = sum m_i x_i / M,
//   x_j = ax_j + com - sum_{l from j on} (m_l / eta_le_l) ax_l,
//   x_host = com - sum_{all l} (m_l / eta_le_l) ax_l,
// and their vector-Jacobian products are the transposed sums.
template <bool EVO>
__device__ inline void adjoint_step_jacobi(const AdjArgs &Ar, AdjLds &L, AdjLane &A, int row) {
    const TapeRef &T = Ar.T;
    const int lane = A.lane;
    const int b = A.body;
    const V3 z{0., 0., 0.};
    V3 x0 = z, v0 = z, L0 = z, v1 = z, L1 = z, x2 = z, v2 = z, L2 = z;
    double X1 = 0., X2 = 0.;
    if (A.valid) {
        x0 = tape3(T, row, TP_X, b); v0 = tape3(T, row, TP_V, b); L0 = tape3(T, row, TP_L, b);
        v1 = tape3(T, row, TP_V1, b); L1 = tape3(T, row, TP_L1, b);
        X1 = *T.at(row, TP_XK1, b); X2 = *T.at(row, TP_XK2, b);
        x2 = tape3(T, row + 1, TP_X, b); v2 = tape3(T, row + 1, TP_V, b); L2 = tape3(T, row + 1, TP_L, b);
    }
    step_radii<EVO>(T, L, A, row);
    const double h = 0.5 * A.dt, dt = A.dt;
    const double mu = posadj::G_K2 * A.eta_le;
    const bool orb = A.orbiting();
    const int hl = A.hl();
    const int rank = A.chain - 1;        // 0 for the first body of the chain
    const int first = A.host == 0 ? 1 : 0;   // index of the first body of the chain

    // ---------------- recompute the drifts and the interaction ----------------
    // to Jacobi coordinates from (x0, v1)
    if (A.valid) { put3(L, 0, lane, x0); put3(L, 3, lane, v1); }
    wsync();
    if (EVO && A.evolving) A.Ih = L.I[hl];   // the host's moment of inertia of this step
    const V3 com = (1. / A.M) * gsum3m(L, A, 0, false);
    const V3 comv = (1. / A.M) * gsum3m(L, A, 3, false);
    const V3 Q0 = orb ? x0 - (1. / A.eta_lt) * chain_before_m(L, A, 0) : z;
    const V3 P0 = orb ? v1 - (1. / A.eta_lt) * chain_before_m(L, A, 3) : z;
    wsync();
    V3 Q1 = z, P1 = z;
    if (orb) posadj::kepler_forward_x(mu, h, X1, Q0, P0, Q1, P1);
    const V3 com1 = com + h * comv;
    // inertial positions of the interaction
    if (A.valid) put3(L, 0, lane, Q1);
    wsync();
    const V3 xg = Q1 + com1 - chain_from_w(L, A, 0, orb ? A.chain : 1);
    wsync();
    // gravitational accelerations (the host / first-body pair left out)
    if (A.valid) put3(L, 0, lane, xg);
    wsync();
    V3 acc = z;
    if (A.valid) {
        for (int j = 0; j < A.n; ++j) {
            if (j == A.k) continue;
            if ((A.k == A.host && j == first) || (j == A.host && A.k == first)) continue;
            const int l = A.gbase + j;
            const V3 d = xg - get3(L, 0, l);
            const double d2 = posadj::dot(d, d);
            const double id = 1. / sqrt(d2);
            acc += (-posadj::G_K2 * L.m[l] * id * id * id) * d;
        }
    }
    wsync();
    if (A.valid) put3(L, 0, lane, acc);
    wsync();
    const V3 aj = orb ? acc - (1. / A.eta_lt) * chain_before_m(L, A, 0) : z;
    wsync();
    // the indirect term of the bodies after the first (whfast.rs:955-962)
    double r2 = 1., r3i = 0.;
    if (orb && rank > 0) {
        r2 = posadj::dot(Q1, Q1) + 1e-12;
        r3i = 1. / (r2 * sqrt(r2));
    }
    const double gk = dt * posadj::G_K2 * A.eta_le;
    const V3 P1pp = orb ? P1 + dt * aj + (rank > 0 ? gk * r3i : 0.) * Q1 : z;
    V3 Q3 = z, P3 = z;
    if (orb) posadj::kepler_forward_x(mu, h, X2, Q1, P1pp, Q3, P3);
    // velocities after the second drift in inertial coordinates
    if (A.valid) put3(L, 0, lane, P3);
    wsync();
    const V3 v2p = P3 + comv - chain_from_w(L, A, 0, orb ? A.chain : 1);
    wsync();

    // ---------------- reverse ----------------
    // second kick (at x2, between (v2', L1) and (v2, L2))
    kick_adjoint(L, A, x2, v2p, L1, v2, L2, h);
    // back to inertial coordinates after the second drift
    if (A.valid) { put3(L, 0, lane, ld3(A.lx)); put3(L, 3, lane, ld3(A.lv)); }
    wsync();
    const V3 Q3b = orb ? ld3(A.lx) - (A.m / A.eta_le) * chain_upto(L, A, 0) : z;
    const V3 P3b = orb ? ld3(A.lv) - (A.m / A.eta_le) * chain_upto(L, A, 3) : z;
    V3 com1b = gsum3(L, A, 0, false);   // com2 = com1 + h comv
    V3 comvb = gsum3(L, A, 3, false) + h * com1b;
    wsync();
    // second Kepler drift
    V3 Q1b = z, P1ppb = z;
    if (orb) posadj::kepler_vjp(mu, h, X2, Q1, P1pp, Q3b, P3b, Q1b, P1ppb);
    // interaction: the indirect term and the Jacobi accelerations
    const V3 P1b = P1ppb;
    if (orb && rank > 0) Q1b += (gk * r3i) * P1ppb - (3. * gk * r3i / r2 * posadj::dot(Q1, P1ppb)) * Q1;
    if (A.valid) put3(L, 0, lane, orb ? dt * P1ppb : z);
    wsync();
    const V3 accb = (orb ? dt * P1ppb : z) - A.m * chain_after_inv(L, A, 0);
    wsync();
    // gravity
    if (A.valid) { put3(L, 0, lane, accb); put3(L, 3, lane, xg); }
    wsync();
    V3 xgb = z;
    if (A.valid) {
        for (int j = 0; j < A.n; ++j) {
            if (j == A.k) continue;
            if ((A.k == A.host && j == first) || (j == A.host && A.k == first)) continue;
            const int l = A.gbase + j;
            const V3 d = xg - get3(L, 3, l);
            xgb += posadj::gravity_pair_vjp(L.m[l], d, accb) - posadj::gravity_pair_vjp(A.m, d, get3(L, 0, l));
        }
    }
    wsync();
    // the interaction's inertial positions from (Q1, com1)
    if (A.valid) put3(L, 0, lane, xgb);
    wsync();
    if (orb) Q1b += xgb - (A.m / A.eta_le) * chain_upto(L, A, 0);
    com1b += gsum3(L, A, 0, false);
    wsync();
    // first Kepler drift (com1 = com + h comv)
    V3 Q0b = z, P0b = z;
    if (orb) posadj::kepler_vjp(mu, h, X1, Q0, P0, Q1b, P1b, Q0b, P0b);
    comvb += h * com1b;
    // to Jacobi coordinates from (x0, v1)
    if (A.valid) { put3(L, 0, lane, Q0b); put3(L, 3, lane, P0b); }
    wsync();
    if (A.valid) {
        st3(A.lx, Q0b - A.m * chain_after_inv(L, A, 0) + (A.m / A.M) * com1b);
        st3(A.lv, P0b - A.m * chain_after_inv(L, A, 3) + (A.m / A.M) * comvb);
    }
    wsync();
    // first kick (at x0, between (v0, L0) and (v1, L1))
    kick_adjoint(L, A, x0, v0, L0, v1, L1, h);
}

__device__ inline void adjoint_load(const AdjArgs &Ar, AdjLds &L, AdjLane &A, int wave) {
    const DevEnsemble &E = Ar.E;
    const int lane = threadIdx.x;
    const LaneGroup lg = lane_group(E, wave, lane);
    const int k = lg.k;
    const bool gvalid = lg.gvalid;
    const int s = lg.sys;
    A.lane = lane;
    A.gbase = lg.gbase;
    A.k = k;
    A.sys = s;
    A.n = gvalid ? (E.body_offset[s + 1] - E.body_offset[s]) : 0;
    A.valid = gvalid && k < A.n;
    A.host = gvalid ? E.host[s] : 0;
    A.is_host = A.valid && k == A.host;
    A.body = E.body_offset[s] + (A.valid ? k : 0);
    const uint32_t eff = E.effects[s];
    A.tides = eff & POS_EFFECT_TIDES;
    A.flat = eff & POS_EFFECT_ROTATIONAL_FLATTENING;
    A.gr = eff & POS_EFFECT_GENERAL_RELATIVITY;
    A.spin = eff & (POS_EFFECT_TIDES | POS_EFFECT_ROTATIONAL_FLATTENING | POS_EFFECT_GENERAL_RELATIVITY | POS_EFFECT_EVOLUTION);
    A.evolving = eff & POS_EFFECT_EVOLUTION;
    // (the recording run admits no GR variant other than Kidder 1995)
    A.dt = E.dt[s];
    A.done = Ar.steps_done[s];
    const int b = A.body;
    const int nb = E.n_bodies;
    A.m = A.valid ? E.mass[b] : 0.;
    A.I = E.inertia[b];
    A.R = E.radius[b];
    A.sg = E.sigma[b];
    A.k2 = E.k2[b];
    A.kf = E.k2f[b];
    // the wind torque enters dL/dt only where spins are integrated, as in
    // the forward step (whfast.rs:259-262)
    A.wind = A.valid && A.spin && (eff & POS_EFFECT_WIND) && (E.roles[b] & POS_ROLE_WIND);
    A.wk = A.wind ? E.wind_k[b] : 0.;
    A.wsat = A.wind ? E.wind_sat[b] : 0.;
    A.wshape = A.wind ? sqrt(A.R / R_SUN * 1. / A.m) : 0.;
    L.m[lane] = A.m;
    L.I[lane] = A.I;
    L.R[lane] = A.R;
    L.sg[lane] = A.sg;
    L.k2[lane] = A.k2;
    L.kf[lane] = A.kf;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        A.lx[q] = A.valid ? Ar.lx[q * nb + b] : 0.;
        A.lv[q] = A.valid ? Ar.lv[q * nb + b] : 0.;
        A.lL[q] = A.valid ? Ar.lL[q * nb + b] : 0.;
    }
    A.gs = A.valid ? Ar.gs[b] : 0.;
    A.gk2 = A.valid ? Ar.gk2[b] : 0.;
    A.gkf = A.valid ? Ar.gkf[b] : 0.;
    __syncthreads();
    A.mh = L.m[A.hl()];
    A.Ih = L.I[A.hl()];
    double M = 0.;
    for (int j = 0; j < A.n; ++j) M += L.m[A.gbase + j];
    A.M = M;
    A.jacobi = E.coord[s] == POS_COORD_JACOBI;
    A.whds = E.coord[s] == POS_COORD_WHDS;
    A.fv = (A.mh + A.m) / A.mh;
    A.wj = A.m / (A.mh + A.m);
    A.chain = chain_of(A, A.k);
    double eta = 0.;
    for (int j = 0; j < A.n; ++j)
        if (chain_of(A, j) <= A.chain) eta += L.m[A.gbase + j];
    A.eta_le = eta;
    A.eta_lt = eta - A.m;
    L.eta[lane] = eta;
    for (int q = 0; q < 3; ++q) { L.wv[q][lane] = 0.; L.wL[q][lane] = 0.; }
    __syncthreads();
}

__device__ inline void adjoint_store(const AdjArgs &Ar, const AdjLane &A) {
    if (!A.valid) return;
    const int b = A.body;
    const int nb = Ar.E.n_bodies;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        Ar.lx[q * nb + b] = A.lx[q];
        Ar.lv[q * nb + b] = A.lv[q];
        Ar.lL[q * nb + b] = A.lL[q];
    }
    Ar.gs[b] = A.gs;
    Ar.gk2[b] = A.gk2;
    Ar.gkf[b] = A.gkf;
}

}  // namespace posdev

#pragma clang fp contract(off)
