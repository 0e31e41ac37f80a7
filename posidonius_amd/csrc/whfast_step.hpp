// whfast_step.hpp -- device implementation of one WHFast step for a lane group.
//
// Mapping: a wavefront of 64 lanes holds 64/G systems; the G lanes of a group
// hold the bodies of one system, lane k <-> body k (k < n).  Every body-local
// quantity lives in the lane's registers.  Quantities that couple bodies go
// through a per-wavefront LDS exchange area:
//   * sums over bodies (centre of mass, momentum, host reactions, the
//     implicit-midpoint convergence norms) are re-accumulated by every lane
//     of the group from the exchange area, strictly in body-index order, so
//     the result is the same rounded value the sequential reference computes;
//   * the host's state is read by the orbiting bodies from the same area.
// All arithmetic is binary64 and this file must be compiled with
// -ffp-contract=off so that each operation rounds exactly as in the
// reference (src/integrator/whfast.rs, src/particles/universe.rs,
// src/effects/*).  Operation order inside every formula follows the
// reference; Rust's powi(n) with constant n is the square-and-multiply
// chain reproduced by p2..p10.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/posidonius_hip.h"

namespace posdev {

// ---- constants (src/constants.rs) ----------------------------------------
constexpr double DBL_EPS = 2.2204460492503131e-16;
constexpr double DBL_EPS_2 = DBL_EPS * DBL_EPS;
constexpr double PI = 3.14159265358979323846264338327950288;
constexpr double DAY = 24. * 3600.;
constexpr double M_SUN = 1.9818e30;
constexpr double G_SI = 6.67428e-11;
constexpr double AU = 1.49597870700e11;
constexpr double SPEED_OF_LIGHT = (2.99792458e8 / AU) * DAY;
constexpr double SPEED_OF_LIGHT_2 = SPEED_OF_LIGHT * SPEED_OF_LIGHT;
constexpr double R_SUN = 6.957e8 / AU;
constexpr double G = G_SI / (AU * AU * AU) * M_SUN * (DAY * DAY);
constexpr double K2 = G;
constexpr double MAX_DISTANCE_2 = 100. * 100.;
constexpr int NMAX_QUART = 64;
constexpr int NMAX_NEWT = 32;
constexpr int MIDPOINT_MIN_ITER = 3;
constexpr int MIDPOINT_MAX_ITER = 10;
constexpr int STUMPFF_REDUCTION_CAP = 600;
constexpr int BISECTION_CAP = 4096;
constexpr int WAVE = 64;

__device__ __forceinline__ double p2(double x) { return x * x; }
__device__ __forceinline__ double p5(double x) { double x2 = x * x; double x4 = x2 * x2; return x * x4; }
__device__ __forceinline__ double p7(double x) { double x2 = x * x; double x4 = x2 * x2; return (x * x2) * x4; }
__device__ __forceinline__ double p8(double x) { double x2 = x * x; double x4 = x2 * x2; return x4 * x4; }
__device__ __forceinline__ double p10(double x) { double x2 = x * x; double x4 = x2 * x2; double x8 = x4 * x4; return x2 * x8; }

// ---- device view of an ensemble (structure of arrays, device order) -------
struct DevEnsemble {
    int n_systems, n_bodies;
    // per system
    const int *body_offset;   // [n_systems + 1]
    const int *host;          // most massive body (index within the system)
    const int *coord;
    const uint32_t *effects;
    const double *dt, *time_limit, *hist_period;
    double *t, *last_hist;
    long long *iteration;
    uint32_t *status;
    int *timestep_warning;
    // per body
    double *x, *v, *L, *spin, *err_v, *err_L, *acc;   // [3][n_bodies] component-major
    double *radius, *rg2, *inertia;                    // [n_bodies]
    const double *mass, *sigma, *k2, *k2f, *wind_k, *wind_sat;
    const uint32_t *roles;
    const double *roche;                               // [n_bodies][POS_MAX_BODIES]
    // evolution
    const int *evo_type, *evo_off, *evo_len;
    int *left_index;
    const double *tab_time, *tab_radius, *tab_rg2, *tab_love;
    // launch table: wave w covers device systems wave_first[w] .. +wave_count[w]-1
    const int *wave_first, *wave_count, *wave_group;
    int n_waves;
};

// ---- LDS exchange area of one wavefront -------------------------------------
constexpr int NX = 26;   // exchange slots per lane
struct WaveLds {
    double m[WAVE], mg[WAVE];     // constant per body
    double xb[NX][WAVE];          // phase-dependent exchange
};

__device__ __forceinline__ void wsync() { __syncthreads(); }

// ---- per-lane state ---------------------------------------------------------
struct Lane {
    // identity
    int lane, gbase, k, n, host, coord, sys, body;
    bool valid, is_host;
    uint32_t eff, roles;
    bool tides, rf, gr, wind, evolution, spin_on;
    // system scalars (identical copies in every lane of the group)
    double t, dt, half_dt, T, last_hist, hist_period;
    long long iteration;
    uint32_t status;
    int tw;
    // body
    double m, mg, R, rg2, I, sigma, k2, k2f, wk, wsat, wsat2;
    double x[3], v[3], L[3], s[3], s2, ev[3], eL[3], ag[3];
    // host constants (masses never change)
    double hm, hmg, hsigma, hk2, hk2f;
    // Jacobi constants
    double jac_mu;      // G * cumulative mass up to this body (whfast.rs:642-651)
    double jac_eta;     // cumulative mass up to this body (whfast.rs:574-576)
    int jac_rank;       // position among the non-host bodies
    // evolution
    int evo, evo_off, evo_len, left_index;
    // alternative coordinates of this body; com* hold the host slot (centre of mass)
    double ax[3], av[3], aacc[3];
    double comx[3], comv[3], com_m;
};

// ===========================================================================
// Evolution tables (evolution.rs:458-546, tools.rs:840-905)
// ===========================================================================
__device__ inline double interp_table(const double *tx, const double *ty, int n, double target, int *left_out) {
    int first = -1;
    for (int i = 0; i < n; ++i) {
        if (tx[i] > target) { first = i; break; }
    }
    int left, right;
    if (first < 0) {
        if (tx[n - 1] > target) { left = right = 0; } else { left = right = n - 1; }
    } else if (first == 0) {
        left = right = 0;
    } else {
        left = first - 1; right = first;
    }
    *left_out = left;
    if (left == right) return ty[left];
    const double x_left = tx[left];
    const double p = (target - x_left) / (tx[right] - x_left);
    return ty[left] * (1. - p) + ty[right] * p;
}

// Evolver::idx() (evolution.rs:449-456); clamped to the last row, which a
// consistent left_index never reaches.
__device__ __forceinline__ int evolver_idx(int left_index, int len) {
    const int i0 = left_index > 10 ? left_index - 10 : 0;
    return i0 < len ? i0 : len - 1;
}

__device__ inline void evolve_body(const DevEnsemble &E, Lane &B, double t) {
    if (!B.valid || B.evo == POS_EVO_NON_EVOLVING) return;
    const double *tt = E.tab_time + B.evo_off;
    int i0 = evolver_idx(B.left_index, B.evo_len);
    int left = 0;
    double newR = interp_table(tt + i0, E.tab_radius + B.evo_off + i0, B.evo_len - i0, t, &left);
    B.left_index += left;
    double newrg2 = B.rg2;
    if (B.evo == POS_EVO_BARAFFE2015 || B.evo == POS_EVO_LECONTE2011 || B.evo == POS_EVO_LECONTE_CHABRIER2013) {
        i0 = evolver_idx(B.left_index, B.evo_len);
        newrg2 = interp_table(tt + i0, E.tab_rg2 + B.evo_off + i0, B.evo_len - i0, t, &left);
        B.left_index += left;
    }
    if (newR != B.R || newrg2 != B.rg2) {
        B.R = newR;
        B.rg2 = newrg2;
        B.I = B.m * B.rg2 * p2(B.R);
    }
    if (B.evo == POS_EVO_LECONTE_CHABRIER2013 && (B.roles & (POS_ROLE_TIDES_CENTRAL | POS_ROLE_TIDES_ORBITING))) {
        i0 = evolver_idx(B.left_index, B.evo_len);
        (void)interp_table(tt + i0, E.tab_love + B.evo_off + i0, B.evo_len - i0, t, &left);
        B.left_index += left;
    }
}

// particles/common.rs:3-14 for this lane's body
__device__ __forceinline__ void own_spin(Lane &B) {
    if (!B.valid) return;
    if (B.I == 0.) { B.status |= POS_STATUS_ZERO_INERTIA; return; }
    B.s[0] = B.L[0] / B.I;
    B.s[1] = B.L[1] / B.I;
    B.s[2] = B.L[2] / B.I;
    B.s2 = p2(B.s[0]) + p2(B.s[1]) + p2(B.s[2]);
}

// The zero-inertia flag is raised by whichever lane meets it; every lane of
// the group must agree on the status word, so it is OR-reduced through LDS
// at the end of each step.

// ===========================================================================
// Additional effects + implicit midpoint (whfast.rs:322-466, universe.rs:353-614)
// ===========================================================================
// Exchange slots used inside one implicit-midpoint iteration.
enum {
    X_X0 = 0, X_X1, X_X2, X_V0, X_V1, X_V2, X_S0, X_S1, X_S2,   // round 1
    X_FS0 = 9, X_FS1, X_FS2, X_FS2N, X_R, X_I, X_L0, X_L1, X_L2  // round 2
};
enum {   // round 3: orbiting body -> host contributions
    C_TF = 0,    // tidal force (3)
    C_RF = 3,    // flattening force (3)
    C_TN = 6,    // -1 * tidal torque on the host (3)
    C_RN = 9,    // -1 * flattening torque on the host (3)
    C_G1 = 12,   // (m/M) * 1PN acceleration (3)
    C_G2 = 15,   // (m/M) * 2PN acceleration (3)
    C_G3 = 18,   // (m/M) * spin-orbit acceleration (3)
    C_GL = 21,   // host dL/dt contribution from GR (3)
    C_COUNT = 24
};

__device__ inline void additional_effects(const DevEnsemble &E, WaveLds &lds, Lane &B, bool evolution,
                                          double a_add[3], double dLdt[3]) {
    const int lane = B.lane;
    // ---- round 1: inertial state and the (stale) spin of every body ----
    if (B.valid) {
        lds.xb[X_X0][lane] = B.x[0]; lds.xb[X_X1][lane] = B.x[1]; lds.xb[X_X2][lane] = B.x[2];
        lds.xb[X_V0][lane] = B.v[0]; lds.xb[X_V1][lane] = B.v[1]; lds.xb[X_V2][lane] = B.v[2];
        lds.xb[X_S0][lane] = B.s[0]; lds.xb[X_S1][lane] = B.s[1]; lds.xb[X_S2][lane] = B.s[2];
    }
    wsync();
    const int hl = B.gbase + B.host;
    double px[3] = {0., 0., 0.}, pv[3] = {0., 0., 0.};
    double d = 0., vr = 0., nv2 = 0., nv = 0.;
    double rs_star = 0., rs_planet = 0.;
    if (B.valid && !B.is_host) {
        // universe.rs:318-351 (the host's heliocentric velocity is zero here)
        px[0] = B.x[0] - lds.xb[X_X0][hl]; px[1] = B.x[1] - lds.xb[X_X1][hl]; px[2] = B.x[2] - lds.xb[X_X2][hl];
        pv[0] = B.v[0] - lds.xb[X_V0][hl]; pv[1] = B.v[1] - lds.xb[X_V1][hl]; pv[2] = B.v[2] - lds.xb[X_V2][hl];
        d = sqrt(p2(px[0]) + p2(px[1]) + p2(px[2]));
        vr = (px[0] * pv[0] + px[1] * pv[1] + px[2] * pv[2]) / d;
        nv2 = p2(pv[0]) + p2(pv[1]) + p2(pv[2]);
        nv = sqrt(nv2);
        // tides/common.rs:143-170 and rotational_flattening/common.rs:93-120:
        // scalar products with the spins of the previous evaluation
        rs_star = px[0] * lds.xb[X_S0][hl] + px[1] * lds.xb[X_S1][hl] + px[2] * lds.xb[X_S2][hl];
        rs_planet = px[0] * B.s[0] + px[1] * B.s[1] + px[2] * B.s[2];
    }
    wsync();
    // ---- evolution and spin = L / I (universe.rs:305-316) ----
    if (evolution && B.evolution) evolve_body(E, B, B.t);
    own_spin(B);
    // ---- round 2: fresh host spin, radius, inertia, angular momentum ----
    if (B.valid) {
        lds.xb[X_FS0][lane] = B.s[0]; lds.xb[X_FS1][lane] = B.s[1]; lds.xb[X_FS2][lane] = B.s[2];
        lds.xb[X_FS2N][lane] = B.s2; lds.xb[X_R][lane] = B.R; lds.xb[X_I][lane] = B.I;
    }
    wsync();
    double hs[3], hs2 = 0., hR = 0., hI = 0.;
    hs[0] = lds.xb[X_FS0][hl]; hs[1] = lds.xb[X_FS1][hl]; hs[2] = lds.xb[X_FS2][hl];
    hs2 = lds.xb[X_FS2N][hl]; hR = lds.xb[X_R][hl]; hI = lds.xb[X_I][hl];
    wsync();

    double td_acc[3] = {0., 0., 0.}, td_dL[3] = {0., 0., 0.};
    double rf_acc[3] = {0., 0., 0.}, rf_dL[3] = {0., 0., 0.};
    double gr_acc[3] = {0., 0., 0.}, gr_dL[3] = {0., 0., 0.};
    double wind_dL[3] = {0., 0., 0.};
    double c[C_COUNT];
#pragma unroll
    for (int q = 0; q < C_COUNT; ++q) c[q] = 0.;

    // ---- wind (wind.rs:61-91), own body ----
    if (B.valid && B.wind && B.spin_on && (B.roles & POS_ROLE_WIND)) {
        const double threshold = sqrt(B.s2);
        const double shape = sqrt(B.R / R_SUN * 1. / B.m);
        if (threshold >= B.wsat) {
            wind_dL[0] = -1. * B.wk * B.s[0] * B.wsat2 * shape;
            wind_dL[1] = -1. * B.wk * B.s[1] * B.wsat2 * shape;
            wind_dL[2] = -1. * B.wk * B.s[2] * B.wsat2 * shape;
        } else {
            wind_dL[0] = -1. * B.wk * B.s[0] * B.s2 * shape;
            wind_dL[1] = -1. * B.wk * B.s[1] * B.s2 * shape;
            wind_dL[2] = -1. * B.wk * B.s[2] * B.s2 * shape;
        }
    }

    if (B.valid && !B.is_host) {
        const double m = B.m, hm = B.hm;
        // ---- tides: orthogonal components (constant_time_lag.rs:206-264) ----
        double os = 0., op = 0.;
        if (B.tides) {
            const double d7 = p7(d);
            os = 4.5 * p2(m) * p10(hR) * B.hsigma / d7;
            op = 4.5 * p2(hm) * p10(B.R) * B.sigma / d7;
        }
        // ---- flattening: orthogonal components (oblate_spheroid.rs:21-49) ----
        double fs = 0., fp = 0., ofs = 0., ofp = 0.;
        if (B.rf) {
            fs = m * B.hk2f * hs2 * p5(hR) / 6.;
            ofs = -6. * fs * rs_star / (hs2 * p5(d));
            fp = hm * B.k2f * B.s2 * p5(B.R) / 6.;
            ofp = -6. * fp * rs_planet / (B.s2 * p5(d));
        }
        if (B.tides) {
            // radial component (constant_time_lag.rs:266-307)
            const double hm2 = hm * hm;
            const double pm2 = m * m;
            const double conservative = (-3.0 * K2) / p7(d) * (pm2 * p5(hR) * B.hk2 + hm2 * p5(B.R) * B.k2);
            const double factor1 = -13.5 * vr / p8(d);
            const double term1 = pm2 * p10(hR) * B.hsigma;
            const double term2 = hm2 * p10(B.R) * B.sigma;
            const double diss_pm = factor1 * term2;
            const double dissipative = diss_pm + factor1 * term1;
            const double radial = conservative + dissipative;
            // force (constant_time_lag.rs:309-332)
            const double factor3 = radial + (os + op) * vr / d;
            double F[3];
            F[0] = factor3 * px[0] / d + os / d * (hs[1] * px[2] - hs[2] * px[1] - pv[0])
                 + op / d * (B.s[1] * px[2] - B.s[2] * px[1] - pv[0]);
            F[1] = factor3 * px[1] / d + os / d * (hs[2] * px[0] - hs[0] * px[2] - pv[1])
                 + op / d * (B.s[2] * px[0] - B.s[0] * px[2] - pv[1]);
            F[2] = factor3 * px[2] / d + os / d * (hs[0] * px[1] - hs[1] * px[0] - pv[2])
                 + op / d * (B.s[0] * px[1] - B.s[1] * px[0] - pv[2]);
            const double f1m = 1. / m;
            td_acc[0] = f1m * F[0]; td_acc[1] = f1m * F[1]; td_acc[2] = f1m * F[2];
            c[C_TF] = F[0]; c[C_TF + 1] = F[1]; c[C_TF + 2] = F[2];
        }
        if (B.rf) {
            // radial component (oblate_spheroid.rs:51-60) and force (83-97)
            const double radial = -3. / p5(d) * (fp + fs)
                + 15. / p7(d) * (fs * rs_star * rs_star / hs2 + fp * rs_planet * rs_planet / B.s2);
            double F[3];
            F[0] = radial * px[0] + ofp * B.s[0] + ofs * hs[0];
            F[1] = radial * px[1] + ofp * B.s[1] + ofs * hs[1];
            F[2] = radial * px[2] + ofp * B.s[2] + ofs * hs[2];
            const double f1m = 1. / m;
            rf_acc[0] = f1m * F[0]; rf_acc[1] = f1m * F[1]; rf_acc[2] = f1m * F[2];
            c[C_RF] = F[0]; c[C_RF + 1] = F[1]; c[C_RF + 2] = F[2];
        }
        if (B.spin_on && B.tides) {
            // torques (constant_time_lag.rs:171-204, tides/common.rs:223-261)
            const double cx = px[1] * pv[2] - px[2] * pv[1];
            const double cy = px[2] * pv[0] - px[0] * pv[2];
            const double cz = px[0] * pv[1] - px[1] * pv[0];
            double N0 = op * (d * B.s[0] - rs_planet * px[0] / d - 1.0 / d * cx);
            double N1 = op * (d * B.s[1] - rs_planet * px[1] / d - 1.0 / d * cy);
            double N2 = op * (d * B.s[2] - rs_planet * px[2] / d - 1.0 / d * cz);
            td_dL[0] = -1.0 * N0; td_dL[1] = -1.0 * N1; td_dL[2] = -1.0 * N2;
            N0 = os * (d * hs[0] - rs_star * px[0] / d - 1.0 / d * cx);
            N1 = os * (d * hs[1] - rs_star * px[1] / d - 1.0 / d * cy);
            N2 = os * (d * hs[2] - rs_star * px[2] / d - 1.0 / d * cz);
            c[C_TN] = -1.0 * N0; c[C_TN + 1] = -1.0 * N1; c[C_TN + 2] = -1.0 * N2;
        }
        if (B.spin_on && B.rf) {
            // torques (oblate_spheroid.rs:62-81, rotational_flattening/common.rs:165-201)
            double N0 = ofp * (px[1] * B.s[2] - px[2] * B.s[1]);
            double N1 = ofp * (px[2] * B.s[0] - px[0] * B.s[2]);
            double N2 = ofp * (px[0] * B.s[1] - px[1] * B.s[0]);
            rf_dL[0] = -1.0 * N0; rf_dL[1] = -1.0 * N1; rf_dL[2] = -1.0 * N2;
            N0 = ofs * (px[1] * hs[2] - px[2] * hs[1]);
            N1 = ofs * (px[2] * hs[0] - px[0] * hs[2]);
            N2 = ofs * (px[0] * hs[1] - px[1] * hs[0]);
            c[C_RN] = -1.0 * N0; c[C_RN + 1] = -1.0 * N1; c[C_RN + 2] = -1.0 * N2;
        }
        if (B.gr) {
            // Kidder 1995 (general_relativity.rs:177-456)
            const double f = B.hmg * B.mg / p2(B.hmg + B.mg);
            const double mg = B.hmg + B.mg;
            const double distance_2 = p2(d);
            const double mOverM = m / hm;
            {   // 1PN
                const double radial_velocity_2 = p2(vr);
                const double radial = -mg / (distance_2 * SPEED_OF_LIGHT_2)
                    * ((1.0 + 3.0 * f) * nv2 - 2.0 * (2.0 + f) * mg / d - 1.5 * f * radial_velocity_2);
                const double orthogonal = mg / (distance_2 * SPEED_OF_LIGHT_2) * 2.0 * (2.0 - f) * vr * nv;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double a = radial * px[q] / d + orthogonal * pv[q] / nv;
                    gr_acc[q] = a;
                    c[C_G1 + q] = mOverM * a;
                }
            }
            {   // 2PN
                const double nv4 = p2(nv2);
                const double vr2 = p2(vr);
                const double vr4 = p2(vr2);
                const double f2 = p2(f);
                const double radial = -mg / (distance_2 * SPEED_OF_LIGHT_2)
                    * (3.0 / 4.0 * (12.0 + 29.0 * f) * (p2(mg) / distance_2)
                       + f * (3.0 - 4.0 * f) * nv4
                       + 15.0 / 8.0 * f * (1.0 - 3.0 * f) * vr4
                       - 3.0 / 2.0 * f * (3.0 - 4.0 * f) * vr2 * nv2
                       - 0.5 * f * (13.0 - 4.0 * f) * (mg / d) * nv2
                       - (2.0 + 25.0 * f + 2.0 * f2) * (mg / d) * vr2);
                const double orthogonal = -mg / (distance_2 * SPEED_OF_LIGHT_2) * (-0.5) * vr
                    * (f * (15.0 + 4.0 * f) * nv2
                       - (4.0 + 41.0 * f + 8.0 * f2) * (mg / d)
                       - 3.0 * f * (3.0 + 2.0 * f) * vr2);
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double a = radial * px[q] / d + orthogonal * pv[q];
                    gr_acc[q] = gr_acc[q] + a;
                    c[C_G2 + q] = mOverM * a;
                }
            }
            {   // 1.5PN spin-orbit
                const double Ls[3] = {hI * hs[0], hI * hs[1], hI * hs[2]};
                const double M = hm + m;
                const double mass_factor = (hm - m) / M;
                const double Lp[3] = {B.I * B.s[0], B.I * B.s[1], B.I * B.s[2]};
                const double n3[3] = {px[0] / d, px[1] / d, px[2] / d};
                double msf[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) msf[q] = mass_factor * M * (Lp[q] / m - Ls[q] / hm);
                double e1[3];
                e1[0] = 6. * n3[0] * ((n3[1] * pv[2] - n3[2] * pv[1]) * (2. * (Ls[0] + Lp[0]) + msf[0]));
                e1[1] = 6. * n3[1] * ((n3[2] * pv[0] - n3[0] * pv[2]) * (2. * (Ls[1] + Lp[1]) + msf[1]));
                e1[2] = 6. * n3[2] * ((n3[0] * pv[1] - n3[1] * pv[0]) * (2. * (Ls[2] + Lp[2]) + msf[2]));
                double e7[3], e3s[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    e7[q] = 7. * (Ls[q] + Lp[q]) + 3. * msf[q];
                    e3s[q] = 3. * (Ls[q] + Lp[q]) + msf[q];
                }
                const double e2[3] = {pv[1] * e7[2] - pv[2] * e7[1], pv[2] * e7[0] - pv[0] * e7[2], pv[0] * e7[1] - pv[1] * e7[0]};
                double e3[3];
                e3[0] = 3. * vr * (n3[1] * e3s[2] - n3[2] * e3s[1]);
                e3[1] = 3. * vr * (n3[2] * e3s[0] - n3[0] * e3s[2]);
                e3[2] = 3. * vr * (n3[0] * e3s[1] - n3[1] * e3s[0]);
                const double factor_a = G / SPEED_OF_LIGHT_2;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double a = factor_a * (e1[q] - e2[q] + e3[q]);
                    gr_acc[q] = gr_acc[q] + a;
                    c[C_G3 + q] = mOverM * a;
                }
                const double mu = (hm * m) / M;
                const double Lo[3] = {mu * (px[1] * pv[2] - px[2] * pv[1]), mu * (px[2] * pv[0] - px[0] * pv[2]),
                                      mu * (px[0] * pv[1] - px[1] * pv[0])};
                double fm = 2. + 3. / 2. * m / hm;
                const double sp = n3[0] * Lp[0] + n3[1] * Lp[1] + n3[2] * Lp[2];
                c[C_GL + 0] = factor_a * (fm * (Lo[1] * Ls[2] - Lo[2] * Ls[1]) - (Lp[1] * Ls[2] - Lp[2] * Ls[1]) + 3. * sp * (n3[1] * Ls[2] - n3[2] * Ls[1]));
                c[C_GL + 1] = factor_a * (fm * (Lo[2] * Ls[0] - Lo[0] * Ls[2]) - (Lp[2] * Ls[0] - Lp[0] * Ls[2]) + 3. * sp * (n3[2] * Ls[0] - n3[0] * Ls[2]));
                c[C_GL + 2] = factor_a * (fm * (Lo[0] * Ls[1] - Lo[1] * Ls[0]) - (Lp[0] * Ls[1] - Lp[1] * Ls[0]) + 3. * sp * (n3[0] * Ls[1] - n3[1] * Ls[0]));
                fm = 2. + 3. / 2. * hm / m;
                const double sh = n3[0] * Ls[0] + n3[1] * Ls[1] + n3[2] * Ls[2];
                gr_dL[0] = factor_a * (fm * (Lo[1] * Lp[2] - Lo[2] * Lp[1]) - (Ls[1] * Lp[2] - Ls[2] * Lp[1]) + 3. * sh * (n3[1] * Lp[2] - n3[2] * Lp[1]));
                gr_dL[1] = factor_a * (fm * (Lo[2] * Lp[0] - Lo[0] * Lp[2]) - (Ls[2] * Lp[0] - Ls[0] * Lp[2]) + 3. * sh * (n3[2] * Lp[0] - n3[0] * Lp[2]));
                gr_dL[2] = factor_a * (fm * (Lo[0] * Lp[1] - Lo[1] * Lp[0]) - (Ls[0] * Lp[1] - Ls[1] * Lp[0]) + 3. * sh * (n3[0] * Lp[1] - n3[1] * Lp[0]));
            }
        }
    }
    // ---- round 3: host reactions, accumulated in body order ----
    if (B.valid && !B.is_host) {
#pragma unroll
        for (int q = 0; q < C_COUNT; ++q) lds.xb[q][lane] = c[q];
    }
    wsync();
    if (B.valid && B.is_host) {
        double sT[3] = {0., 0., 0.}, sR[3] = {0., 0., 0.}, sTN[3] = {0., 0., 0.}, sRN[3] = {0., 0., 0.};
        double s1[3] = {0., 0., 0.}, s2[3] = {0., 0., 0.}, s3[3] = {0., 0., 0.}, sGL[3] = {0., 0., 0.};
        for (int j = 0; j < B.n; ++j) {
            if (j == B.host) continue;
            const int l = B.gbase + j;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                sT[q] = sT[q] + lds.xb[C_TF + q][l];
                sR[q] = sR[q] + lds.xb[C_RF + q][l];
                sTN[q] = sTN[q] + lds.xb[C_TN + q][l];
                sRN[q] = sRN[q] + lds.xb[C_RN + q][l];
                s1[q] = s1[q] + lds.xb[C_G1 + q][l];
                s2[q] = s2[q] + lds.xb[C_G2 + q][l];
                s3[q] = s3[q] + lds.xb[C_G3 + q][l];
                sGL[q] = sGL[q] + lds.xb[C_GL + q][l];
            }
        }
        const double factor2 = 1. / B.m;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (B.tides) { td_acc[q] = -1.0 * factor2 * sT[q]; if (B.spin_on) td_dL[q] = sTN[q]; }
            if (B.rf) { rf_acc[q] = -1.0 * factor2 * sR[q]; if (B.spin_on) rf_dL[q] = sRN[q]; }
            if (B.gr) {
                gr_acc[q] = -1.0 * s1[q];
                gr_acc[q] = gr_acc[q] + -1.0 * s2[q];
                gr_acc[q] = gr_acc[q] + -1.0 * s3[q];
                gr_dL[q] = sGL[q];
            }
        }
    }
    wsync();
    // ---- outputs (universe.rs:540-614) ----
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        dLdt[q] = td_dL[q] + rf_dL[q] + gr_dL[q] + wind_dL[q];
        double a = 0.;
        if (B.tides) a = a + td_acc[q];
        if (B.rf) a = a + rf_acc[q];
        if (B.gr) a = a + gr_acc[q];
        a_add[q] = a;
    }
}

// Ordered sum over the group of slot `q` (every lane obtains the same value).
__device__ __forceinline__ double group_sum(const WaveLds &lds, const Lane &B, int q) {
    double s = 0.;
    for (int j = 0; j < B.n; ++j) s = s + lds.xb[q][B.gbase + j];
    return s;
}

__device__ inline void integrate_velocity_dependent_forces(const DevEnsemble &E, WaveLds &lds, Lane &B, bool evolution) {
    const double h = B.half_dt;
    double v_orig[3], L_orig[3], v_fin[3], L_fin[3], dv[3] = {0., 0., 0.}, dL[3] = {0., 0., 0.};
#pragma unroll
    for (int q = 0; q < 3; ++q) { v_orig[q] = B.v[q]; L_orig[q] = B.L[q]; v_fin[q] = B.v[q]; L_fin[q] = B.L[q]; }
    bool converged = false;
    for (int it = 0; it < MIDPOINT_MAX_ITER; ++it) {
        double v_prev[3], L_prev[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) { v_prev[q] = v_fin[q]; L_prev[q] = L_fin[q]; }
        double a_add[3], dLdt[3];
        additional_effects(E, lds, B, evolution && it == 0, a_add, dLdt);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            dv[q] = h * a_add[q] - B.ev[q];
            v_fin[q] = v_orig[q] + dv[q];
            if (B.spin_on) {
                dL[q] = h * dLdt[q] - B.eL[q];
                L_fin[q] = L_orig[q] + dL[q];
            }
        }
        if (it >= MIDPOINT_MIN_ITER - 1) {
            // whfast.rs:424-451, sums in body order
            if (B.valid) {
                const double dvx = v_fin[0] - v_prev[0], dvy = v_fin[1] - v_prev[1], dvz = v_fin[2] - v_prev[2];
                lds.xb[0][B.lane] = p2(dvx) + p2(dvy) + p2(dvz);
                lds.xb[1][B.lane] = p2(v_fin[0]) + p2(v_fin[1]) + p2(v_fin[2]);
                const double dsx = L_fin[0] - L_prev[0], dsy = L_fin[1] - L_prev[1], dsz = L_fin[2] - L_prev[2];
                lds.xb[2][B.lane] = p2(dsx) + p2(dsy) + p2(dsz);
                lds.xb[3][B.lane] = p2(L_fin[0]) + p2(L_fin[1]) + p2(L_fin[2]);
            }
            wsync();
            const double dv2 = group_sum(lds, B, 0), fv2 = group_sum(lds, B, 1);
            const double dL2 = group_sum(lds, B, 2), fL2 = group_sum(lds, B, 3);
            wsync();
            const bool ok = (B.spin_on && dv2 / fv2 < DBL_EPS_2 && dL2 / fL2 < DBL_EPS_2) ||
                            (!B.spin_on && dv2 / fv2 < DBL_EPS_2);
            if (ok) { converged = true; break; }
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            B.v[q] = 0.5 * (v_orig[q] + v_fin[q]);
            if (B.spin_on) B.L[q] = 0.5 * (L_orig[q] + L_fin[q]);
        }
    }
    if (!converged) B.status |= POS_STATUS_MIDPOINT_NOT_CONVERGED;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        B.v[q] = v_fin[q];
        B.ev[q] = (B.v[q] - v_orig[q]) - dv[q];
        if (B.spin_on) {
            B.L[q] = L_fin[q];
            B.eL[q] = (B.L[q] - L_orig[q]) - dL[q];
        }
    }
}

// ===========================================================================
// Kepler drift (whfast.rs:676-876)
// ===========================================================================
__device__ inline void stumpff_cs3(double z, double cs[4], uint32_t &status) {
    const double invf[14] = {1., 1., 1. / 2., 1. / 6., 1. / 24., 1. / 120., 1. / 720., 1. / 5040.,
                             1. / 40320., 1. / 362880., 1. / 3628800., 1. / 39916800.,
                             1. / 479001600., 1. / 6227020800.};
    int n = 0;
    while (fabs(z) > 0.1) {
        if (n >= STUMPFF_REDUCTION_CAP) { status |= POS_STATUS_ITERATION_CAP; break; }
        z = z / 4.;
        n += 1;
    }
    double c_odd = invf[13];
    double c_even = invf[12];
#pragma unroll
    for (int np = 11; np >= 3; np -= 2) {
        c_odd = invf[np] - z * c_odd;
        c_even = invf[np - 1] - z * c_even;
    }
    cs[3] = c_odd;
    cs[2] = c_even;
    cs[1] = invf[1] - z * c_odd;
    cs[0] = invf[0] - z * c_even;
    while (n > 0) {
        cs[3] = (cs[2] + cs[0] * cs[3]) * 0.25;
        cs[2] = cs[1] * cs[1] * 0.5;
        cs[1] = cs[0] * cs[1];
        cs[0] = 2. * cs[0] * cs[0] - 1.;
        n -= 1;
    }
}

__device__ __forceinline__ void stiefel_gs3(double beta, double x, double gs[4], uint32_t &status) {
    const double x2 = x * x;
    stumpff_cs3(beta * x2, gs, status);
    gs[1] = gs[1] * x;
    gs[2] = gs[2] * x2;
    gs[3] = gs[3] * (x2 * x);
}

__device__ inline void kepler_individual_step(Lane &B, double mass_g, double dt) {
    const double x0[3] = {B.ax[0], B.ax[1], B.ax[2]};
    const double v0[3] = {B.av[0], B.av[1], B.av[2]};
    const double r0 = sqrt(p2(x0[0]) + p2(x0[1]) + p2(x0[2]));
    const double r0i = 1. / r0;
    const double v2 = p2(v0[0]) + p2(v0[1]) + p2(v0[2]);
    const double beta = 2. * mass_g * r0i - v2;
    const double eta0 = x0[0] * v0[0] + x0[1] * v0[1] + x0[2] * v0[2];
    const double zeta0 = mass_g - beta * r0;
    double x, gs[4];
    double invperiod = 0.;
    double x_per_period;
    const bool elliptic = beta > 0.;
    if (elliptic) {
        const double sqrt_beta = sqrt(beta);
        invperiod = sqrt_beta * beta / (2. * PI * mass_g);
        x_per_period = (2. * PI) / sqrt_beta;
        if (fabs(dt) * invperiod > 1. && B.tw == 0) {
            B.status |= POS_STATUS_TIMESTEP_WARNING;
            B.tw += 1;
        }
        const double dtr0i = dt * r0i;
        x = dtr0i * (1. - dtr0i * eta0 * 0.5 * r0i);
    } else {
        x = 0.;
        x_per_period = __builtin_nan("");
    }
    int converged = 0;
    double old_x = x;
    stiefel_gs3(beta, x, gs, B.status);
    double e = eta0 * gs[1] + zeta0 * gs[2];
    double ri = 1. / (r0 + e);
    x = ri * (x * e - eta0 * gs[2] - zeta0 * gs[3] + dt);
    if (!__builtin_isnan(x_per_period) && fabs(x - old_x) > 0.01 * x_per_period) {
        // Laguerre-Conway (rare: only when the step is a sizeable part of an orbit)
        x = beta * dt / mass_g;
        double prev_x[NMAX_QUART + 1];
        for (int q = 0; q <= NMAX_QUART; ++q) prev_x[q] = 0.;
        for (int n_lag = 1; n_lag < NMAX_QUART; ++n_lag) {
            stiefel_gs3(beta, x, gs, B.status);
            const double f = r0 * x + eta0 * gs[2] + zeta0 * gs[3] - dt;
            const double fp = r0 + eta0 * gs[1] + zeta0 * gs[2];
            const double fpp = eta0 * gs[0] + zeta0 * gs[1];
            const double denom = fp + sqrt(fabs(16. * fp * fp - 20. * f * fpp));
            x = (x * denom - 5. * f) / denom;
            bool hit = false;
            for (int q = 1; q < n_lag; ++q)
                if (x == prev_x[q]) { hit = true; break; }
            if (hit) { converged = 1; break; }
            prev_x[n_lag] = x;
        }
        e = eta0 * gs[1] + zeta0 * gs[2];
        ri = 1. / (r0 + e);
    } else {
        double old_x2;
        for (int q = 1; q < NMAX_NEWT; ++q) {
            old_x2 = old_x;
            old_x = x;
            stiefel_gs3(beta, x, gs, B.status);
            e = eta0 * gs[1] + zeta0 * gs[2];
            ri = 1. / (r0 + e);
            x = ri * (x * e - eta0 * gs[2] - zeta0 * gs[3] + dt);
            if (x == old_x || x == old_x2) { converged = 1; break; }
        }
    }
    if (converged == 0) {
        B.status |= POS_STATUS_KEPLER_BISECTION;
        double x_min, x_max;
        if (elliptic) {
            x_min = x_per_period * floor(dt * invperiod);
            x_max = x_min + x_per_period;
        } else {
            const double h2 = r0 * r0 * v2 - eta0 * eta0;
            const double q = h2 / mass_g / (1. + sqrt(1. - h2 * beta / (mass_g * mass_g)));
            const double vq = sqrt(h2) / q;
            x_min = 1. / (vq + r0 / dt);
            x_max = dt / q;
        }
        x = (x_max + x_min) / 2.;
        int it = 0;
        for (;;) {
            if (++it > BISECTION_CAP) { B.status |= POS_STATUS_ITERATION_CAP; break; }
            stiefel_gs3(beta, x, gs, B.status);
            const double s = r0 * x + eta0 * gs[2] + zeta0 * gs[3] - dt;
            if (s >= 0.) x_max = x; else x_min = x;
            x = (x_max + x_min) / 2.;
            if (fabs(x_max - x_min) / x_max <= 1e-15) break;
        }
        e = eta0 * gs[1] + zeta0 * gs[2];
        ri = 1. / (r0 + e);
    }
    if (__builtin_isnan(ri)) {
        ri = 0.; gs[1] = 0.; gs[2] = 0.; gs[3] = 0.;
    }
    const double f = -mass_g * gs[2] * r0i;
    const double g = dt - mass_g * gs[3];
    const double fd = -mass_g * gs[1] * r0i * ri;
    const double gd = -mass_g * gs[2] * ri;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        B.ax[q] = B.ax[q] + (f * x0[q] + g * v0[q]);
        B.av[q] = B.av[q] + (fd * x0[q] + gd * v0[q]);
    }
}

__device__ inline void kepler_steps(Lane &B, double dt) {
    if (B.valid && !B.is_host) {
        double mu;
        if (B.coord == POS_COORD_JACOBI) mu = B.jac_mu;
        else if (B.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) mu = B.hmg;
        else mu = B.hmg + B.mg;
        kepler_individual_step(B, mu, dt);
    }
    // the host slot (centre of mass) drifts in a straight line (whfast.rs:634-639);
    // every lane keeps the same copy
#pragma unroll
    for (int q = 0; q < 3; ++q) B.comx[q] = B.comx[q] + dt * B.comv[q];
}

// ===========================================================================
// Coordinate transforms, jump and interaction (whfast.rs:495-625, 881-1155)
// ===========================================================================
__device__ inline void inertial_to_alternative(WaveLds &lds, Lane &B) {
    const int lane = B.lane;
    if (B.valid) {
        lds.xb[0][lane] = B.x[0]; lds.xb[1][lane] = B.x[1]; lds.xb[2][lane] = B.x[2];
        lds.xb[3][lane] = B.v[0]; lds.xb[4][lane] = B.v[1]; lds.xb[5][lane] = B.v[2];
    }
    wsync();
    const int hl = B.gbase + B.host;
    if (B.coord == POS_COORD_JACOBI) {
        // whfast.rs:889-933: prefix recurrence, recomputed by every lane up to itself
        double eta = lds.m[hl];
        double s[3], sv[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) { s[q] = eta * lds.xb[q][hl]; sv[q] = eta * lds.xb[3 + q][hl]; }
        for (int j = 0; j < B.n; ++j) {
            if (j == B.host) continue;
            const int l = B.gbase + j;
            const double mj = lds.m[l];
            const double ei = 1. / eta;
            eta = eta + mj;
            const double pme = eta * ei;
            double axj[3], avj[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                axj[q] = lds.xb[q][l] - s[q] * ei;
                avj[q] = lds.xb[3 + q][l] - sv[q] * ei;
            }
            if (j == B.k) {
#pragma unroll
                for (int q = 0; q < 3; ++q) { B.ax[q] = axj[q]; B.av[q] = avj[q]; }
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                s[q] = s[q] * pme + mj * axj[q];
                sv[q] = sv[q] * pme + mj * avj[q];
            }
        }
        const double mtot_i = 1. / eta;
        B.com_m = eta;
#pragma unroll
        for (int q = 0; q < 3; ++q) { B.comx[q] = s[q] * mtot_i; B.comv[q] = sv[q] * mtot_i; }
    } else {
        // whfast.rs:973-1023: centre of mass, host first then the others in order
        double cx[3] = {0., 0., 0.}, cv[3] = {0., 0., 0.}, cm = 0.;
        for (int jj = -1; jj < B.n; ++jj) {
            const int j = jj < 0 ? B.host : jj;
            if (jj >= 0 && jj == B.host) continue;
            const int l = B.gbase + j;
            const double mj = lds.m[l];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                cx[q] = cx[q] + lds.xb[q][l] * mj;
                cv[q] = cv[q] + lds.xb[3 + q][l] * mj;
            }
            cm = cm + mj;
        }
        B.com_m = cm;
#pragma unroll
        for (int q = 0; q < 3; ++q) { B.comx[q] = cx[q] / cm; B.comv[q] = cv[q] / cm; }
        if (B.valid && !B.is_host) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                B.ax[q] = B.x[q] - lds.xb[q][hl];
                B.av[q] = B.v[q] - B.comv[q];
            }
            if (B.coord == POS_COORD_WHDS) {
                const double factor = (B.hm + B.m) / B.hm;
#pragma unroll
                for (int q = 0; q < 3; ++q) B.av[q] = B.av[q] * factor;
            }
        }
    }
    wsync();
}

__device__ inline void jump_step(WaveLds &lds, Lane &B, double dt) {
    if (B.coord == POS_COORD_JACOBI) return;
    const int lane = B.lane;
    if (B.valid && !B.is_host) {
        lds.xb[0][lane] = B.av[0]; lds.xb[1][lane] = B.av[1]; lds.xb[2][lane] = B.av[2];
    }
    wsync();
    const double m0 = B.hm;
    double p[3] = {0., 0., 0.};
    for (int j = 0; j < B.n; ++j) {
        if (j == B.host) continue;
        const int l = B.gbase + j;
        const double mj = lds.m[l];
        if (B.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) {
#pragma unroll
            for (int q = 0; q < 3; ++q) p[q] = p[q] + mj * lds.xb[q][l];
        } else {
            const double factor = m0 + mj;
#pragma unroll
            for (int q = 0; q < 3; ++q) p[q] = p[q] + mj * lds.xb[q][l] / factor;
        }
    }
    wsync();
    if (B.valid && !B.is_host) {
        if (B.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) {
#pragma unroll
            for (int q = 0; q < 3; ++q) B.ax[q] = B.ax[q] + dt * p[q] / m0;
        } else {
            const double factor = m0 + B.m;
#pragma unroll
            for (int q = 0; q < 3; ++q) B.ax[q] = B.ax[q] + dt * (p[q] - (B.m * B.av[q] / factor));
        }
    }
}

__device__ inline void interaction_step(WaveLds &lds, Lane &B, double dt) {
    if (B.coord == POS_COORD_JACOBI) {
        // whfast.rs:935-963 then 566-592
        const int lane = B.lane;
        if (B.valid) { lds.xb[0][lane] = B.ag[0]; lds.xb[1][lane] = B.ag[1]; lds.xb[2][lane] = B.ag[2]; }
        wsync();
        const int hl = B.gbase + B.host;
        double eta = lds.m[hl];
        double s[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) s[q] = eta * lds.xb[q][hl];
        for (int j = 0; j < B.n; ++j) {
            if (j == B.host) continue;
            const int l = B.gbase + j;
            const double mj = lds.m[l];
            const double ei = 1. / eta;
            eta = eta + mj;
            const double pme = eta * ei;
            double aj[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) aj[q] = lds.xb[q][l] - s[q] * ei;
            if (j == B.k) {
#pragma unroll
                for (int q = 0; q < 3; ++q) B.aacc[q] = aj[q];
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) s[q] = s[q] * pme + mj * aj[q];
        }
        wsync();
        if (B.valid && !B.is_host) {
#pragma unroll
            for (int q = 0; q < 3; ++q) B.av[q] = B.av[q] + dt * B.aacc[q];
            if (B.jac_rank > 0) {
                const double rj2i = 1. / (p2(B.ax[0]) + p2(B.ax[1]) + p2(B.ax[2]) + 1e-12);
                const double rji = sqrt(rj2i);
                const double rj3im = rji * rj2i * G * B.jac_eta;
                const double prefac = dt * rj3im;
#pragma unroll
                for (int q = 0; q < 3; ++q) B.av[q] = B.av[q] + prefac * B.ax[q];
            }
        }
        return;
    }
    if (B.valid && !B.is_host) {
        if (B.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) {
#pragma unroll
            for (int q = 0; q < 3; ++q) B.av[q] = B.av[q] + dt * B.ag[q];
        } else {
            const double factor = B.hm + B.m;
#pragma unroll
            for (int q = 0; q < 3; ++q) B.av[q] = B.av[q] + dt * factor * B.ag[q] / B.hm;
        }
    }
}

__device__ inline void alternative_to_inertial(WaveLds &lds, Lane &B) {
    const int lane = B.lane;
    if (B.valid && !B.is_host) {
        lds.xb[0][lane] = B.ax[0]; lds.xb[1][lane] = B.ax[1]; lds.xb[2][lane] = B.ax[2];
        lds.xb[3][lane] = B.av[0]; lds.xb[4][lane] = B.av[1]; lds.xb[5][lane] = B.av[2];
    }
    wsync();
    if (B.coord == POS_COORD_JACOBI) {
        // whfast.rs:1034-1080, reverse recurrence
        double eta = B.com_m;
        double s[3], sv[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) { s[q] = eta * B.comx[q]; sv[q] = eta * B.comv[q]; }
        for (int j = B.n - 1; j >= 0; --j) {
            if (j == B.host) continue;
            const int l = B.gbase + j;
            const double mj = lds.m[l];
            const double ei = 1. / eta;
            double xj[3], vj[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                s[q] = (s[q] - mj * lds.xb[q][l]) * ei;
                sv[q] = (sv[q] - mj * lds.xb[3 + q][l]) * ei;
                xj[q] = lds.xb[q][l] + s[q];
                vj[q] = lds.xb[3 + q][l] + sv[q];
            }
            if (j == B.k) {
#pragma unroll
                for (int q = 0; q < 3; ++q) { B.x[q] = xj[q]; B.v[q] = vj[q]; }
            }
            eta = eta - mj;
#pragma unroll
            for (int q = 0; q < 3; ++q) { s[q] = s[q] * eta; sv[q] = sv[q] * eta; }
        }
        if (B.valid && B.is_host) {
            const double mtot_i = 1. / eta;
#pragma unroll
            for (int q = 0; q < 3; ++q) { B.x[q] = s[q] * mtot_i; B.v[q] = sv[q] * mtot_i; }
        }
    } else {
        // whfast.rs:1090-1155
        const double mtot = B.com_m;
        double nsp[3] = {B.comx[0], B.comx[1], B.comx[2]};
        double nsv[3] = {B.comv[0], B.comv[1], B.comv[2]};
        const double m0 = B.hm;
        for (int j = 0; j < B.n; ++j) {
            if (j == B.host) continue;
            const int l = B.gbase + j;
            const double mj = lds.m[l];
#pragma unroll
            for (int q = 0; q < 3; ++q) nsp[q] = nsp[q] - lds.xb[q][l] * mj / mtot;
        }
        for (int j = 0; j < B.n; ++j) {
            if (j == B.host) continue;
            const int l = B.gbase + j;
            const double mj = lds.m[l];
            const double factor = (B.coord == POS_COORD_WHDS) ? mj / (m0 + mj) : mj / m0;
#pragma unroll
            for (int q = 0; q < 3; ++q) nsv[q] = nsv[q] - lds.xb[3 + q][l] * factor;
        }
        if (B.valid) {
            if (B.is_host) {
#pragma unroll
                for (int q = 0; q < 3; ++q) { B.x[q] = nsp[q]; B.v[q] = nsv[q]; }
            } else {
#pragma unroll
                for (int q = 0; q < 3; ++q) B.x[q] = B.ax[q] + nsp[q];
                if (B.coord == POS_COORD_WHDS) {
                    const double f = (m0 + B.m) / m0;
#pragma unroll
                    for (int q = 0; q < 3; ++q) B.v[q] = B.av[q] / f + B.comv[q];
                } else {
#pragma unroll
                    for (int q = 0; q < 3; ++q) B.v[q] = B.av[q] + B.comv[q];
                }
            }
        }
    }
    wsync();
}

// ===========================================================================
// Newtonian gravity (universe.rs:198-303)
// ===========================================================================
__device__ inline void gravity(const DevEnsemble &E, WaveLds &lds, Lane &B) {
    const int lane = B.lane;
    if (B.valid) { lds.xb[0][lane] = B.x[0]; lds.xb[1][lane] = B.x[1]; lds.xb[2][lane] = B.x[2]; lds.xb[3][lane] = B.R; }
    wsync();
    if (B.valid) {
        const int i = B.k, mm = B.host;
        const bool whfast_one = (B.coord == POS_COORD_JACOBI);
        double acc[3] = {0., 0., 0.};
        for (int j = 0; j < B.n; ++j) {
            if (j == i) continue;
            const int l = B.gbase + j;
            const double dx = B.x[0] - lds.xb[0][l];
            const double dy = B.x[1] - lds.xb[1][l];
            const double dz = B.x[2] - lds.xb[2][l];
            const double distance_2 = dx * dx + dy * dy + dz * dz;
            if (i < j) {
                const double rr = E.roche[(size_t)B.body * POS_MAX_BODIES + j];
                if (distance_2 <= rr * rr) B.status |= POS_STATUS_ROCHE;
                if (distance_2 <= p2(B.R + lds.xb[3][l])) B.status |= POS_STATUS_COLLISION;
                if (i == mm && distance_2 > MAX_DISTANCE_2) B.status |= POS_STATUS_EJECTION;
            }
            if (whfast_one && ((i == mm && mm == 0 && j == 1) || (i == mm && mm > 0 && j == 0) ||
                               (j == mm && mm == 0 && i == 1) || (j == mm && mm > 0 && i == 0)))
                continue;
            if (!whfast_one && (i == mm || j == mm)) continue;
            const double distance = sqrt(distance_2);
            const double prefact = -G / (distance * distance * distance) * lds.m[l];
            acc[0] = acc[0] + prefact * dx;
            acc[1] = acc[1] + prefact * dy;
            acc[2] = acc[2] + prefact * dz;
        }
        B.ag[0] = acc[0]; B.ag[1] = acc[1]; B.ag[2] = acc[2];
    }
    wsync();
}

// Status bits raised by one lane (gravity checks, zero inertia, Kepler) are
// OR-ed over the group so that every lane agrees.
__device__ inline void merge_status(WaveLds &lds, Lane &B) {
    lds.xb[0][B.lane] = __builtin_bit_cast(double, (unsigned long long)B.status);
    wsync();
    uint32_t st = 0;
    int tw = 0;
    for (int j = 0; j < B.n; ++j) {
        st |= (uint32_t)__builtin_bit_cast(unsigned long long, lds.xb[0][B.gbase + j]);
    }
    lds.xb[1][B.lane] = (double)B.tw;
    wsync();
    for (int j = 0; j < B.n; ++j) tw = max(tw, (int)lds.xb[1][B.gbase + j]);
    wsync();
    B.status = st;
    B.tw = tw;
}

// ===========================================================================
// WHFast::iterate (whfast.rs:235-305)
// ===========================================================================
__device__ inline void iterate(const DevEnsemble &E, WaveLds &lds, Lane &B) {
    const bool first = B.last_hist < 0.;
    const bool trig = B.last_hist + B.hist_period <= B.t;
    if (first || trig) {
        if (B.evolution) evolve_body(E, B, B.t);
        own_spin(B);
        if (!first) B.last_hist += B.hist_period;
        else B.last_hist = 0.;
    }
    integrate_velocity_dependent_forces(E, lds, B, true);
    inertial_to_alternative(lds, B);
    kepler_steps(B, B.half_dt);
    jump_step(lds, B, B.half_dt);
    alternative_to_inertial(lds, B);
    gravity(E, lds, B);
    interaction_step(lds, B, B.dt);
    jump_step(lds, B, B.half_dt);
    kepler_steps(B, B.half_dt);
    alternative_to_inertial(lds, B);
    integrate_velocity_dependent_forces(E, lds, B, false);
    B.t = B.t + B.dt;
    B.iteration += 1;
    merge_status(lds, B);
    if (B.t + B.dt > B.T) B.status |= POS_STATUS_COMPLETED;
}

}  // namespace posdev
