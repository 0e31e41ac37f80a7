// whfast_step.hpp -- device implementation of one WHFast step for a lane group.
//
// Mapping: a wavefront of 64 lanes holds 64/G systems; the G lanes of a group
// hold the bodies of one system, lane k <-> body k (k < n).
//
// Storage per lane:
//   * registers: the hot state (position, velocity, spin angular momentum,
//     compensated-summation errors, mass, time step);
//   * a persistent per-wavefront LDS area: per-body parameters, the spin of
//     the last evaluation (the reference reads it "stale", see
//     additional_effects), radius, moment of inertia and the gravitational
//     acceleration.  Orbiting bodies read their host's values from the host
//     lane's slot;
//   * global memory: per-system bookkeeping touched once per step.
// Quantities that couple bodies go through an exchange area of the same LDS
// block.  Sums over bodies (centre of mass, momentum, host reactions, the
// implicit-midpoint convergence norms) are re-accumulated strictly in
// body-index order, so every result is the same rounded value the sequential
// reference computes.
//
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
#include "crmath.hpp"

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
constexpr double SUN_DYN_FREQ_2 = K2 / (R_SUN * R_SUN * R_SUN);
constexpr double SMOOTHING_FACTOR_DYN_TIDE_COROTATION = 1.0e-5 * 86400.;
constexpr double BOLTZMANN_CONSTANT = 1.380649e-23 / (M_SUN * AU * AU) * DAY * DAY;
constexpr double MASS_HYDROGEN_ATOM = 1.672162841e-27 / M_SUN;
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

// ---- division by a shared divisor ------------------------------------------
// Many quotients of one evaluation share a divisor (the distance, |v|, a mass,
// the moment of inertia).  For a divisor b with correctly rounded reciprocal
// r = RN(1/b), q0 = RN(a*r) is within one ulp of a/b, the residual
// e = a - b*q0 is exact in one FMA, and RN(q0 + e*r) is the correctly rounded
// quotient RN(a/b) (Markstein's theorem; Muller et al., Handbook of
// Floating-Point Arithmetic, 2nd ed., sec. 4.7), barring overflow and
// underflow, which the step's magnitudes are far from.  The result is thus
// bit-identical to a/b at about a third of the instruction cost.  The sign is
// taken from q0: RN(a r) has the sign of a/b (also when it underflows to
// zero), and so does the corrected quotient, except that q0 = -0 with a zero
// residual would come back as +0 from the FMA; copysign restores it (one
// bitfield insert instead of a compare and a select).
struct Rcp {
    double b, r;
};
__device__ __forceinline__ Rcp rcp(double b) { return Rcp{b, 1.0 / b}; }
__device__ __forceinline__ double divr(double a, const Rcp &R) {
    const double q0 = a * R.r;
    const double e = __builtin_fma(-R.b, q0, a);
    return __builtin_copysign(__builtin_fma(e, R.r, q0), q0);
}

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
    const double *sigma_scale, *sigma_factor;          // the two factors of sigma
    const uint32_t *roles;
    const double *roche;                               // [n_bodies][POS_MAX_BODIES]
    // evolution
    const int *evo_type, *evo_off, *evo_len;
    int *left_index;
    const double *tab_time, *tab_radius, *tab_rg2, *tab_love, *tab_itq;
    // evolving dissipation and disk: lag angle of each body (evolution.rs:548-567),
    // migration timescale of its last disk evaluation (disk.rs:236), and the
    // disk host's DiskProperties per system (6 doubles)
    double *lag, *disk_migration;
    double *pair_ph;   // an orbiting body's own pair-dependent factor map entry, NaN = none
    const double *disk_props;
    // launch table: wave w covers device systems wave_first[w] .. +wave_count[w]-1
    // in groups of wave_group[w] lanes; a wave with wave_group[w] == 0 holds
    // systems of different sizes, wave_sys[wave_first[w] ..
    // +wave_count[w]-1], N lanes each, in that order
    const int *wave_first, *wave_count, *wave_group, *wave_sys;
    int n_waves;
    // historic snapshots (optional): record r of system s is written at
    // hist[(r * HIST_FIELDS + f) * n_bodies + body], hist_count[s] records so far
    double *hist;
    int *hist_count;
    int hist_cap;
};

// Fields of one body's historic-snapshot record on the device (the host adds
// the constant ones: time step, mass, dissipation factor, lag angle and
// migration timescale, src/integrator/output.rs:90-134).
enum HistField {
    HF_TIME = 0, HF_X = 1, HF_SPIN = 4, HF_V = 7, HF_RADIUS = 10, HF_RG2 = 11, HF_LOVE = 12, HF_DEDT = 13,
    HF_LAG = 14, HF_MIGRATION = 15, HIST_FIELDS = 16
};

// evolving-dissipation models (constant_time_lag.rs:26, 100)
__host__ __device__ __forceinline__ bool evo_dynamical(int type) {
    return type == POS_EVO_BOLMONT_MATHIS2016 || type == POS_EVO_GALLET_BOLMONT2017 ||
           type == POS_EVO_LECONTE_CHABRIER2013_DYN;
}

// ---- LDS of one wavefront ---------------------------------------------------
constexpr int NX = 24;   // exchange slots per lane
struct WaveLds {
    // persistent per body (lane)
    double m[WAVE + 1];   // entry WAVE is +0, the mass of the exchange area's zero column
    double inv_m[WAVE], sigma[WAVE], k2[WAVE], k2f[WAVE];
    // of an orbiting body's pair with its host (masses only): the Kidder mass
    // factor f = G m_h G m / (G m_h + G m)^2 (general_relativity.rs:91) and
    // 1 / (m_h + m) (general_relativity.rs:326)
    double kf[WAVE], inv_M[WAVE];
    double R[WAVE], I[WAVE];
    double s[3][WAVE];             // spin of the last evaluation
    double ag[3][WAVE];            // gravitational acceleration (universe.rs:300-302)
    // phase-dependent exchange; column WAVE holds +0 in every row (written
    // once per wavefront, never by a lane): the ordered folds read it for
    // the slots they skip.  The odd row stride also spreads the folds' reads
    // of one column in several rows over the LDS banks.
    double xb[NX][WAVE + 1];
};

// Every kernel that uses the exchange area runs one wavefront per workgroup
// (__launch_bounds__(64)), so the lanes that publish and the lanes that read
// a value are in the same wavefront.  The LDS executes the operations of one
// wavefront in the order they are issued, so a read issued after a write sees
// it: ordering the two at wavefront scope is enough, which costs no wait
// (a workgroup-scope barrier would wait for every outstanding LDS write).
#ifndef POS_WSYNC_WORKGROUP
#define POS_WSYNC_WORKGROUP 0
#endif
__device__ __forceinline__ void wsync() {
#if POS_WSYNC_WORKGROUP
    __syncthreads();
#else
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#endif
}

// The lane group of `lane` in wave `wave`: its device system, first lane and
// this lane's index in it.  A lane past the wave's groups gets gvalid =
// false, the wave's first system and its own lane as group start (so that
// every LDS index it forms stays inside the wave's columns).
struct LaneGroup {
    int sys, gbase, k;
    bool gvalid;
};
__device__ inline LaneGroup lane_group(const DevEnsemble &E, int wave, int lane) {
    const int lanes = E.wave_group[wave];   // lanes per group, 0 = sizes mixed
    const int count = E.wave_count[wave];
    LaneGroup r;
    if (lanes > 0) {
        const int g = lane / lanes;
        r.gvalid = g < count;
        r.sys = E.wave_first[wave] + (r.gvalid ? g : 0);
        r.gbase = g * lanes;
        r.k = lane - g * lanes;
        return r;
    }
    const int *list = E.wave_sys + E.wave_first[wave];
    r.sys = list[0];
    r.gbase = lane;
    r.k = 0;
    r.gvalid = false;
    int base = 0;
    for (int i = 0; i < count; ++i) {
        const int s = list[i];
        const int n = E.body_offset[s + 1] - E.body_offset[s];
        if (lane < base + n) {
            r.sys = s;
            r.gbase = base;
            r.k = lane - base;
            r.gvalid = true;
            break;
        }
        base += n;
    }
    return r;
}

// G m of the body in LDS slot l (the reference's mass_g = mass * K2)
__device__ __forceinline__ double mg_of(const WaveLds &lds, int l) { return lds.m[l] * K2; }

// ---- per-lane registers -----------------------------------------------------
enum : uint32_t {
    F_VALID = 1u << 0, F_HOST = 1u << 1, F_TIDES = 1u << 2, F_RF = 1u << 3, F_GR = 1u << 4,
    F_WIND = 1u << 5, F_EVOLUTION = 1u << 6, F_SPIN = 1u << 7, F_WIND_ROLE = 1u << 8, F_LC_LOVE = 1u << 9,
    F_DISK = 1u << 10,        // the system has the disk effect
    F_DISK_ROLE = 1u << 11,   // this body is a disk orbiting body
    F_DYN_HOST = 1u << 12,    // the host has an evolving-dissipation model
    F_DYN = 1u << 13,         // this body has an evolving-dissipation model
    // an orbiting body of the effect (TidesEffect / RotationalFlatteningEffect /
    // GeneralRelativityEffect::OrbitingBody) in a system where it is on; the
    // effect loops of the reference skip every other body
    // (constant_time_lag.rs:217-218, oblate_spheroid.rs:24, general_relativity.rs:96-97)
    F_TIDES_ROLE = 1u << 14,
    F_RF_ROLE = 1u << 15,
    F_GR_ROLE = 1u << 16
};
struct Lane {
    int lane, gbase, k, n, host, coord, sys, body;
    uint32_t flags, status;
    int tw, evo, left_index, hcount;
    uint32_t seen_status;   // the status and warning count the group last agreed on
    int seen_tw;
    long long iteration;
    double t, dt, last_hist;
    double hist_period, time_limit;   // per-system constants, read once per launch
    // the largest squared Roche radius of this body's pairs with the bodies
    // after it (rr * rr, as the Roche test squares it), read once per launch:
    // a pair farther apart than that cannot meet its own Roche test
    double roche_max2;
    double m, inv_m;
    double x[3], v[3], L[3], ev[3], eL[3];
    // alternative coordinates of this body; com* hold the host slot (centre of mass)
    double ax[3], av[3];
    double comx[3], comv[3], com_m;
    double kx;   // universal anomaly of this body's last Kepler drift

    __device__ __forceinline__ bool valid() const { return flags & F_VALID; }
    __device__ __forceinline__ bool is_host() const { return flags & F_HOST; }
    __device__ __forceinline__ bool orbiting() const { return (flags & (F_VALID | F_HOST)) == F_VALID; }
    __device__ __forceinline__ bool tides() const { return flags & F_TIDES; }
    __device__ __forceinline__ bool rf() const { return flags & F_RF; }
    __device__ __forceinline__ bool gr() const { return flags & F_GR; }
    __device__ __forceinline__ bool tides_role() const { return flags & F_TIDES_ROLE; }
    __device__ __forceinline__ bool rf_role() const { return flags & F_RF_ROLE; }
    __device__ __forceinline__ bool gr_role() const { return flags & F_GR_ROLE; }
    __device__ __forceinline__ bool spin_on() const { return flags & F_SPIN; }
    __device__ __forceinline__ int host_body() const { return body - k + host; }
    __device__ __forceinline__ int hl() const { return gbase + host; }
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

// calculate_particles_non_spin_dependent_evolving_quantities for this body
__device__ inline void evolve_body(const DevEnsemble &E, WaveLds &lds, Lane &B, double t) {
    if (!B.valid() || B.evo == POS_EVO_NON_EVOLVING) return;
    const int off = E.evo_off[B.body], len = E.evo_len[B.body];
    const double *tt = E.tab_time + off;
    int i0 = evolver_idx(B.left_index, len);
    int left = 0;
    const double R = lds.R[B.lane], rg2 = E.rg2[B.body];
    const double newR = interp_table(tt + i0, E.tab_radius + off + i0, len - i0, t, &left);
    B.left_index += left;
    double newrg2 = rg2;
    if (B.evo == POS_EVO_BARAFFE2015 || B.evo == POS_EVO_LECONTE2011 || B.evo == POS_EVO_LECONTE_CHABRIER2013 ||
        B.evo == POS_EVO_LECONTE_CHABRIER2013_DYN) {
        i0 = evolver_idx(B.left_index, len);
        newrg2 = interp_table(tt + i0, E.tab_rg2 + off + i0, len - i0, t, &left);
        B.left_index += left;
    }
    if (newR != R || newrg2 != rg2) {
        lds.R[B.lane] = newR;
        E.rg2[B.body] = newrg2;
        lds.I[B.lane] = B.m * newrg2 * p2(newR);
    }
    if (B.flags & F_LC_LOVE) {   // love number evaluated and discarded (evolution.rs:536-538)
        i0 = evolver_idx(B.left_index, len);
        (void)interp_table(tt + i0, E.tab_love + off + i0, len - i0, t, &left);
        B.left_index += left;
    }
}

// particles/common.rs:3-14 for this lane's body
__device__ __forceinline__ void own_spin(WaveLds &lds, Lane &B) {
    if (!B.valid()) return;
    const double I = lds.I[B.lane];
    if (I == 0.) { B.status |= POS_STATUS_ZERO_INERTIA; return; }
    const Rcp RI = rcp(I);
    const double s0 = divr(B.L[0], RI), s1 = divr(B.L[1], RI), s2 = divr(B.L[2], RI);
    lds.s[0][B.lane] = s0; lds.s[1][B.lane] = s1; lds.s[2][B.lane] = s2;
}

// |spin|^2 of lane l's last evaluation (the same rounded value as when the
// spin is formed: the same products summed in the same order)
__device__ __forceinline__ double spin2(const WaveLds &lds, int l) {
    return p2(lds.s[0][l]) + p2(lds.s[1][l]) + p2(lds.s[2][l]);
}

// evolution.rs:548-567 for this lane's body (after its spin is refreshed):
// the lag angle of an evolving-dissipation model, from the inverse tidal Q
// factor interpolated at t (which moves the evolver's left_index too).
__device__ inline void evolve_spin_dependent(const DevEnsemble &E, WaveLds &lds, Lane &B, double t) {
    if (!B.valid() || !(B.flags & F_DYN)) return;
    const int off = E.evo_off[B.body], len = E.evo_len[B.body];
    const int i0 = evolver_idx(B.left_index, len);
    int left = 0;
    const double itq = interp_table(E.tab_time + off + i0, E.tab_itq + off + i0, len - i0, t, &left);
    B.left_index += left;
    const double epsilon_squared = spin2(lds, B.lane) / SUN_DYN_FREQ_2;
    E.lag[B.body] = 3.0 * epsilon_squared * itq / 4.0;
}

// ===========================================================================
// Additional effects + implicit midpoint (whfast.rs:322-466, universe.rs:353-614)
// ===========================================================================
enum {   // orbiting body -> host contributions
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

__device__ __forceinline__ void put3(WaveLds &lds, int slot, int lane, double x, double y, double z) {
    lds.xb[slot][lane] = x; lds.xb[slot + 1][lane] = y; lds.xb[slot + 2][lane] = z;
}

// R reductions over the bodies of a lane's group, each in body order.  The
// terms of FOLD_BATCH bodies of every row are formed first (their LDS loads
// issue together and one wait covers them), then folded row by row in body
// order.  Slots past the group's last body, and the host's slot when
// SKIP_HOST, are given index WAVE: the exchange area's zero column and the
// zero mass past the last lane, so the callback's term there is +0 (or -0
// where it negates its terms).  Adding that term leaves the sum unchanged bit
// for bit, so each result is the sequential sum over the bodies it names:
// adding -0 changes no value, and adding +0 changes only -0, which a sum
// that starts at +0 never is (in round-to-nearest x + y is -0 only when both
// x and y are).  Every caller starts at +0 or adds -0 terms.
// FOLD_BATCH and the rows per pass of the reaction sums (REACTION_ROWS) are
// the fastest of the settings measured on the benchmark ensemble: the loads in
// flight cost registers.
constexpr int FOLD_BATCH = 4;
constexpr int REACTION_ROWS = 3;
template <int R, bool SKIP_HOST, class Load>
__device__ __forceinline__ void fold_rows(const Lane &B, double (&acc)[R], Load load) {
    for (int j0 = 0; j0 < B.n; j0 += FOLD_BATCH) {
        double t[R][FOLD_BATCH];
#pragma unroll
        for (int u = 0; u < FOLD_BATCH; ++u) {
            const int j = j0 + u;
            const bool use = j < B.n && !(SKIP_HOST && j == B.host);
            const int l = use ? B.gbase + j : WAVE;
#pragma unroll
            for (int r = 0; r < R; ++r) t[r][u] = load(r, l);
        }
#pragma unroll
        for (int u = 0; u < FOLD_BATCH; ++u)
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = acc[r] + t[r][u];
    }
}

// Ordered sums over the orbiting bodies of the group of the contribution rows
// [0, C_COUNT), spread over the group's lanes: lane k sums rows k, k+n, ...,
// REACTION_ROWS of them per pass.  Every sum runs over the bodies in index
// order, so it is exactly the sequential reference sum.  The host's column of
// the contribution rows is never an input (the host contributes nothing), so
// each result is written there; after the next wsync() the host reads its
// reactions in its own column.
__device__ __forceinline__ void host_reaction_sums(WaveLds &lds, const Lane &B) {
    const int hl = B.hl();
    for (int q0 = B.k; B.valid() && q0 < C_COUNT; q0 += REACTION_ROWS * B.n) {
        int q[REACTION_ROWS];
        double acc[REACTION_ROWS];
#pragma unroll
        for (int r = 0; r < REACTION_ROWS; ++r) {
            q[r] = q0 + r * B.n < C_COUNT ? q0 + r * B.n : q0;   // a row past the last reads row q0, unused
            acc[r] = 0.;
        }
        fold_rows<REACTION_ROWS, true>(B, acc, [&](int r, int l) { return lds.xb[q[r]][l]; });
#pragma unroll
        for (int r = 0; r < REACTION_ROWS; ++r)
            if (q0 + r * B.n < C_COUNT) lds.xb[q[r]][hl] = acc[r];
    }
}

// ===========================================================================
// General-relativity variants other than Kidder 1995.  They run after the
// host reactions of the other effects (their accelerations are added last,
// universe.rs:540-566) and use the exchange rows freely.
// ===========================================================================
enum GrVariant { GRV_KIDDER = 0, GRV_ANDERSON = 1, GRV_NEWHALL = 2 };
// Kernel variant: the GR variant plus VAR_EXT for systems with the disk effect
// or an evolving-dissipation host (the extended effects pass).  The device
// copy of a system's effect mask carries its variant in bits 16-18.
constexpr int VAR_EXT = 4;
// VAR_FAST: every system of the launch has exactly tides + flattening + Kidder
// GR (spins integrated), no wind, disk or evolution, democratic heliocentric
// coordinates (with VAR_JACOBI: Jacobi coordinates); the flags and the
// coordinate type are then compile-time constants and the per-system branches
// fold away (same arithmetic).
constexpr int VAR_FAST = 8;
constexpr int VAR_JACOBI = 16;   // only with VAR_FAST
constexpr int N_VARIANTS = 32;
constexpr uint32_t EFF_VAR_SHIFT = 16;
__host__ __device__ __forceinline__ int eff_variant(uint32_t eff) { return (int)((eff >> EFF_VAR_SHIFT) & 31u); }
__host__ __device__ __forceinline__ int eff_gr_variant(uint32_t eff) { return eff_variant(eff) & 3; }

// The effect and role tests of a lane for kernel variant V.  Under VAR_FAST
// every effect is on and every orbiting body has every orbiting role, so the
// tests are compile-time constants there (the same values the flags hold) and
// the pair terms of an orbiting lane form one block the compiler can schedule
// as a whole.
template <int V>
struct LaneTests {
    bool tides, rf, gr, spin, tides_role, rf_role, gr_role;
    __device__ __forceinline__ explicit LaneTests(const Lane &B) {
        constexpr bool F = (V & VAR_FAST) != 0;
        const bool orb = B.orbiting();
        tides = F || B.tides();
        rf = F || B.rf();
        gr = F || B.gr();
        spin = F || B.spin_on();
        tides_role = F ? orb : B.tides_role();
        rf_role = F ? orb : B.rf_role();
        gr_role = F ? orb : B.gr_role();
    }
};

// Chained order of general_relativity.rs (host, then the other bodies by
// index): position of body k, and body at position c.
__device__ __forceinline__ int chained_pos(const Lane &B, int k) { return k == B.host ? 0 : (k < B.host ? k + 1 : k); }
__device__ __forceinline__ int chained_body(const Lane &B, int c) { return c == 0 ? B.host : (c - 1 < B.host ? c - 1 : c); }

// general_relativity.rs:641-681: the accelerations of the last gravity
// evaluation plus the host pair terms the splitting ignored; the separation
// mixes the host's inertial position with the body's heliocentric position,
// as the reference does.  Rows 0-2 must hold every body's position.  Leaves
// this body's acceleration in an[]; uses rows 3-5.
__device__ inline void gr_newtonian(WaveLds &lds, Lane &B, double an[3]) {
    const int lane = B.lane, hl = B.hl();
#pragma unroll
    for (int q = 0; q < 3; ++q) an[q] = lds.ag[q][lane];
    // the ignored pairs: the first body after the host in chained order
    // (Jacobi) or every body (heliocentric splittings)
    const int np = B.coord == POS_COORD_JACOBI ? 1 : B.n - 1;
    const int c = chained_pos(B, B.k);
    double host_term[3] = {0., 0., 0.};
    if (B.orbiting() && c <= np) {   // every body of a GR system is a GR body
        const double hx[3] = {B.x[0] - lds.xb[0][hl], B.x[1] - lds.xb[1][hl], B.x[2] - lds.xb[2][hl]};
        const double dx = lds.xb[0][hl] - hx[0], dy = lds.xb[1][hl] - hx[1], dz = lds.xb[2][hl] - hx[2];
        const double r2 = p2(dx) + p2(dy) + p2(dz);
        const double r = sqrt(r2);
        const double prefac = G / (r2 * r);
        const double pms = prefac * lds.m[hl], pmp = prefac * B.m;
        an[0] = an[0] + pms * dx; an[1] = an[1] + pms * dy; an[2] = an[2] + pms * dz;
        host_term[0] = pmp * dx; host_term[1] = pmp * dy; host_term[2] = pmp * dz;
    }
    if (B.valid()) { lds.xb[3][lane] = host_term[0]; lds.xb[4][lane] = host_term[1]; lds.xb[5][lane] = host_term[2]; }
    wsync();
    if (B.valid() && B.is_host()) {
        for (int cc = 1; cc <= np && cc < B.n; ++cc) {
            const int l = B.gbase + chained_body(B, cc);
#pragma unroll
            for (int q = 0; q < 3; ++q) an[q] = an[q] - lds.xb[3 + q][l];
        }
    }
    wsync();
}

// Anderson et al. 1975 through Jacobi coordinates (general_relativity.rs:461-639)
__device__ inline void gr_anderson1975(WaveLds &lds, Lane &B, double g[3]) {
    const int lane = B.lane, hl = B.hl();
    if (B.valid()) { lds.xb[0][lane] = B.x[0]; lds.xb[1][lane] = B.x[1]; lds.xb[2][lane] = B.x[2]; }
    wsync();
    double an[3];
    gr_newtonian(lds, B, an);
    if (B.valid()) {
#pragma unroll
        for (int q = 0; q < 3; ++q) { lds.xb[3 + q][lane] = B.v[q]; lds.xb[6 + q][lane] = an[q]; }
    }
    wsync();
    // Jacobi prefix recurrence up to this body, recomputed by every lane
    double eta = lds.m[hl];
    double s[3], sv[3], sa[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) { s[q] = eta * lds.xb[q][hl]; sv[q] = eta * lds.xb[3 + q][hl]; sa[q] = eta * lds.xb[6 + q][hl]; }
    double jx[3] = {0., 0., 0.}, jv[3] = {0., 0., 0.}, ja[3] = {0., 0., 0.};
    for (int j = 0; j < B.n; ++j) {
        if (j == B.host) continue;
        const int l = B.gbase + j;
        const double mj = lds.m[l];
        const double ei = 1. / eta;
        eta = eta + mj;
        const double pme = eta * ei;
        double x[3], v[3], a[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            x[q] = lds.xb[q][l] - s[q] * ei;
            v[q] = lds.xb[3 + q][l] - sv[q] * ei;
            a[q] = lds.xb[6 + q][l] - sa[q] * ei;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            s[q] = s[q] * pme + mj * x[q];
            sv[q] = sv[q] * pme + mj * v[q];
            sa[q] = sa[q] * pme + mj * a[q];
        }
        if (j == B.k) {
#pragma unroll
            for (int q = 0; q < 3; ++q) { jx[q] = x[q]; jv[q] = v[q]; ja[q] = a[q]; }
        }
    }
    const double mtot = eta;
    wsync();
    if (B.orbiting()) {
        const double mu = mg_of(lds, hl);
        double vi[3] = {jv[0], jv[1], jv[2]};
        double vi2 = p2(jv[0]) + p2(jv[1]) + p2(jv[2]);
        const double ri = sqrt(p2(jx[0]) + p2(jx[1]) + p2(jx[2]));
        double fa = (0.5 * vi2 + 3. * mu / ri) / SPEED_OF_LIGHT_2;
        for (int it = 0; it < 10; ++it) {
            const double o0 = vi[0], o1 = vi[1], o2 = vi[2];
            vi[0] = jv[0] / (1. - fa); vi[1] = jv[1] / (1. - fa); vi[2] = jv[2] / (1. - fa);
            vi2 = vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2];
            fa = (0.5 * vi2 + 3. * mu / ri) / SPEED_OF_LIGHT_2;
            const double dvx = vi[0] - o0, dvy = vi[1] - o1, dvz = vi[2] - o2;
            if ((dvx * dvx + dvy * dvy + dvz * dvz) / vi2 < DBL_EPS_2) break;
        }
        const double ri3 = ri * ri * ri;
        const double fb = (mu / ri - 1.5 * vi2) * mu / ri3 / SPEED_OF_LIGHT_2;
        const double rdr = jx[0] * jv[0] + jx[1] * jv[1] + jx[2] * jv[2];
        const double vd0 = ja[0] + fb * jx[0], vd1 = ja[1] + fb * jx[1], vd2 = ja[2] + fb * jx[2];
        const double vdv = vi[0] * vd0 + vi[1] * vd1 + vi[2] * vd2;
        const double fd = (vdv - 3. * mu / ri3 * rdr) / SPEED_OF_LIGHT_2;
#pragma unroll
        for (int q = 0; q < 3; ++q) ja[q] = fb * (1. - fa) * jx[q] - fa * ja[q] - fd * vi[q];
        lds.xb[9][lane] = ja[0]; lds.xb[10][lane] = ja[1]; lds.xb[11][lane] = ja[2];
    }
    wsync();
    // reverse recurrence back to inertial accelerations (Jacobi star acceleration zero)
    eta = mtot;
    double t[3] = {eta * 0., eta * 0., eta * 0.};
    g[0] = g[1] = g[2] = 0.;
    for (int j = B.n - 1; j >= 0; --j) {
        if (j == B.host) continue;
        const int l = B.gbase + j;
        const double mj = lds.m[l];
        const double ei = 1. / eta;
        double a[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            t[q] = (t[q] - mj * lds.xb[9 + q][l]) * ei;
            a[q] = lds.xb[9 + q][l] + t[q];
        }
        if (j == B.k) { g[0] = a[0]; g[1] = a[1]; g[2] = a[2]; }
        eta = eta - mj;
#pragma unroll
        for (int q = 0; q < 3; ++q) t[q] = t[q] * eta;
    }
    if (B.valid() && B.is_host()) {
        const double mi = 1. / eta;
        g[0] = t[0] * mi; g[1] = t[1] * mi; g[2] = t[2] * mi;
    }
    wsync();
}

// Newhall et al. 1983, every pair (general_relativity.rs:683-895)
__device__ inline void gr_newhall1983(WaveLds &lds, Lane &B, double g[3]) {
    const int lane = B.lane;
    const int n = B.n;
    if (B.valid()) { lds.xb[0][lane] = B.x[0]; lds.xb[1][lane] = B.x[1]; lds.xb[2][lane] = B.x[2]; }
    wsync();
    double an[3];
    gr_newtonian(lds, B, an);
    if (B.valid()) {
#pragma unroll
        for (int q = 0; q < 3; ++q) { lds.xb[3 + q][lane] = B.v[q]; lds.xb[6 + q][lane] = an[q]; }
    }
    const double c2 = SPEED_OF_LIGHT_2;
    // distance to the body at chained position c (0 for itself)
    auto dist = [&](int c) {
        const int l = B.gbase + chained_body(B, c);
        const double dx = B.x[0] - lds.xb[0][l], dy = B.x[1] - lds.xb[1][l], dz = B.x[2] - lds.xb[2][l];
        return sqrt(p2(dx) + p2(dy) + p2(dz));
    };
    const int me = chained_pos(B, B.k);
    wsync();
    // a2 depends only on the second body j: sum_l G m_l / (c^2 r_lj), which
    // body j evaluates from its own distances (|x_l - x_j| = |x_j - x_l|)
    double a1 = 0., a2own = 0.;
    if (B.valid()) {
        for (int c = 0; c < n; ++c) {
            if (c == me) continue;
            const double mk = lds.m[B.gbase + chained_body(B, c)];
            const double r = dist(c);
            a1 = a1 + (4. / c2) * G * mk / r;
            a2own = a2own + (1. / c2) * G * mk / r;
        }
        lds.xb[9][lane] = a2own;
    }
    wsync();
    double ac[3] = {0., 0., 0.};
    if (B.valid()) {
        const double vi2 = p2(B.v[0]) + p2(B.v[1]) + p2(B.v[2]);
        for (int c = 0; c < n; ++c) {
            if (c == me) continue;
            const int l = B.gbase + chained_body(B, c);
            const double dx = B.x[0] - lds.xb[0][l], dy = B.x[1] - lds.xb[1][l], dz = B.x[2] - lds.xb[2][l];
            const double rij = sqrt(p2(dx) + p2(dy) + p2(dz));
            const double rij2 = p2(rij);
            const double rij3 = rij2 * rij;
            const double vj0 = lds.xb[3][l], vj1 = lds.xb[4][l], vj2c = lds.xb[5][l];
            const double a2 = lds.xb[9][l];
            const double a3 = -vi2 / c2;
            const double vj2 = p2(vj0) + p2(vj1) + p2(vj2c);
            const double a4 = -2. * vj2 / c2;
            const double a5 = (4. / c2) * (B.v[0] * vj0 + B.v[1] * vj1 + B.v[2] * vj2c);
            const double a60 = dx * vj0 + dy * vj1 + dz * vj2c;
            const double a6 = (3. / (2. * c2)) * p2(a60) / rij2;
            const double f1 = a1 + a2 + a3 + a4 + a5 + a6;
            const double gm = G * lds.m[l];
            ac[0] = ac[0] + gm * dx * f1 / rij3;
            ac[1] = ac[1] + gm * dy * f1 / rij3;
            ac[2] = ac[2] + gm * dz * f1 / rij3;
            const double dvx = B.v[0] - vj0, dvy = B.v[1] - vj1, dvz = B.v[2] - vj2c;
            const double f2 = dx * (4. * B.v[0] - 3. * vj0) + dy * (4. * B.v[1] - 3. * vj1) + dz * (4. * B.v[2] - 3. * vj2c);
            ac[0] = ac[0] + gm * f2 * dvx / rij3 / c2;
            ac[1] = ac[1] + gm * f2 * dvy / rij3 / c2;
            ac[2] = ac[2] + gm * f2 * dvz / rij3 / c2;
        }
    }
    // fixed point: rows 10-12 hold the previous iterate of every body, rows
    // 13-15 the new one.  A group stops iterating when its own test passes
    // (every lane of the group evaluates it identically); the wavefront
    // leaves the loop once every group has.
    double anew[3] = {0., 0., 0.};
    bool done = !B.valid();
    for (int k = 0; k < 10; ++k) {
        if (B.valid()) { lds.xb[10][lane] = anew[0]; lds.xb[11][lane] = anew[1]; lds.xb[12][lane] = anew[2]; }
        wsync();
        double nc[3] = {0., 0., 0.};
        if (!done) {
            for (int c = 0; c < n; ++c) {
                if (c == me) continue;
                const int l = B.gbase + chained_body(B, c);
                const double dx = B.x[0] - lds.xb[0][l], dy = B.x[1] - lds.xb[1][l], dz = B.x[2] - lds.xb[2][l];
                const double rij = sqrt(p2(dx) + p2(dy) + p2(dz));
                const double rij2 = p2(rij);
                const double rij3 = rij2 * rij;
                const double s0 = lds.xb[6][l] + lds.xb[10][l], s1 = lds.xb[7][l] + lds.xb[11][l], s2 = lds.xb[8][l] + lds.xb[12][l];
                const double proj = dx * s0 + dy * s1 + dz * s2;
                const double gm = G * lds.m[l];
                nc[0] = nc[0] + (gm * dx / rij3) * proj / (2. * c2) + (7. / (2. * c2)) * G * lds.m[l] * s0 / rij;
                nc[1] = nc[1] + (gm * dy / rij3) * proj / (2. * c2) + (7. / (2. * c2)) * G * lds.m[l] * s1 / rij;
                nc[2] = nc[2] + (gm * dz / rij3) * proj / (2. * c2) + (7. / (2. * c2)) * G * lds.m[l] * s2 / rij;
            }
        }
        if (!done) { anew[0] = ac[0] + nc[0]; anew[1] = ac[1] + nc[1]; anew[2] = ac[2] + nc[2]; }
        if (B.valid()) { lds.xb[13][lane] = anew[0]; lds.xb[14][lane] = anew[1]; lds.xb[15][lane] = anew[2]; }
        wsync();
        // convergence test of the reference, in its order (the deviations
        // persist across bodies and are only updated for tiny components)
        const double dev_limit = 1.0e-30;
        double maxdev = 0., ddx = 0., ddy = 0., ddz = 0.;
        for (int c = 0; !done && c < n; ++c) {
            const int l = B.gbase + chained_body(B, c);
            const double nx = lds.xb[13][l], ny = lds.xb[14][l], nz = lds.xb[15][l];
            if (fabs(nx) < dev_limit) ddx = fabs(nx - lds.xb[10][l]) / nx;
            if (fabs(ny) < dev_limit) ddy = fabs(ny - lds.xb[11][l]) / ny;
            if (fabs(nz) < dev_limit) ddz = fabs(nz - lds.xb[12][l]) / nz;
            if (ddx > maxdev) maxdev = ddx;
            if (ddy > maxdev) maxdev = ddy;
            if (ddz > maxdev) maxdev = ddz;
        }
        if (maxdev < dev_limit) done = true;
        wsync();
        if (__all(done)) break;
    }
    g[0] = anew[0]; g[1] = anew[1]; g[2] = anew[2];
}

// tools.rs:251-285 (Mercury's perihelion distance and eccentricity)
__device__ inline void perihelion_distance_and_eccentricity(double gm, const double x[3], const double u[3], double &q,
                                                            double &e) {
    const double hx = x[1] * u[2] - x[2] * u[1];
    const double hy = x[2] * u[0] - x[0] * u[2];
    const double hz = x[0] * u[1] - x[1] * u[0];
    const double h2 = hx * hx + hy * hy + hz * hz;
    const double v2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
    const double sv = h2 / gm;
    const double temp = 1. + sv * (v2 / gm - 2. / r);
    e = temp <= 0. ? 0. : sqrt(temp);
    q = sv / (1. + e);
}

// constant_time_lag.rs:57-92 / 121-156: the pair value of an evolving-
// dissipation body (norm_spin, radius, lag angle, sigma's two factors) when
// the orbit's mean motion excites its dynamical tide.
__device__ inline bool excited_pair_factor(double mean_motion, double norm_spin, double radius, double lag,
                                           double scale, double factor, double &value) {
    double half = fabs(norm_spin - mean_motion);
    if (!(half < norm_spin)) return false;
    if (half < SMOOTHING_FACTOR_DYN_TIDE_COROTATION) half = SMOOTHING_FACTOR_DYN_TIDE_COROTATION;
    const double inverse_of_half = 1. / half;
    value = scale * (2.0 * K2 / (3.0 * p5(radius)) * lag * inverse_of_half + factor);
    return true;
}

// Disk force on an orbiting body (disk.rs:155-268, Alibert et al. 2013) from
// its heliocentric position and velocity; stores its migration timescale.
__device__ inline void disk_force(const DevEnsemble &E, const WaveLds &lds, const Lane &B, const double px[3],
                                  const double pv[3], double d, double nv2, double F[3]) {
    const double *dp = E.disk_props + 6 * B.sys;   // inner, outer, lifetime, alpha, sigma0, mu
    const int hl = B.hl();
    const double hm = lds.m[hl], hmg = mg_of(lds, hl), m = B.m, mg = mg_of(lds, B.lane);
    const double sma = (mg + hmg) * d / (2.0 * (mg + hmg) - d * nv2);
    // gas radial velocity from the alpha viscosity (disk.rs:206-226)
    const double keplerian_frequency = sqrt(hmg / (d * (d * d)));
    const double temperature = 280.0 * poscr::cr_rsqrt(d) * hm;
    const double speed_of_sound_2 = BOLTZMANN_CONSTANT * temperature / (dp[5] * MASS_HYDROGEN_ATOM);
    const double viscosity = dp[3] * speed_of_sound_2 / keplerian_frequency;
    const double gas_radial_velocity = -3.0 * viscosity / (2.0 * sma);
    // surface density (disk.rs:196-205)
    const double initial = dp[4] * (1. / d) * poscr::cr_exp(-1.0 * d / dp[1]) * (1.0 - sqrt(dp[0] / d));
    double surface_density = 0.0;
    if (d > dp[0]) surface_density = initial * poscr::cr_exp(-B.t / dp[2]);
    const double ratio = 2.0 * surface_density * p2(sma) / m;
    const double a_dot = gas_radial_velocity * ((ratio != ratio || ratio > 1.0) ? 1.0 : ratio);   // f64::min(1, .)
    const double timescale = -sma / a_dot;
    E.disk_migration[B.body] = timescale;
    const double factor_migration = -m / timescale;
    const double eccentricity_timescale = 0.1 * timescale;
    const double factor_eccentricity = -m * 2.0 / eccentricity_timescale;
    const double factor_inclination = -m * 2.0 / eccentricity_timescale;
    const double sp = 1.0 / (d * d) * (px[0] * pv[0] + px[1] * pv[1] + px[2] * pv[2]);
    F[0] = factor_migration * pv[0] + factor_eccentricity * sp * px[0] + 0.0;
    F[1] = factor_migration * pv[1] + factor_eccentricity * sp * px[1] + 0.0;
    F[2] = factor_migration * pv[2] + factor_eccentricity * sp * px[2] + factor_inclination * pv[2];
}

// Terms of an orbiting body's pair with its host that depend only on the
// positions, masses, radii and fixed dissipation parameters.  Within one
// implicit-midpoint solve the positions do not move, and in the fixed-effect-
// set kernel (no evolution, no evolving dissipation) the radii and factors do
// not change either, so these are evaluated once per solve instead of once
// per iteration.  Each is the same rounded value the in-loop expression gives
// (same operations in the same order), and a quotient by a hoisted divisor
// uses its correctly rounded reciprocal (divr), which is exact.
struct PairGeo {
    double px[3];            // heliocentric position
    Rcp Rd;                  // distance and its reciprocal
    Rcp Rp8;                 // distance^8 and its reciprocal
    double d5;               // distance^5
    double os, op;           // tidal orthogonal components (constant_time_lag.rs:206-264)
    double od, pd, id;       // os / d, op / d, 1 / d
    double conservative;     // conservative radial tidal component (266-307)
    double term1, term2;     // the two dissipative radial terms without their velocity factor
    double m3d5, c15;        // -3 / d^5 and 15 / d^7 (oblate_spheroid.rs:51-60)
    double mgc, mg_d;        // -(G m_h + G m) / (d^2 c^2) and (G m_h + G m) / d
    double c1pn, o1pn;       // 1PN: 2 (2 + f) mg / d and (mg / (d^2 c^2)) 2 (2 - f)
    double c2pn;             // 2PN: 3/4 (12 + 29 f) (mg^2 / d^2)
    double n3[3];            // unit vector of the separation
    // carried from one evaluation to the next: the scalar products of the
    // separation with the host's and the body's spins of the previous
    // evaluation (tides/common.rs:143-170, rotational_flattening/common.rs:93-120)
    double rs_star, rs_planet;
};

// The pair geometry of this lane's body (orbiting lanes; zeros elsewhere),
// with every lane's position exchanged through rows 0-2, and the scalar
// products with the spins currently in LDS (those of the last evaluation).
__device__ inline void pair_geometry(WaveLds &lds, const Lane &B, PairGeo &g) {
    const int lane = B.lane;
    const int hl = B.hl();
    if (B.valid()) { lds.xb[0][lane] = B.x[0]; lds.xb[1][lane] = B.x[1]; lds.xb[2][lane] = B.x[2]; }
    wsync();
    double d = 1.;
#pragma unroll
    for (int q = 0; q < 3; ++q) g.px[q] = 0.;
    g.rs_star = 0.;
    g.rs_planet = 0.;
    if (B.orbiting()) {
#pragma unroll
        for (int q = 0; q < 3; ++q) g.px[q] = B.x[q] - lds.xb[q][hl];
        d = sqrt(p2(g.px[0]) + p2(g.px[1]) + p2(g.px[2]));
        g.rs_star = g.px[0] * lds.s[0][hl] + g.px[1] * lds.s[1][hl] + g.px[2] * lds.s[2][hl];
        g.rs_planet = g.px[0] * lds.s[0][lane] + g.px[1] * lds.s[1][lane] + g.px[2] * lds.s[2][lane];
    }
    wsync();
    const double m = B.m, hm = lds.m[hl];
    const double hR = lds.R[hl], R = lds.R[lane];
    g.Rd = rcp(d);
    const double d7 = p7(d);
    const Rcp Rd7 = rcp(d7);
    g.d5 = p5(d);
    g.Rp8 = rcp(p8(d));
    const double sigma_host = lds.sigma[hl], sigma_pair = lds.sigma[lane];
    g.os = divr(4.5 * p2(m) * p10(hR) * sigma_host, Rd7);
    g.op = divr(4.5 * p2(hm) * p10(R) * sigma_pair, Rd7);
    const double hm2 = hm * hm;
    const double pm2 = m * m;
    g.conservative = divr(-3.0 * K2, Rd7) * (pm2 * p5(hR) * lds.k2[hl] + hm2 * p5(R) * lds.k2[lane]);
    g.term1 = pm2 * p10(hR) * sigma_host;
    g.term2 = hm2 * p10(R) * sigma_pair;
    g.od = divr(g.os, g.Rd);
    g.pd = divr(g.op, g.Rd);
    g.id = divr(1.0, g.Rd);
    g.m3d5 = -3. / g.d5;
    g.c15 = divr(15., Rd7);
    const double hmg = mg_of(lds, hl), pmg = mg_of(lds, lane);
    const double f = lds.kf[lane];
    const double mg = hmg + pmg;
    const double distance_2 = p2(d);
    const Rcp Rdc = rcp(distance_2 * SPEED_OF_LIGHT_2);
    g.mgc = divr(-mg, Rdc);
    g.mg_d = divr(mg, g.Rd);
    g.c1pn = divr(2.0 * (2.0 + f) * mg, g.Rd);
    g.o1pn = divr(mg, Rdc) * 2.0 * (2.0 - f);
    g.c2pn = 3.0 / 4.0 * (12.0 + 29.0 * f) * (p2(mg) / distance_2);
#pragma unroll
    for (int q = 0; q < 3; ++q) g.n3[q] = divr(g.px[q], g.Rd);
}

// Accelerations and spin derivatives of every body of the group at the current
// (averaged) state.  Outputs are accumulated in the reference's order:
// a_add = ((0 + tides) + flattening) + GR and
// dL/dt = ((tides + flattening) + GR) + wind (universe.rs:540-614).
// With H (the fixed-effect-set kernel inside an implicit-midpoint solve) the
// pair terms come from geo, which pair_geometry filled at the start of the
// solve, and the spins' scalar products are carried in geo from one
// evaluation to the next; without H geo is not read.
template <int V, bool H = false>
__device__ inline void additional_effects(const DevEnsemble &E, WaveLds &lds, Lane &B, bool evolution,
                                          double a_add[3], double dLdt[3], PairGeo &geo) {
    constexpr int GRV = V & 3;
    constexpr bool EXT = (V & VAR_EXT) != 0;
    static_assert(!H || ((V & VAR_FAST) != 0 && !EXT), "hoisted pair terms need fixed radii and factors");
    const LaneTests<V> LT(B);
    const int lane = B.lane;
    const int hl = B.hl();
    const bool orbiting = B.orbiting();
    // ---- inertial -> heliocentric (universe.rs:318-351); the host's
    //      heliocentric velocity is zero at this point ----
    if (B.valid()) {
        if (!H) { lds.xb[0][lane] = B.x[0]; lds.xb[1][lane] = B.x[1]; lds.xb[2][lane] = B.x[2]; }
        lds.xb[3][lane] = B.v[0]; lds.xb[4][lane] = B.v[1]; lds.xb[5][lane] = B.v[2];
    }
    // With H the old spins are not read from LDS (their scalar products are
    // carried), so the new spins are published in the same exchange as the
    // velocities.
    if (H) own_spin(lds, B);
    wsync();
    double px[3] = {0., 0., 0.}, pv[3] = {0., 0., 0.};
    double d = 1., nv2 = 0., nv = 1.;
    double rs_star = 0., rs_planet = 0.;
    if (H) {
#pragma unroll
        for (int q = 0; q < 3; ++q) px[q] = geo.px[q];
        d = geo.Rd.b;
    }
    if (orbiting) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (!H) px[q] = B.x[q] - lds.xb[q][hl];
            pv[q] = B.v[q] - lds.xb[3 + q][hl];
        }
        if (!H) d = sqrt(p2(px[0]) + p2(px[1]) + p2(px[2]));
        nv2 = p2(pv[0]) + p2(pv[1]) + p2(pv[2]);
        nv = sqrt(nv2);
        // Universe::initialize runs before the spins are refreshed
        // (universe.rs:429-430), so the scalar products of tides/common.rs:143-170
        // and rotational_flattening/common.rs:93-120 use the spins of the
        // previous evaluation.
        if (H) {
            rs_star = geo.rs_star;
            rs_planet = geo.rs_planet;
        } else {
            rs_star = px[0] * lds.s[0][hl] + px[1] * lds.s[1][hl] + px[2] * lds.s[2][hl];
            rs_planet = px[0] * lds.s[0][lane] + px[1] * lds.s[1][lane] + px[2] * lds.s[2][lane];
        }
    }
    const Rcp Rd = H ? geo.Rd : rcp(d);
    const double vr = orbiting ? divr(px[0] * pv[0] + px[1] * pv[1] + px[2] * pv[2], Rd) : 0.;
    if (!H) {
        wsync();
        // ---- evolution and spin = L / I (universe.rs:305-316) ----
        if (evolution && (B.flags & F_EVOLUTION)) evolve_body(E, lds, B, B.t);
        own_spin(lds, B);
    }
    double dk[3] = {0., 0., 0.};   // disk acceleration of this body
    double host_lag = 0.;
    if (EXT) {
        if (evolution && (B.flags & F_EVOLUTION)) evolve_spin_dependent(E, lds, B, B.t);
        // disk forces of the orbiting bodies (rows 0-2, row 3 = 1 for a disk
        // body) and the host's lag angle (row 3 of its column) for the host
        double F[3] = {0., 0., 0.};
        if (orbiting && (B.flags & F_DISK) && (B.flags & F_DISK_ROLE)) {
            disk_force(E, lds, B, px, pv, d, nv2, F);
#pragma unroll
            for (int q = 0; q < 3; ++q) dk[q] = B.inv_m * F[q];
        }
        if (B.valid()) {
            lds.xb[0][lane] = F[0]; lds.xb[1][lane] = F[1]; lds.xb[2][lane] = F[2];
            lds.xb[3][lane] = B.is_host() ? E.lag[B.body] : ((B.flags & F_DISK_ROLE) ? 1. : 0.);
        }
    }
    if (!H) wsync();
    if (EXT) {
        if (B.valid() && B.is_host() && (B.flags & F_DISK)) {   // disk.rs:259-265, bodies in index order
            double sum[3] = {0., 0., 0.};
            for (int j = 0; j < B.n; ++j) {
                const int l = B.gbase + j;
                if (j == B.host || lds.xb[3][l] == 0.) continue;
#pragma unroll
                for (int q = 0; q < 3; ++q) sum[q] = sum[q] + lds.xb[q][l];
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) dk[q] = -1.0 * B.inv_m * sum[q];
        }
        host_lag = lds.xb[3][hl];
        wsync();
    }

#pragma unroll
    for (int q = 0; q < 3; ++q) { a_add[q] = 0.; dLdt[q] = 0.; }

    if (orbiting) {
        const double m = B.m, hm = lds.m[hl];
        const double hs[3] = {lds.s[0][hl], lds.s[1][hl], lds.s[2][hl]};
        const double s[3] = {lds.s[0][lane], lds.s[1][lane], lds.s[2][lane]};
        const double hs2 = p2(hs[0]) + p2(hs[1]) + p2(hs[2]), s2 = p2(s[0]) + p2(s[1]) + p2(s[2]);
        const double hR = lds.R[hl], R = lds.R[lane];
        const double d7 = H ? 0. : p7(d);
        const Rcp Rd7 = H ? Rcp{0., 0.} : rcp(d7);
        const double d5 = H ? geo.d5 : p5(d);
        // ---- the dissipation factors of this host-body pair: the bodies' own
        //      scaled factors, or for evolving-dissipation models the pair
        //      values of a dynamical tide this orbit excites
        //      (constant_time_lag.rs:20-194).  The host's value is recomputed
        //      here every time; the body's own map entry persists in
        //      pair_ph, and a body that is not excited removes the host's
        //      entry, as the reference does (157-160). ----
        double sigma_host = lds.sigma[hl], sigma_pair = lds.sigma[lane];
        if (EXT && LT.tides_role && (B.flags & (F_DYN_HOST | F_DYN))) {
            const double gm = mg_of(lds, hl) + mg_of(lds, lane);
            double q, e;
            perihelion_distance_and_eccentricity(gm, px, pv, q, e);
            const double mean_motion = sqrt(gm) * poscr::cr_rsqrt3(q / (1.0 - e));
            bool host_entry = false;
            double host_value = 0.;
            if (B.flags & F_DYN_HOST) {
                const int hb = B.host_body();
                host_entry = excited_pair_factor(mean_motion, sqrt(hs2), hR, host_lag, E.sigma_scale[hb],
                                                 E.sigma_factor[hb], host_value);
            }
            if (B.flags & F_DYN) {
                double v;
                if (excited_pair_factor(mean_motion, sqrt(s2), R, E.lag[B.body], E.sigma_scale[B.body],
                                        E.sigma_factor[B.body], v))
                    E.pair_ph[B.body] = v;
                else
                    host_entry = false;
                const double own = E.pair_ph[B.body];
                if (own == own) sigma_pair = own;   // an entry exists (NaN marks none)
            }
            if (host_entry) sigma_host = host_value;
        }
        // ---- tides: orthogonal components (constant_time_lag.rs:206-264) ----
        double os = 0., op = 0.;
        if (H) {
            os = geo.os;
            op = geo.op;
        } else if (LT.tides_role) {
            os = divr(4.5 * p2(m) * p10(hR) * sigma_host, Rd7);
            op = divr(4.5 * p2(hm) * p10(R) * sigma_pair, Rd7);
        }
        // ---- flattening: orthogonal components (oblate_spheroid.rs:21-49) ----
        double fs = 0., fp = 0., ofs = 0., ofp = 0.;
        if (LT.rf_role) {
            constexpr Rcp R6{6., 1. / 6.};
            fs = divr(m * lds.k2f[hl] * hs2 * p5(hR), R6);
            ofs = -6. * fs * rs_star / (hs2 * d5);
            fp = divr(hm * lds.k2f[lane] * s2 * p5(R), R6);
            ofp = -6. * fp * rs_planet / (s2 * d5);
        }
        // A body the tidal loops skip adds nothing to the host's reaction sums:
        // its rows hold +0, which leaves an ordered sum that starts at +0
        // unchanged (such a sum is never -0).
        if (LT.tides && !LT.tides_role) put3(lds, C_TF, lane, 0., 0., 0.);
        if (LT.tides_role) {
            // radial component (constant_time_lag.rs:266-307)
            const double hm2 = hm * hm;
            const double pm2 = m * m;
            const double conservative = H ? geo.conservative
                                          : divr(-3.0 * K2, Rd7) * (pm2 * p5(hR) * lds.k2[hl] + hm2 * p5(R) * lds.k2[lane]);
            const double factor1 = H ? divr(-13.5 * vr, geo.Rp8) : -13.5 * vr / p8(d);
            const double term1 = H ? geo.term1 : pm2 * p10(hR) * sigma_host;
            const double term2 = H ? geo.term2 : hm2 * p10(R) * sigma_pair;
            const double diss_pm = factor1 * term2;
            const double dissipative = diss_pm + factor1 * term1;
            const double radial = conservative + dissipative;
            // force (constant_time_lag.rs:309-332)
            const double factor3 = radial + divr((os + op) * vr, Rd);
            const double od = H ? geo.od : divr(os, Rd), pd = H ? geo.pd : divr(op, Rd);
            const double F0 = divr(factor3 * px[0], Rd) + od * (hs[1] * px[2] - hs[2] * px[1] - pv[0])
                            + pd * (s[1] * px[2] - s[2] * px[1] - pv[0]);
            const double F1 = divr(factor3 * px[1], Rd) + od * (hs[2] * px[0] - hs[0] * px[2] - pv[1])
                            + pd * (s[2] * px[0] - s[0] * px[2] - pv[1]);
            const double F2 = divr(factor3 * px[2], Rd) + od * (hs[0] * px[1] - hs[1] * px[0] - pv[2])
                            + pd * (s[0] * px[1] - s[1] * px[0] - pv[2]);
            put3(lds, C_TF, lane, F0, F1, F2);
            a_add[0] = a_add[0] + B.inv_m * F0;
            a_add[1] = a_add[1] + B.inv_m * F1;
            a_add[2] = a_add[2] + B.inv_m * F2;
        }
        if (EXT && (B.flags & F_DISK)) {   // universe.rs:550-554: after tides, before flattening
#pragma unroll
            for (int q = 0; q < 3; ++q) a_add[q] = a_add[q] + dk[q];
        }
        if (LT.rf && !LT.rf_role) put3(lds, C_RF, lane, 0., 0., 0.);
        if (LT.rf_role) {
            // radial component (oblate_spheroid.rs:51-60) and force (83-97)
            const double radial = (H ? geo.m3d5 : -3. / d5) * (fp + fs)
                + (H ? geo.c15 : divr(15., Rd7)) * (fs * rs_star * rs_star / hs2 + fp * rs_planet * rs_planet / s2);
            const double F0 = radial * px[0] + ofp * s[0] + ofs * hs[0];
            const double F1 = radial * px[1] + ofp * s[1] + ofs * hs[1];
            const double F2 = radial * px[2] + ofp * s[2] + ofs * hs[2];
            put3(lds, C_RF, lane, F0, F1, F2);
            a_add[0] = a_add[0] + B.inv_m * F0;
            a_add[1] = a_add[1] + B.inv_m * F1;
            a_add[2] = a_add[2] + B.inv_m * F2;
        }
        if (LT.spin) {
            // torques (constant_time_lag.rs:171-204, tides/common.rs:223-261;
            // oblate_spheroid.rs:62-81, rotational_flattening/common.rs:165-201)
            double tdL[3] = {0., 0., 0.}, rdL[3] = {0., 0., 0.};
            if (LT.tides && !LT.tides_role) {
                // The host's tidal torque sums over every other body, whatever
                // its role (tides/common.rs:244-255, the test is on the host's
                // effect), and a body the tidal loops skip keeps the zero
                // distance, position and components it was created with, so
                // its term is 0 * (0 * s - 0 * 0 / 0 - 1 / 0 * 0) = NaN
                // (constant_time_lag.rs:171-204).
                const double nan = __builtin_nan("");
                put3(lds, C_TN, lane, nan, nan, nan);
            }
            if (LT.tides_role) {
                const double cx = px[1] * pv[2] - px[2] * pv[1];
                const double cy = px[2] * pv[0] - px[0] * pv[2];
                const double cz = px[0] * pv[1] - px[1] * pv[0];
                const double id = H ? geo.id : divr(1.0, Rd);
                tdL[0] = -1.0 * (op * (d * s[0] - divr(rs_planet * px[0], Rd) - id * cx));
                tdL[1] = -1.0 * (op * (d * s[1] - divr(rs_planet * px[1], Rd) - id * cy));
                tdL[2] = -1.0 * (op * (d * s[2] - divr(rs_planet * px[2], Rd) - id * cz));
                put3(lds, C_TN, lane,
                     -1.0 * (os * (d * hs[0] - divr(rs_star * px[0], Rd) - id * cx)),
                     -1.0 * (os * (d * hs[1] - divr(rs_star * px[1], Rd) - id * cy)),
                     -1.0 * (os * (d * hs[2] - divr(rs_star * px[2], Rd) - id * cz)));
            }
            if (LT.rf && !LT.rf_role) put3(lds, C_RN, lane, 0., 0., 0.);   // its torque term is -1 * 0 * 0
            if (LT.rf_role) {
                rdL[0] = -1.0 * (ofp * (px[1] * s[2] - px[2] * s[1]));
                rdL[1] = -1.0 * (ofp * (px[2] * s[0] - px[0] * s[2]));
                rdL[2] = -1.0 * (ofp * (px[0] * s[1] - px[1] * s[0]));
                put3(lds, C_RN, lane,
                     -1.0 * (ofs * (px[1] * hs[2] - px[2] * hs[1])),
                     -1.0 * (ofs * (px[2] * hs[0] - px[0] * hs[2])),
                     -1.0 * (ofs * (px[0] * hs[1] - px[1] * hs[0])));
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) dLdt[q] = tdL[q] + rdL[q];
        }
        if (GRV == GRV_KIDDER && LT.gr && !LT.gr_role) {   // skipped by the Kidder loops
            put3(lds, C_G1, lane, 0., 0., 0.);
            put3(lds, C_G2, lane, 0., 0., 0.);
            put3(lds, C_G3, lane, 0., 0., 0.);
            put3(lds, C_GL, lane, 0., 0., 0.);
        }
        if (GRV == GRV_KIDDER && LT.gr_role) {
            // Kidder 1995 (general_relativity.rs:177-456)
            const Rcp Rhm{hm, lds.inv_m[hl]};
            const Rcp Rm{m, B.inv_m};
            const double hmg = mg_of(lds, hl), pmg = mg_of(lds, lane);
            const double f = lds.kf[lane];
            const double mg = hmg + pmg;
            const double distance_2 = H ? 0. : p2(d);
            const Rcp Rdc = H ? Rcp{0., 0.} : rcp(distance_2 * SPEED_OF_LIGHT_2);
            const double mgc = H ? geo.mgc : divr(-mg, Rdc);          // -mg / (d^2 c^2)
            const double mg_d = H ? geo.mg_d : divr(mg, Rd);
            const double mOverM = divr(m, Rhm);
            const Rcp Rnv = rcp(nv);
            double gr[3];
            {   // 1PN
                const double radial_velocity_2 = p2(vr);
                const double radial = mgc * ((1.0 + 3.0 * f) * nv2 - (H ? geo.c1pn : divr(2.0 * (2.0 + f) * mg, Rd)) - 1.5 * f * radial_velocity_2);
                const double orthogonal = (H ? geo.o1pn : divr(mg, Rdc) * 2.0 * (2.0 - f)) * vr * nv;
#pragma unroll
                for (int q = 0; q < 3; ++q) gr[q] = divr(radial * px[q], Rd) + divr(orthogonal * pv[q], Rnv);
                put3(lds, C_G1, lane, mOverM * gr[0], mOverM * gr[1], mOverM * gr[2]);
            }
            {   // 2PN
                const double nv4 = p2(nv2);
                const double vr2 = p2(vr);
                const double vr4 = p2(vr2);
                const double f2 = p2(f);
                const double radial = mgc
                    * ((H ? geo.c2pn : 3.0 / 4.0 * (12.0 + 29.0 * f) * (p2(mg) / distance_2))
                       + f * (3.0 - 4.0 * f) * nv4
                       + 15.0 / 8.0 * f * (1.0 - 3.0 * f) * vr4
                       - 3.0 / 2.0 * f * (3.0 - 4.0 * f) * vr2 * nv2
                       - 0.5 * f * (13.0 - 4.0 * f) * mg_d * nv2
                       - (2.0 + 25.0 * f + 2.0 * f2) * mg_d * vr2);
                const double orthogonal = mgc * (-0.5) * vr
                    * (f * (15.0 + 4.0 * f) * nv2
                       - (4.0 + 41.0 * f + 8.0 * f2) * mg_d
                       - 3.0 * f * (3.0 + 2.0 * f) * vr2);
                double a[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    a[q] = divr(radial * px[q], Rd) + orthogonal * pv[q];
                    gr[q] = gr[q] + a[q];
                }
                put3(lds, C_G2, lane, mOverM * a[0], mOverM * a[1], mOverM * a[2]);
            }
            {   // 1.5PN spin-orbit
                const double hI = lds.I[hl], I = lds.I[lane];
                const double Ls[3] = {hI * hs[0], hI * hs[1], hI * hs[2]};
                const double M = hm + m;
                const Rcp RM{M, lds.inv_M[lane]};
                const double mass_factor = divr(hm - m, RM);
                const double Lp[3] = {I * s[0], I * s[1], I * s[2]};
                const double n3[3] = {H ? geo.n3[0] : divr(px[0], Rd), H ? geo.n3[1] : divr(px[1], Rd),
                                      H ? geo.n3[2] : divr(px[2], Rd)};
                const double factor_a = G / SPEED_OF_LIGHT_2;
                double a[3];
                {
                    double msf[3], e7[3], e3s[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        msf[q] = mass_factor * M * (divr(Lp[q], Rm) - divr(Ls[q], Rhm));
                        e7[q] = 7. * (Ls[q] + Lp[q]) + 3. * msf[q];
                        e3s[q] = 3. * (Ls[q] + Lp[q]) + msf[q];
                    }
                    const double e10 = 6. * n3[0] * ((n3[1] * pv[2] - n3[2] * pv[1]) * (2. * (Ls[0] + Lp[0]) + msf[0]));
                    const double e11 = 6. * n3[1] * ((n3[2] * pv[0] - n3[0] * pv[2]) * (2. * (Ls[1] + Lp[1]) + msf[1]));
                    const double e12 = 6. * n3[2] * ((n3[0] * pv[1] - n3[1] * pv[0]) * (2. * (Ls[2] + Lp[2]) + msf[2]));
                    a[0] = factor_a * (e10 - (pv[1] * e7[2] - pv[2] * e7[1]) + 3. * vr * (n3[1] * e3s[2] - n3[2] * e3s[1]));
                    a[1] = factor_a * (e11 - (pv[2] * e7[0] - pv[0] * e7[2]) + 3. * vr * (n3[2] * e3s[0] - n3[0] * e3s[2]));
                    a[2] = factor_a * (e12 - (pv[0] * e7[1] - pv[1] * e7[0]) + 3. * vr * (n3[0] * e3s[1] - n3[1] * e3s[0]));
                }
#pragma unroll
                for (int q = 0; q < 3; ++q) gr[q] = gr[q] + a[q];
                put3(lds, C_G3, lane, mOverM * a[0], mOverM * a[1], mOverM * a[2]);
                const double mu = divr(hm * m, RM);
                const double Lo[3] = {mu * (px[1] * pv[2] - px[2] * pv[1]), mu * (px[2] * pv[0] - px[0] * pv[2]),
                                      mu * (px[0] * pv[1] - px[1] * pv[0])};
                double fm = 2. + divr(3. / 2. * m, Rhm);
                const double sp = n3[0] * Lp[0] + n3[1] * Lp[1] + n3[2] * Lp[2];
                put3(lds, C_GL, lane,
                     factor_a * (fm * (Lo[1] * Ls[2] - Lo[2] * Ls[1]) - (Lp[1] * Ls[2] - Lp[2] * Ls[1]) + 3. * sp * (n3[1] * Ls[2] - n3[2] * Ls[1])),
                     factor_a * (fm * (Lo[2] * Ls[0] - Lo[0] * Ls[2]) - (Lp[2] * Ls[0] - Lp[0] * Ls[2]) + 3. * sp * (n3[2] * Ls[0] - n3[0] * Ls[2])),
                     factor_a * (fm * (Lo[0] * Ls[1] - Lo[1] * Ls[0]) - (Lp[0] * Ls[1] - Lp[1] * Ls[0]) + 3. * sp * (n3[0] * Ls[1] - n3[1] * Ls[0])));
                fm = 2. + divr(3. / 2. * hm, Rm);
                const double sh = n3[0] * Ls[0] + n3[1] * Ls[1] + n3[2] * Ls[2];
                dLdt[0] = dLdt[0] + factor_a * (fm * (Lo[1] * Lp[2] - Lo[2] * Lp[1]) - (Ls[1] * Lp[2] - Ls[2] * Lp[1]) + 3. * sh * (n3[1] * Lp[2] - n3[2] * Lp[1]));
                dLdt[1] = dLdt[1] + factor_a * (fm * (Lo[2] * Lp[0] - Lo[0] * Lp[2]) - (Ls[2] * Lp[0] - Ls[0] * Lp[2]) + 3. * sh * (n3[2] * Lp[0] - n3[0] * Lp[2]));
                dLdt[2] = dLdt[2] + factor_a * (fm * (Lo[0] * Lp[1] - Lo[1] * Lp[0]) - (Ls[0] * Lp[1] - Ls[1] * Lp[0]) + 3. * sh * (n3[0] * Lp[1] - n3[1] * Lp[0]));
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) a_add[q] = a_add[q] + gr[q];
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) dLdt[q] = dLdt[q] + 0.;
        }
        if (H) {   // for the next evaluation, whose spins are refreshed before it reads them
            geo.rs_star = px[0] * hs[0] + px[1] * hs[1] + px[2] * hs[2];
            geo.rs_planet = px[0] * s[0] + px[1] * s[1] + px[2] * s[2];
        }
    }
    wsync();
    // ---- host reactions, accumulated in body order ----
    if (B.valid()) host_reaction_sums(lds, B);
    wsync();
    if (B.valid() && B.is_host()) {
        auto row = [&](int q) { return lds.xb[q][lane]; };   // this lane's column
        if (LT.tides) {   // tides/common.rs:312-345 and 223-261
#pragma unroll
            for (int q = 0; q < 3; ++q) a_add[q] = a_add[q] + -1.0 * B.inv_m * row(C_TF + q);
            if (LT.spin) {
#pragma unroll
                for (int q = 0; q < 3; ++q) dLdt[q] = row(C_TN + q);
            }
        }
        if (EXT && (B.flags & F_DISK)) {
#pragma unroll
            for (int q = 0; q < 3; ++q) a_add[q] = a_add[q] + dk[q];
        }
        if (LT.rf) {      // rotational_flattening/common.rs:203-237 and 165-201
#pragma unroll
            for (int q = 0; q < 3; ++q) a_add[q] = a_add[q] + -1.0 * B.inv_m * row(C_RF + q);
#pragma unroll
            for (int q = 0; q < 3; ++q) dLdt[q] = dLdt[q] + (LT.spin ? row(C_RN + q) : 0.);
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) dLdt[q] = dLdt[q] + 0.;
        }
        if (GRV == GRV_KIDDER && LT.gr) {      // general_relativity.rs:236-238, 295-297, 453-455, 312-315
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double g = -1.0 * row(C_G1 + q);
                g = g + -1.0 * row(C_G2 + q);
                g = g + -1.0 * row(C_G3 + q);
                a_add[q] = a_add[q] + g;
                dLdt[q] = dLdt[q] + row(C_GL + q);
            }
        } else {
#pragma unroll
            for (int q = 0; q < 3; ++q) dLdt[q] = dLdt[q] + 0.;
        }
    }
    wsync();
    // ---- the other GR implementations: accelerations only, added last ----
    if (GRV != GRV_KIDDER) {   // uniform: such a launch holds only systems of this variant
        double g[3];
        if (GRV == GRV_ANDERSON) gr_anderson1975(lds, B, g);
        else gr_newhall1983(lds, B, g);
        if (B.valid() && LT.gr) {
#pragma unroll
            for (int q = 0; q < 3; ++q) a_add[q] = a_add[q] + g[q];
        }
    }
    // ---- wind (wind.rs:61-91), own body, added last ----
    double w[3] = {0., 0., 0.};
    if (B.valid() && (B.flags & F_WIND) && LT.spin && (B.flags & F_WIND_ROLE)) {
        const double wk = E.wind_k[B.body], wsat = E.wind_sat[B.body];
        const double s2 = spin2(lds, lane);
        const double threshold = sqrt(s2);
        const double shape = sqrt(lds.R[lane] / R_SUN * 1. / B.m);
        const double f = threshold >= wsat ? wsat * wsat : s2;
        w[0] = -1. * wk * lds.s[0][lane] * f * shape;
        w[1] = -1. * wk * lds.s[1][lane] * f * shape;
        w[2] = -1. * wk * lds.s[2][lane] * f * shape;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) dLdt[q] = dLdt[q] + w[q];
}

template <int V>
__device__ inline void integrate_velocity_dependent_forces(const DevEnsemble &E, WaveLds &lds, Lane &B, bool evolution) {
    const double h = 0.5 * B.dt;
    const bool spin_on = LaneTests<V>(B).spin;
    // The final state of an iteration is v_orig + dv (and L_orig + dL), so only
    // the increments are carried from one iteration to the next; the previous
    // final state and the midpoint are recomputed from them, bit for bit.
    double v_orig[3], L_orig[3], dv[3] = {0., 0., 0.}, dL[3] = {0., 0., 0.};
#pragma unroll
    for (int q = 0; q < 3; ++q) { v_orig[q] = B.v[q]; L_orig[q] = B.L[q]; }
    constexpr bool H = (V & VAR_FAST) != 0;   // fixed radii and factors: hoist the pair terms
    PairGeo geo;
    if (H) pair_geometry(lds, B, geo);
    bool converged = false;
    for (int it = 0; it < MIDPOINT_MAX_ITER; ++it) {
        double a_add[3], dLdt[3];
        additional_effects<V, H>(E, lds, B, evolution && it == 0, a_add, dLdt, geo);
        double pdv = 0., pdL = 0.;   // |change of the final state|^2 of this body
        double fv2 = 0., fL2 = 0.;   // |final state|^2 of this body
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double v_prev = it == 0 ? v_orig[q] : v_orig[q] + dv[q];
            dv[q] = h * a_add[q] - B.ev[q];
            const double vf = v_orig[q] + dv[q];
            pdv = pdv + p2(vf - v_prev);
            fv2 = fv2 + p2(vf);
            B.v[q] = 0.5 * (v_orig[q] + vf);
            if (spin_on) {
                const double L_prev = it == 0 ? L_orig[q] : L_orig[q] + dL[q];
                dL[q] = h * dLdt[q] - B.eL[q];
                const double Lf = L_orig[q] + dL[q];
                pdL = pdL + p2(Lf - L_prev);
                fL2 = fL2 + p2(Lf);
                B.L[q] = 0.5 * (L_orig[q] + Lf);
            }
        }
        if (it >= MIDPOINT_MIN_ITER - 1) {
            // whfast.rs:424-451: four sums over the bodies in body order, which
            // every lane of the group forms itself from the exchange rows
            if (B.valid()) {
                lds.xb[0][B.lane] = pdv;
                lds.xb[1][B.lane] = fv2;
                lds.xb[2][B.lane] = pdL;
                lds.xb[3][B.lane] = fL2;
            }
            wsync();
            double sums[4] = {0., 0., 0., 0.};
            fold_rows<4, false>(B, sums, [&](int r, int l) { return lds.xb[r][l]; });
            const double dv2 = sums[0], sv2 = sums[1], dL2 = sums[2], sL2 = sums[3];
            wsync();
            const bool ok = (spin_on && dv2 / sv2 < DBL_EPS_2 && dL2 / sL2 < DBL_EPS_2) ||
                            (!spin_on && dv2 / sv2 < DBL_EPS_2);
            if (ok) { converged = true; break; }
        }
    }
    if (!converged) B.status |= POS_STATUS_MIDPOINT_NOT_CONVERGED;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        B.v[q] = v_orig[q] + dv[q];
        B.ev[q] = (B.v[q] - v_orig[q]) - dv[q];
        if (spin_on) {
            B.L[q] = L_orig[q] + dL[q];
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
    B.kx = x;
}

// Jacobi prefix sums up to body k (whfast.rs:574-576, 642-651): the
// cumulative G-mass and mass of the host and the bodies before k, k included.
__device__ inline void jacobi_prefix(const WaveLds &lds, const Lane &B, double *mu, double *eta, int *rank) {
    const int hl = B.hl();
    double m_g = mg_of(lds, hl), e = lds.m[hl];
    int r = 0;
    for (int j = 0; j < B.n; ++j) {
        if (j == B.host) continue;
        m_g = m_g + mg_of(lds, B.gbase + j);
        e = e + lds.m[B.gbase + j];
        if (j == B.k) break;
        ++r;
    }
    *mu = m_g; *eta = e; *rank = r;
}

__device__ inline void kepler_steps(const WaveLds &lds, Lane &B, double dt) {
    if (B.orbiting()) {
        double mu;
        if (B.coord == POS_COORD_JACOBI) {
            double eta;
            int rank;
            jacobi_prefix(lds, B, &mu, &eta, &rank);
        } else if (B.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) {
            mu = mg_of(lds, B.hl());
        } else {
            mu = mg_of(lds, B.hl()) + mg_of(lds, B.lane);
        }
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
    if (B.valid()) {
        lds.xb[0][lane] = B.x[0]; lds.xb[1][lane] = B.x[1]; lds.xb[2][lane] = B.x[2];
        lds.xb[3][lane] = B.v[0]; lds.xb[4][lane] = B.v[1]; lds.xb[5][lane] = B.v[2];
    }
    wsync();
    const int hl = B.hl();
    if (B.coord == POS_COORD_JACOBI) {
        // whfast.rs:889-933: prefix recurrence, recomputed by every lane
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
        // whfast.rs:973-1023: centre of mass, host first then the others in
        // order; the seven ordered sums (x, v weighted by the masses, and the
        // masses themselves as 1 * m from row 13) are spread over the group's
        // lanes and left in rows 6-12 of column gbase
        if (B.valid()) lds.xb[13][lane] = 1.;
        wsync();
        for (int q = B.k; B.valid() && q < 7; q += B.n) {
            const int rq = q < 6 ? q : 13;
            double acc[1] = {0. + lds.xb[rq][hl] * lds.m[hl]};
            fold_rows<1, true>(B, acc, [&](int, int l) { return lds.xb[rq][l] * lds.m[l]; });
            lds.xb[6 + q][B.gbase] = acc[0];
        }
        wsync();
        const double cx[3] = {lds.xb[6][B.gbase], lds.xb[7][B.gbase], lds.xb[8][B.gbase]};
        const double cv[3] = {lds.xb[9][B.gbase], lds.xb[10][B.gbase], lds.xb[11][B.gbase]};
        const double cm = lds.xb[12][B.gbase];
        B.com_m = cm;
        const Rcp Rcm = rcp(cm);
#pragma unroll
        for (int q = 0; q < 3; ++q) { B.comx[q] = divr(cx[q], Rcm); B.comv[q] = divr(cv[q], Rcm); }
        if (B.orbiting()) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                B.ax[q] = B.x[q] - lds.xb[q][hl];
                B.av[q] = B.v[q] - B.comv[q];
            }
            if (B.coord == POS_COORD_WHDS) {
                const double hm = lds.m[hl];
                const double factor = divr(hm + B.m, Rcp{hm, lds.inv_m[hl]});
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
    if (B.orbiting()) {
        lds.xb[0][lane] = B.av[0]; lds.xb[1][lane] = B.av[1]; lds.xb[2][lane] = B.av[2];
    }
    wsync();
    const int hl = B.hl();
    const double m0 = lds.m[hl];
    const Rcp Rm0{m0, lds.inv_m[hl]};
    // three ordered momentum sums, spread over the lanes (valid lanes only:
    // an empty group has n = 0), left in the host's column
    for (int q = B.k; B.valid() && q < 3; q += B.n) {
        double acc[1] = {0.};
        if (B.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC)
            fold_rows<1, true>(B, acc, [&](int, int l) { return lds.m[l] * lds.xb[q][l]; });
        else
            fold_rows<1, true>(B, acc, [&](int, int l) {
                const double mj = lds.m[l];
                return divr(mj * lds.xb[q][l], rcp(m0 + mj));
            });
        lds.xb[q][hl] = acc[0];
    }
    wsync();
    const double p[3] = {lds.xb[0][hl], lds.xb[1][hl], lds.xb[2][hl]};
    wsync();
    if (B.orbiting()) {
        if (B.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) {
#pragma unroll
            for (int q = 0; q < 3; ++q) B.ax[q] = B.ax[q] + divr(dt * p[q], Rm0);
        } else {
            const Rcp Rf = rcp(m0 + B.m);
#pragma unroll
            for (int q = 0; q < 3; ++q) B.ax[q] = B.ax[q] + dt * (p[q] - divr(B.m * B.av[q], Rf));
        }
    }
}

__device__ inline void interaction_step(WaveLds &lds, Lane &B, double dt) {
    const int lane = B.lane;
    const int hl = B.hl();
    if (B.coord == POS_COORD_JACOBI) {
        // whfast.rs:935-963 then 566-592; the gravitational accelerations are
        // in the persistent LDS area
        double eta = lds.m[hl];
        double s[3], aj_own[3] = {0., 0., 0.};
#pragma unroll
        for (int q = 0; q < 3; ++q) s[q] = eta * lds.ag[q][hl];
        for (int j = 0; j < B.n; ++j) {
            if (j == B.host) continue;
            const int l = B.gbase + j;
            const double mj = lds.m[l];
            const double ei = 1. / eta;
            eta = eta + mj;
            const double pme = eta * ei;
            double aj[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) aj[q] = lds.ag[q][l] - s[q] * ei;
            if (j == B.k) {
#pragma unroll
                for (int q = 0; q < 3; ++q) aj_own[q] = aj[q];
            }
#pragma unroll
            for (int q = 0; q < 3; ++q) s[q] = s[q] * pme + mj * aj[q];
        }
        if (B.orbiting()) {
#pragma unroll
            for (int q = 0; q < 3; ++q) B.av[q] = B.av[q] + dt * aj_own[q];
            double mu, jeta;
            int rank;
            jacobi_prefix(lds, B, &mu, &jeta, &rank);
            if (rank > 0) {
                const double rj2i = 1. / (p2(B.ax[0]) + p2(B.ax[1]) + p2(B.ax[2]) + 1e-12);
                const double rji = sqrt(rj2i);
                const double rj3im = rji * rj2i * G * jeta;
                const double prefac = dt * rj3im;
#pragma unroll
                for (int q = 0; q < 3; ++q) B.av[q] = B.av[q] + prefac * B.ax[q];
            }
        }
        return;
    }
    if (B.orbiting()) {
        if (B.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) {
#pragma unroll
            for (int q = 0; q < 3; ++q) B.av[q] = B.av[q] + dt * lds.ag[q][lane];
        } else {
            const double hm = lds.m[hl];
            const double factor = hm + B.m;
            const Rcp Rhm{hm, lds.inv_m[hl]};
#pragma unroll
            for (int q = 0; q < 3; ++q) B.av[q] = B.av[q] + divr(dt * factor * lds.ag[q][lane], Rhm);
        }
    }
}

__device__ inline void alternative_to_inertial(WaveLds &lds, Lane &B) {
    const int lane = B.lane;
    if (B.orbiting()) {
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
        if (B.valid() && B.is_host()) {
            const double mtot_i = 1. / eta;
#pragma unroll
            for (int q = 0; q < 3; ++q) { B.x[q] = s[q] * mtot_i; B.v[q] = sv[q] * mtot_i; }
        }
    } else {
        // whfast.rs:1090-1155
        const int hl = B.hl();
        const Rcp Rmt = rcp(B.com_m);
        const double m0 = lds.m[hl];
        const Rcp Rm0{m0, lds.inv_m[hl]};
        // six ordered chains (host position and velocity), one per lane,
        // left in the host's column
        // (acc - t is acc + (-t) bit for bit; a skipped slot's term is -0,
        // which leaves every value, -0 included, unchanged)
        for (int q = B.k; B.valid() && q < 6; q += B.n) {
            // constant indices only: a dynamic index would move comx/comv to the stack
            double acc[1] = {q == 0 ? B.comx[0] : q == 1 ? B.comx[1] : q == 2 ? B.comx[2]
                             : q == 3 ? B.comv[0] : q == 4 ? B.comv[1] : B.comv[2]};
            const bool whds = B.coord == POS_COORD_WHDS;
            fold_rows<1, true>(B, acc, [&](int, int l) {
                const double mj = lds.m[l], xq = lds.xb[q][l];
                const double tx = divr(xq * mj, Rmt);
                const double tv = xq * (whds ? mj / (m0 + mj) : divr(mj, Rm0));
                return -(q < 3 ? tx : tv);
            });
            lds.xb[q][hl] = acc[0];
        }
        wsync();
        const double nsp[3] = {lds.xb[0][hl], lds.xb[1][hl], lds.xb[2][hl]};
        const double nsv[3] = {lds.xb[3][hl], lds.xb[4][hl], lds.xb[5][hl]};
        if (B.valid()) {
            if (B.is_host()) {
#pragma unroll
                for (int q = 0; q < 3; ++q) { B.x[q] = nsp[q]; B.v[q] = nsv[q]; }
            } else {
#pragma unroll
                for (int q = 0; q < 3; ++q) B.x[q] = B.ax[q] + nsp[q];
                if (B.coord == POS_COORD_WHDS) {
                    const Rcp Rf = rcp(divr(m0 + B.m, Rm0));
#pragma unroll
                    for (int q = 0; q < 3; ++q) B.v[q] = divr(B.av[q], Rf) + B.comv[q];
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
    if (B.valid()) { lds.xb[0][lane] = B.x[0]; lds.xb[1][lane] = B.x[1]; lds.xb[2][lane] = B.x[2]; }
    wsync();
    if (B.valid()) {
        const int i = B.k, mm = B.host;
        const bool whfast_one = (B.coord == POS_COORD_JACOBI);
        const double Ri = lds.R[lane];
        double acc[3] = {0., 0., 0.};
        for (int j = 0; j < B.n; ++j) {
            if (j == i) continue;
            const int l = B.gbase + j;
            const double dx = B.x[0] - lds.xb[0][l];
            const double dy = B.x[1] - lds.xb[1][l];
            const double dz = B.x[2] - lds.xb[2][l];
            const double distance_2 = dx * dx + dy * dy + dz * dz;
            if (i < j) {
                // the table is read only when the screen cannot exclude the pair
                if (distance_2 <= B.roche_max2) {
                    const double rr = E.roche[(size_t)B.body * POS_MAX_BODIES + j];
                    if (distance_2 <= rr * rr) B.status |= POS_STATUS_ROCHE;
                }
                if (distance_2 <= p2(Ri + lds.R[l])) B.status |= POS_STATUS_COLLISION;
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
        lds.ag[0][lane] = acc[0]; lds.ag[1][lane] = acc[1]; lds.ag[2][lane] = acc[2];
    }
    wsync();
}

// Status bits raised by one lane (gravity checks, zero inertia, Kepler) are
// OR-ed over the group so that every lane agrees, and the time-step warning
// count is the group's largest.  Both change only on rare events: when no
// active lane of the wavefront holds a status or count other than the one
// its group agreed on last (a wavefront-wide vote), every lane already holds
// the group's OR and maximum, and the exchange is skipped.
__device__ inline void merge_status(WaveLds &lds, Lane &B) {
    if (!__any(B.status != B.seen_status || B.tw != B.seen_tw)) return;
    lds.xb[0][B.lane] = __builtin_bit_cast(double, (unsigned long long)B.status);
    lds.xb[1][B.lane] = (double)B.tw;
    wsync();
    uint32_t st = 0;
    int tw = 0;
    for (int j = 0; j < B.n; ++j) {
        st |= (uint32_t)__builtin_bit_cast(unsigned long long, lds.xb[0][B.gbase + j]);
        tw = max(tw, (int)lds.xb[1][B.gbase + j]);
    }
    wsync();
    B.status = st;
    B.tw = tw;
    B.seen_status = st;
    B.seen_tw = tw;
}

// ===========================================================================
// Historic snapshot record (whfast.rs:240-256, output.rs:90-134): the state
// after the snapshot branch's spin and evolution refresh.  denergy_dt
// (tides/common.rs:263-281, Bolmont et al. 2015 eq. 32) is evaluated at this
// state: the reference evaluates it from the tidal internals of the last
// effects evaluation of the previous step, half a step earlier.
// ===========================================================================
template <bool EXT>
__device__ inline void write_history_record(const DevEnsemble &E, WaveLds &lds, Lane &B) {
    const int lane = B.lane;
    if (B.valid()) {
        lds.xb[0][lane] = B.x[0]; lds.xb[1][lane] = B.x[1]; lds.xb[2][lane] = B.x[2];
        lds.xb[3][lane] = B.v[0]; lds.xb[4][lane] = B.v[1]; lds.xb[5][lane] = B.v[2];
    }
    wsync();
    double dedt = 0.;   // a body the tidal loops skip keeps its initial 0 (tides/common.rs:263-281)
    if (B.tides_role()) {
        const int hl = B.hl();
        double r[3], u[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) { r[q] = B.x[q] - lds.xb[q][hl]; u[q] = B.v[q] - lds.xb[3 + q][hl]; }
        const double d = sqrt(p2(r[0]) + p2(r[1]) + p2(r[2]));
        const double vr = (r[0] * u[0] + r[1] * u[1] + r[2] * u[2]) / d;
        const double hm = lds.m[hl];
        const double R = lds.R[lane];
        double sigma = lds.sigma[lane];   // or its pair-dependent map entry
        if (EXT && (B.flags & F_DYN) && E.pair_ph[B.body] == E.pair_ph[B.body]) sigma = E.pair_ph[B.body];
        const double op = 4.5 * p2(hm) * p10(R) * sigma / p7(d);
        const double diss_pm = -13.5 * vr / p8(d) * (p2(hm) * p10(R) * sigma);
        const double f2 = op / d;
        const double sp[3] = {lds.s[0][lane], lds.s[1][lane], lds.s[2][lane]};
        const double rs = r[0] * sp[0] + r[1] * sp[1] + r[2] * sp[2];
        const double c[3] = {r[1] * u[2] - r[2] * u[1], r[2] * u[0] - r[0] * u[2], r[0] * u[1] - r[1] * u[0]};
        double torque_dot_spin = 0.;
#pragma unroll
        for (int q = 0; q < 3; ++q) torque_dot_spin += -op * (d * sp[q] - rs * r[q] / d - c[q] / d) * sp[q];
        const double ru = r[0] * u[0] + r[1] * u[1] + r[2] * u[2];
        const double w0 = sp[1] * r[2] - sp[2] * r[1] - u[0];
        const double w1 = sp[2] * r[0] - sp[0] * r[2] - u[1];
        const double w2 = sp[0] * r[1] - sp[1] * r[0] - u[2];
        dedt = -((1.0 / d * (diss_pm + f2 * vr)) * ru + f2 * (w0 * u[0] + w1 * u[1] + w2 * u[2])) - torque_dot_spin;
    }
    wsync();
    const int slot = B.hcount;
    if (slot >= E.hist_cap) {
        B.status |= POS_STATUS_HISTORY_OVERFLOW;
        return;
    }
    if (B.valid()) {
        const size_t nb = (size_t)E.n_bodies;
        double *rec = E.hist + (size_t)slot * HIST_FIELDS * nb + B.body;
        rec[HF_TIME * nb] = B.t;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            rec[(HF_X + q) * nb] = B.x[q];
            rec[(HF_SPIN + q) * nb] = lds.s[q][lane];
            rec[(HF_V + q) * nb] = B.v[q];
        }
        rec[HF_RADIUS * nb] = lds.R[lane];
        rec[HF_RG2 * nb] = E.rg2[B.body];
        rec[HF_LOVE * nb] = lds.k2[lane];
        rec[HF_DEDT * nb] = dedt;
        rec[HF_LAG * nb] = EXT ? E.lag[B.body] : 0.;
        rec[HF_MIGRATION * nb] = EXT ? E.disk_migration[B.body] : 0.;
    }
    B.hcount = slot + 1;
}

// ===========================================================================
// Step tape: what the reverse sweep of the adjoint needs from each step,
// written by the recording variant of the step.  Row r of a segment holds,
// per body (component-major, body index fastest), the state at the start of
// step r (x, v, L), the velocity and angular momentum after the first kick
// (v1, L1) and the universal anomalies of the two Kepler drifts.  Row
// n_steps holds the state after the last step in its x, v, L fields.
// ===========================================================================
constexpr int TAPE_FIELDS = 19;
enum TapeField { TP_X = 0, TP_V = 3, TP_L = 6, TP_V1 = 9, TP_L1 = 12, TP_XK1 = 15, TP_XK2 = 16, TP_R = 17, TP_I = 18 };

struct TapeRef {
    double *base;      // [rows][TAPE_FIELDS][n_bodies]
    int n_bodies;
    __device__ __forceinline__ double *at(int row, int field, int body) const {
        return base + ((size_t)row * TAPE_FIELDS + field) * (size_t)n_bodies + body;
    }
};

__device__ __forceinline__ void tape_put3(const TapeRef &T, int row, int field, int body, const double v[3]) {
#pragma unroll
    for (int q = 0; q < 3; ++q) *T.at(row, field + q, body) = v[q];
}

// ===========================================================================
// WHFast::iterate (whfast.rs:235-305).  With REC, the step also writes its
// tape row `row` (and the end state into row + 1).
// ===========================================================================
template <bool REC, int V>
__device__ inline void iterate(const DevEnsemble &E, WaveLds &lds, Lane &B, const TapeRef &T, int row) {
    {
        const bool first = B.last_hist < 0.;
        const double period = B.hist_period;
        if (first || B.last_hist + period <= B.t) {
            if (B.flags & F_EVOLUTION) evolve_body(E, lds, B, B.t);
            own_spin(lds, B);
            if ((V & VAR_EXT) && (B.flags & F_EVOLUTION)) evolve_spin_dependent(E, lds, B, B.t);
            wsync();
            if (E.hist) write_history_record<(V & VAR_EXT) != 0>(E, lds, B);
            B.last_hist = first ? 0. : B.last_hist + period;
        }
    }
    const double half_dt = 0.5 * B.dt;
    if (REC && B.valid()) {
        tape_put3(T, row, TP_X, B.body, B.x);
        tape_put3(T, row, TP_V, B.body, B.v);
        tape_put3(T, row, TP_L, B.body, B.L);
    }
    integrate_velocity_dependent_forces<V>(E, lds, B, true);
    if (REC && B.valid()) {
        tape_put3(T, row, TP_V1, B.body, B.v);
        tape_put3(T, row, TP_L1, B.body, B.L);
        // the radius and moment of inertia both kicks of the step use (the
        // evolution tables are evaluated at the step's time in the first)
        *T.at(row, TP_R, B.body) = lds.R[B.lane];
        *T.at(row, TP_I, B.body) = lds.I[B.lane];
    }
    inertial_to_alternative(lds, B);
    kepler_steps(lds, B, half_dt);
    if (REC && B.valid()) *T.at(row, TP_XK1, B.body) = B.orbiting() ? B.kx : 0.;
    jump_step(lds, B, half_dt);
    alternative_to_inertial(lds, B);
    gravity(E, lds, B);
    interaction_step(lds, B, B.dt);
    jump_step(lds, B, half_dt);
    kepler_steps(lds, B, half_dt);
    if (REC && B.valid()) *T.at(row, TP_XK2, B.body) = B.orbiting() ? B.kx : 0.;
    alternative_to_inertial(lds, B);
    integrate_velocity_dependent_forces<V>(E, lds, B, false);
    if (REC && B.valid()) {
        tape_put3(T, row + 1, TP_X, B.body, B.x);
        tape_put3(T, row + 1, TP_V, B.body, B.v);
        tape_put3(T, row + 1, TP_L, B.body, B.L);
    }
    B.t = B.t + B.dt;
    B.iteration += 1;
    merge_status(lds, B);
    if (B.t + B.dt > B.time_limit) B.status |= POS_STATUS_COMPLETED;
}

}  // namespace posdev
