// posidonius_oracle.hpp -- CPU restatement of Posidonius' WHFast step.
//
// TEST INFRASTRUCTURE ONLY.  This header is the oracle that the HIP path is
// checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may use it; the product library never links it.
//
// It restates, for one system at a time, the reference algorithm of
//   src/integrator/whfast.rs        (WHFast::iterate and its operators)
//   src/particles/universe.rs       (gravity, heliocentric transform,
//                                    additional-effects orchestration)
//   src/particles/common.rs         (spin = L / I)
//   src/effects/tides/{common,constant_time_lag}.rs
//   src/effects/rotational_flattening/{common,oblate_spheroid}.rs
//   src/effects/general_relativity.rs (Kidder 1995)
//   src/effects/general_relativity.rs (Anderson 1975, Newhall 1983)
//   src/effects/wind.rs
//   src/effects/disk.rs
//   src/effects/evolution.rs:449-567 and src/tools.rs:840-907 (tables,
//   inverse tidal Q factor and lag angle of evolving dissipation)
//   src/effects/tides/constant_time_lag.rs:20-165 (pair-dependent factors)
// Data layout and decomposition are this project's own; what is kept alike
// is the order of every floating-point operation, because the golden
// fixtures are compared at 1e-14 absolute and the GPU path bitwise.  Rust's
// f64::powi with a constant exponent is expanded by LLVM into the binary
// square-and-multiply chain; pw2..pw10 below reproduce those chains.
// Compile with -ffp-contract=off: the reference never fuses multiply-adds.
//
// Deliberate departures from the reference:
//   * panic!() paths set status bits and freeze the system (the step in which
//     they fire still completes);
//   * loops the reference leaves unbounded (Stumpff argument reduction on an
//     infinite argument, Kepler bisection on NaN brackets) are capped and
//     raise POS_STATUS_ITERATION_CAP;
//   * a system must have its tidal, flattening and disk hosts equal to its
//     most massive body (the reference also allows other hosts), and
//     Anderson 1975 / Newhall 1983 need every other body to be a GR orbiting
//     body.  Other layouts are rejected at load;
//   * the pair-dependent factor map holds only what the step reads: the
//     host's entries are recomputed before every use, the orbiting bodies'
//     entries persist (loaded from and returned to the caller);
//   * exp and powf are correctly rounded (oracle_types.hpp), where the
//     reference's libm is within 0.52 ulp.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <limits>

#include "../include/posidonius_hip.h"

namespace oracle {

// ---- constants (src/constants.rs:3-74) -----------------------------------
constexpr int MAXB = POS_MAX_BODIES;
constexpr double DBL_EPS = 2.2204460492503131e-16;
constexpr double DBL_EPS_2 = DBL_EPS * DBL_EPS;
constexpr double PI = 3.14159265358979323846264338327950288;
constexpr double HOUR = 3600.;
constexpr double DAY = 24. * HOUR;
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
constexpr int WHFAST_NMAX_QUART = 64;
constexpr int WHFAST_NMAX_NEWT = 32;
constexpr int IMPLICIT_MIDPOINT_MIN_ITER = 3;
constexpr int IMPLICIT_MIDPOINT_MAX_ITER = 10;
// Caps for loops that are unbounded in the reference (see header comment).
constexpr int STUMPFF_REDUCTION_CAP = 600;   // 4^512 overflows binary64
constexpr int BISECTION_CAP = 4096;          // > bisection steps over the full exponent range

// ---- arithmetic helpers, generic over the scalar type ---------------------
template <class R> inline R pw2(const R &x) { return x * x; }
template <class R> inline R pw3(const R &x) { R x2 = x * x; return x * x2; }
template <class R> inline R pw4(const R &x) { R x2 = x * x; return x2 * x2; }
template <class R> inline R pw5(const R &x) { R x2 = x * x; R x4 = x2 * x2; return x * x4; }
template <class R> inline R pw7(const R &x) { R x2 = x * x; R x4 = x2 * x2; return (x * x2) * x4; }
template <class R> inline R pw8(const R &x) { R x2 = x * x; R x4 = x2 * x2; return x4 * x4; }
template <class R> inline R pw10(const R &x) { R x2 = x * x; R x4 = x2 * x2; R x8 = x4 * x4; return x2 * x8; }

// Scalar hooks: the FlopCount and Dual types provide their own overloads.
inline double val(double x) { return x; }
inline double sqrt_r(double x) { return std::sqrt(x); }
inline double abs_r(double x) { return std::fabs(x); }
inline double rfabs(double x) { return std::fabs(x); }
inline double rfloor(double x) { return std::floor(x); }
inline bool risnan(double x) { return std::isnan(x); }

template <class R> struct V3 {
    R x, y, z;
};
template <class R> inline V3<R> v3zero() { return V3<R>{R(0.), R(0.), R(0.)}; }

// ---- Evolver (src/effects/evolution.rs:19-27, 449-513) --------------------
struct Evolver {
    int type = POS_EVO_NON_EVOLVING;
    const double *time = nullptr, *radius = nullptr, *rg2 = nullptr, *love = nullptr, *itq = nullptr;
    int len = 0;
    int64_t left_index = 0;

    // evolution.rs:449-456; clamped to the last row, which a consistent
    // left_index never reaches (the reference would panic on the slice).
    int idx() const {
        const int i0 = left_index > 10 ? (int)(left_index - 10) : 0;
        return i0 < len ? i0 : len - 1;
    }

    // tools.rs:840-860 linear_interpolation + 882-905 find_indices_around_target_value
    template <class R>
    static R interpolate(const R &target, const double *x, const double *y, int n, int *left_out) {
        int first_greater = -1;
        for (int i = 0; i < n; ++i) {
            if (x[i] > val(target)) { first_greater = i; break; }
        }
        int left, right;
        if (first_greater < 0) {
            if (x[n - 1] > val(target)) { left = right = 0; }
            else { left = right = n - 1; }
        } else if (first_greater == 0) {
            left = right = 0;
        } else {
            left = first_greater - 1; right = first_greater;
        }
        *left_out = left;
        if (left == right) return R(y[left]);
        double x_left = x[left];
        R target_percent = (target - R(x_left)) / R(x[right] - x_left);
        return R(y[left]) * (R(1.) - target_percent) + R(y[right]) * target_percent;
    }
    template <class R> R eval_radius(const R &t, const R &current) {
        if (type == POS_EVO_NON_EVOLVING) return current;
        int i0 = idx(), left = 0;
        R r = interpolate(t, time + i0, radius + i0, len - i0, &left);
        left_index += left;
        return r;
    }
    template <class R> R eval_rg2(const R &t, const R &current) {
        if (type == POS_EVO_BARAFFE2015 || type == POS_EVO_LECONTE2011 ||
            type == POS_EVO_LECONTE_CHABRIER2013 || type == POS_EVO_LECONTE_CHABRIER2013_DYN) {
            int i0 = idx(), left = 0;
            R r = interpolate(t, time + i0, rg2 + i0, len - i0, &left);
            left_index += left;
            return r;
        }
        return current;
    }
    bool dynamical() const {
        return type == POS_EVO_BOLMONT_MATHIS2016 || type == POS_EVO_GALLET_BOLMONT2017 ||
               type == POS_EVO_LECONTE_CHABRIER2013_DYN;
    }
    // evolution.rs:487-513
    template <class R> R eval_inverse_tidal_q_factor(const R &t) {
        if (!dynamical()) return R(0.);
        int i0 = idx(), left = 0;
        R r = interpolate(t, time + i0, itq + i0, len - i0, &left);
        left_index += left;
        return r;
    }
    template <class R> R eval_love(const R &t, const R &current) {
        if (type == POS_EVO_LECONTE_CHABRIER2013 || type == POS_EVO_LECONTE_CHABRIER2013_DYN) {
            int i0 = idx(), left = 0;
            R r = interpolate(t, time + i0, love + i0, len - i0, &left);
            left_index += left;
            return r;
        }
        return current;
    }
};

// ---- one body (Particle, particle.rs:16-52, with the effect sub-structs) ---
template <class R> struct Body {
    R m, mg, radius, rg2, inertia;
    V3<R> x, v, a_grav, a_add;
    V3<R> hx, hv;            // heliocentric position / velocity
    R hdist, hvr, hnv2, hnv; // heliocentric distance, radial velocity, |v|^2, |v|
    V3<R> spin;
    R spin2;
    V3<R> L, dLdt;
    uint32_t roles;
    // parameters
    R sigma;      // scaled CTL dissipation factor (tides/common.rs:104-112)
    R sigma_scale, sigma_factor;   // its two factors (ConstantTimeLagParameters)
    R lag;        // lag angle of evolving dissipation (evolution.rs:548-567)
    R sigma_host; // the host's factor in this body's pair (constant_time_lag.rs:20-94, 196-204)
    R sigma_pair; // this body's own factor in that pair (constant_time_lag.rs:96-165, 196-204)
    bool ph_set;  // the map entry id * MAX_PARTICLES + host_id exists (persistent)
    R ph;         // its value
    R k2;         // CTL love number
    R k2f;        // oblate-spheroid (fluid) love number
    R wind_k, wind_sat, wind_sat2;
    // tides coordinates, internal, output.  The coordinates are copied from
    // the heliocentric ones for the host and the tidal orbiting bodies only
    // (tides/common.rs:202-217); every other body keeps the zeros it was
    // created with, and so do its internal values.
    V3<R> td_x, td_v;
    R td_d, td_vr;
    R td_rs_star, td_rs_planet, td_orth_star, td_orth_planet, td_radial, td_diss_pm;
    V3<R> td_acc, td_dLdt;
    // flattening coordinates (rotational_flattening/common.rs:152-163, same
    // rule as the tides), internal, output
    V3<R> rf_x;
    R rf_d;
    R rf_rs_star, rf_rs_planet, rf_radial, rf_fac_star, rf_fac_planet, rf_orth_star, rf_orth_planet;
    V3<R> rf_acc, rf_dLdt;
    // general relativity
    R gr_factor;
    V3<R> gr_acc, gr_dLdt;
    // wind
    V3<R> wind_dLdt;
    // disk (disk.rs:17-34)
    R disk_migration;
    V3<R> disk_acc;
};

template <class R> struct Alt {   // AlternativeCoordinates (whfast.rs:122-130)
    R m, mg;
    V3<R> x, v, a;
};

template <class R> struct System {
    int n = 0, host = 0, coord = POS_COORD_DEMOCRATIC_HELIOCENTRIC, gr_impl = POS_GR_DISABLED;
    bool tides = false, rf = false, gr = false, wind = false, evolution = false, disk = false;
    double disk_props[6] = {0., 0., 0., 0., 0., 0.};   // DiskProperties of the disk host
    R dt, half_dt, t, time_limit;
    double hist_period = std::numeric_limits<double>::infinity();
    double last_hist = -1.;
    int64_t iteration = 0;
    Body<R> p[MAXB];
    Alt<R> alt[MAXB];
    V3<R> err_v[MAXB], err_L[MAXB];
    double roche[MAXB * MAXB];
    Evolver ev[MAXB];
    uint32_t status = 0;
    int timestep_warning = 0;
    int64_t midpoint_iterations = 0;   // diagnostic: effect evaluations executed
};

// ---------------------------------------------------------------------------
// Loading: desc (the case.json content) -> System
// ---------------------------------------------------------------------------
template <class R>
inline bool load_system(const pos_ensemble_desc *d, int s, System<R> &S, std::string *err) {
    const int b0 = d->body_offset[s];
    const int n = d->body_offset[s + 1] - b0;
    if (n < 2 || n > MAXB) { if (err) *err = "body count out of range"; return false; }
    S = System<R>();
    S.n = n;
    const uint32_t eff = d->effects[s];
    S.tides = eff & POS_EFFECT_TIDES;
    S.rf = eff & POS_EFFECT_ROTATIONAL_FLATTENING;
    S.gr = eff & POS_EFFECT_GENERAL_RELATIVITY;
    S.wind = eff & POS_EFFECT_WIND;
    S.evolution = eff & POS_EFFECT_EVOLUTION;
    S.disk = eff & POS_EFFECT_DISK;
    if (S.disk) {
        if (!d->disk_properties) { if (err) *err = "disk effect without disk properties"; return false; }
        for (int q = 0; q < 6; ++q) S.disk_props[q] = d->disk_properties[6 * s + q];
    }
    S.coord = d->coordinates[s];
    S.gr_impl = d->gr_implementation[s];
    if (S.gr && S.gr_impl == POS_GR_DISABLED) S.gr = false;
    S.dt = R(d->time_step[s]);
    S.half_dt = R(0.5) * S.dt;
    S.time_limit = R(d->time_limit[s]);
    S.t = R(d->current_time ? d->current_time[s] : 0.);
    if (d->historic_snapshot_period) S.hist_period = d->historic_snapshot_period[s];
    if (d->last_historic_snapshot_time) S.last_hist = d->last_historic_snapshot_time[s];
    if (d->current_iteration) S.iteration = d->current_iteration[s];

    // most massive body: strictly greater keeps the first maximum (universe.rs:902-909)
    int mm = -1; double mmax = 0.;
    for (int i = 0; i < n; ++i) if (d->mass[b0 + i] > mmax) { mmax = d->mass[b0 + i]; mm = i; }
    if (mm < 0) { if (err) *err = "no positive mass"; return false; }
    S.host = mm;

    for (int i = 0; i < n; ++i) {
        const int b = b0 + i;
        Body<R> &P = S.p[i];
        std::memset((void *)&P, 0, sizeof(P));
        P.m = R(d->mass[b]);
        P.mg = P.m * R(K2);
        P.radius = R(d->radius[b]);
        P.rg2 = R(d->radius_of_gyration_2[b]);
        P.inertia = d->moment_of_inertia ? R(d->moment_of_inertia[b]) : P.m * P.rg2 * pw2(P.radius);
        P.x = V3<R>{R(d->position[3 * b]), R(d->position[3 * b + 1]), R(d->position[3 * b + 2])};
        P.v = V3<R>{R(d->velocity[3 * b]), R(d->velocity[3 * b + 1]), R(d->velocity[3 * b + 2])};
        P.L = V3<R>{R(d->angular_momentum[3 * b]), R(d->angular_momentum[3 * b + 1]), R(d->angular_momentum[3 * b + 2])};
        if (d->spin) P.spin = V3<R>{R(d->spin[3 * b]), R(d->spin[3 * b + 1]), R(d->spin[3 * b + 2])};
        else P.spin = V3<R>{P.L.x / P.inertia, P.L.y / P.inertia, P.L.z / P.inertia};
        P.spin2 = pw2(P.spin.x) + pw2(P.spin.y) + pw2(P.spin.z);
        if (d->heliocentric_velocity)
            P.hv = V3<R>{R(d->heliocentric_velocity[3 * b]), R(d->heliocentric_velocity[3 * b + 1]), R(d->heliocentric_velocity[3 * b + 2])};
        P.roles = d->roles[b];
        P.sigma_scale = R(d->tides_dissipation_factor_scale[b]);
        P.sigma_factor = R(d->tides_dissipation_factor[b]);
        P.sigma = P.sigma_scale * P.sigma_factor;
        if (d->pair_dissipation_factor && !std::isnan(d->pair_dissipation_factor[b])) {
            P.ph_set = true;
            P.ph = R(d->pair_dissipation_factor[b]);
        }
        P.k2 = R(d->tides_love_number[b]);
        P.k2f = R(d->flattening_love_number[b]);
        P.wind_k = R(d->wind_k_factor ? d->wind_k_factor[b] : 0.);
        P.wind_sat = R(d->wind_rotation_saturation ? d->wind_rotation_saturation[b] : 0.);
        P.wind_sat2 = pw2(P.wind_sat);
        if (d->velocity_error) S.err_v[i] = V3<R>{R(d->velocity_error[3 * b]), R(d->velocity_error[3 * b + 1]), R(d->velocity_error[3 * b + 2])};
        else S.err_v[i] = v3zero<R>();
        if (d->angular_momentum_error) S.err_L[i] = V3<R>{R(d->angular_momentum_error[3 * b]), R(d->angular_momentum_error[3 * b + 1]), R(d->angular_momentum_error[3 * b + 2])};
        else S.err_L[i] = v3zero<R>();
        Evolver &E = S.ev[i];
        E = Evolver();
        if (d->evolution_type) {
            E.type = d->evolution_type[b];

            if (E.type != POS_EVO_NON_EVOLVING) {
                const int o = d->evolver_offset[b];
                E.len = d->evolver_offset[b + 1] - o;
                if (E.len <= 0) { if (err) *err = "empty evolver table"; return false; }
                E.time = d->evolver_time + o;
                E.radius = d->evolver_radius + o;
                E.rg2 = d->evolver_radius_of_gyration_2 + o;
                E.love = d->evolver_love_number + o;
                if (E.dynamical()) {
                    if (!d->evolver_inverse_tidal_q_factor) { if (err) *err = "evolving dissipation without a Q-factor table"; return false; }
                    E.itq = d->evolver_inverse_tidal_q_factor + o;
                }
                E.left_index = d->evolver_left_index ? d->evolver_left_index[b] : 0;
            }
        }
    }
    // disable_unnecessary_effects (universe.rs:699-764): an effect no body
    // carries is switched off; find_indices (900-1006): one central body per
    // effect, the GR host is the most massive body, tides and flattening
    // share their host.
    int host_of[4] = {-1, -1, -1, -1};
    const uint32_t central[4] = {POS_ROLE_TIDES_CENTRAL, POS_ROLE_FLAT_CENTRAL, POS_ROLE_GR_CENTRAL, POS_ROLE_DISK_CENTRAL};
    uint32_t any = 0;
    bool evolving = false;
    for (int i = 0; i < n; ++i) {
        any |= S.p[i].roles;
        evolving = evolving || S.ev[i].type != POS_EVO_NON_EVOLVING;
        for (int e = 0; e < 4; ++e)
            if (S.p[i].roles & central[e]) {
                if (host_of[e] >= 0) { if (err) *err = "only one central body is allowed per effect"; return false; }
                host_of[e] = i;
            }
    }
    S.tides = S.tides && host_of[0] >= 0;
    S.rf = S.rf && host_of[1] >= 0;
    S.gr = S.gr && host_of[2] >= 0;
    S.disk = S.disk && host_of[3] >= 0;
    S.wind = S.wind && (any & POS_ROLE_WIND);
    S.evolution = S.evolution && evolving;
    if (S.tides && S.rf && host_of[0] != host_of[1]) { if (err) *err = "tidal and flattening hosts differ"; return false; }
    if (S.gr && host_of[2] != S.host) { if (err) *err = "the GR host is not the most massive body"; return false; }
    // not restated: a tidal, flattening or disk host other than the most
    // massive body (hosts.most_massive.all false, universe.rs:386-419), and
    // Anderson / Newhall with a body whose GR is Disabled (they read every
    // body's GR coordinates, which the reference leaves stale for it)
    if ((S.tides || S.rf) && host_of[S.tides ? 0 : 1] != S.host) { if (err) *err = "tidal host is not the most massive body"; return false; }
    if (S.disk && host_of[3] != S.host) { if (err) *err = "disk host is not the most massive body"; return false; }
    if (S.gr && S.gr_impl != POS_GR_KIDDER1995)
        for (int i = 0; i < n; ++i)
            if (i != S.host && !(S.p[i].roles & POS_ROLE_GR_ORBITING)) {
                if (err) *err = "Anderson1975 / Newhall1983 with a body whose GR is Disabled";
                return false;
            }
    return true;
}

// ---------------------------------------------------------------------------
// Setup: Roche radii (universe.rs:177-196) and initialize_physical_values
// ---------------------------------------------------------------------------
template <class R> inline void calculate_roche_radiuses(System<R> &S) {
    for (int i = 0; i < MAXB * MAXB; ++i) S.roche[i] = 0.;
    for (int i = 0; i < S.n; ++i)
        for (int j = 0; j < S.n; ++j) {
            if (i == j) continue;
            const double ma = val(S.p[i].m), mb = val(S.p[j].m);
            if (ma > mb) S.roche[i * S.n + j] = (val(S.p[j].radius) / 0.462) * std::pow(ma / mb, 1. / 3.);
            else S.roche[i * S.n + j] = (val(S.p[i].radius) / 0.462) * std::pow(mb / ma, 1. / 3.);
        }
}

// evolution.rs:516-546
template <class R> inline void evolve_non_spin_dependent(System<R> &S, const R &t) {
    for (int i = 0; i < S.n; ++i) {
        Body<R> &P = S.p[i];
        Evolver &E = S.ev[i];
        R new_radius = E.eval_radius(t, P.radius);
        R new_rg2 = E.eval_rg2(t, P.rg2);
        if (val(new_radius) != val(P.radius) || val(new_rg2) != val(P.rg2)) {
            P.radius = new_radius;
            P.rg2 = new_rg2;
            P.inertia = P.m * P.rg2 * pw2(P.radius);
        }
        // The love number is evaluated and discarded: the reference binds the
        // CTL parameters by value (evolution.rs:536-538), so only the
        // evolver's left_index moves.
        if (P.roles & (POS_ROLE_TIDES_CENTRAL | POS_ROLE_TIDES_ORBITING)) (void)E.eval_love(t, P.k2);
    }
}

// particles/common.rs:3-14
template <class R> inline void calculate_spin(System<R> &S) {
    for (int i = 0; i < S.n; ++i) {
        Body<R> &P = S.p[i];
        if (val(P.inertia) == 0.) { S.status |= POS_STATUS_ZERO_INERTIA; continue; }
        P.spin.x = P.L.x / P.inertia;
        P.spin.y = P.L.y / P.inertia;
        P.spin.z = P.L.z / P.inertia;
        P.spin2 = pw2(P.spin.x) + pw2(P.spin.y) + pw2(P.spin.z);
    }
}

// evolution.rs:548-567: lag angles of the evolving-dissipation models
template <class R> inline void evolve_spin_dependent(System<R> &S, const R &t) {
    for (int i = 0; i < S.n; ++i) {
        Body<R> &P = S.p[i];
        Evolver &E = S.ev[i];
        if (E.dynamical()) {
            const R inverse_tidal_q_factor = E.eval_inverse_tidal_q_factor(t);
            const R epsilon_squared = P.spin2 / R(SUN_DYN_FREQ_2);
            P.lag = R(3.0) * epsilon_squared * inverse_tidal_q_factor / R(4.0);
        } else {
            P.lag = R(0.);
        }
    }
}

// universe.rs:305-316
template <class R> inline void calculate_spin_and_evolving_quantities(System<R> &S, const R &t, bool evolution) {
    if (evolution && S.evolution) evolve_non_spin_dependent(S, t);
    calculate_spin(S);
    if (evolution && S.evolution) evolve_spin_dependent(S, t);
}

template <class R> inline void initialize_physical_values(System<R> &S) {
    calculate_spin_and_evolving_quantities(S, S.t, true);
    calculate_roche_radiuses(S);
}

// A fresh system (t = 0) runs initialize_physical_values (output.rs:187-193);
// a resumed one keeps the Roche radii its snapshot carries
// (universe.roche_radiuses); entries it lacks (NaN) come from its current radii.
template <class R> inline void start_system(System<R> &S, const pos_ensemble_desc *d, int s) {
    if (val(S.t) == 0.) {
        initialize_physical_values(S);
    } else {
        calculate_roche_radiuses(S);   // the current radii, for entries not carried
        if (d->roche_radius) {
            const int b0 = d->body_offset[s];
            for (int i = 0; i < S.n; ++i)
                for (int j = 0; j < S.n; ++j) {
                    const double r = d->roche_radius[(size_t)(b0 + i) * POS_MAX_BODIES + j];
                    if (i != j && !std::isnan(r)) S.roche[i * S.n + j] = r;
                }
        }
    }
}

// ---------------------------------------------------------------------------
// universe.rs:318-351
// ---------------------------------------------------------------------------
template <class R> inline void inertial_to_heliocentric(System<R> &S) {
    Body<R> &H = S.p[S.host];
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        P.hx.x = P.x.x - H.x.x; P.hx.y = P.x.y - H.x.y; P.hx.z = P.x.z - H.x.z;
        P.hv.x = P.v.x - H.v.x; P.hv.y = P.v.y - H.v.y; P.hv.z = P.v.z - H.v.z;
        P.hdist = sqrt_r(pw2(P.hx.x) + pw2(P.hx.y) + pw2(P.hx.z));
        P.hvr = (P.hx.x * P.hv.x + P.hx.y * P.hv.y + P.hx.z * P.hv.z) / P.hdist;
        P.hnv2 = pw2(P.hv.x - H.hv.x) + pw2(P.hv.y - H.hv.y) + pw2(P.hv.z - H.hv.z);
        P.hnv = sqrt_r(P.hnv2);
    }
    H.hx = v3zero<R>(); H.hv = v3zero<R>();
    H.hdist = R(0.); H.hvr = R(0.); H.hnv2 = R(0.); H.hnv = R(0.);
}

// Gravity terms a WHFast splitting leaves out of gravity_calculate_acceleration
// (universe.rs:66-70): Jacobi drops the host - first body pair, the
// heliocentric splittings every host pair.
enum IgnoreTerms { IGNORE_NONE = 0, IGNORE_WHFAST_ONE = 1, IGNORE_WHFAST_TWO = 2 };

// ---------------------------------------------------------------------------
// Effect kernels
// ---------------------------------------------------------------------------
// tides/common.rs:143-170 (with copy_heliocentric_coordinates, 202-217)
template <class R> inline void tides_initialize(System<R> &S) {
    Body<R> &H = S.p[S.host];
    H.td_x = H.hx; H.td_v = H.hv; H.td_d = H.hdist; H.td_vr = H.hvr;
    H.td_rs_star = R(0.); H.td_rs_planet = R(0.);
    H.td_acc = v3zero<R>(); H.td_dLdt = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_TIDES_ORBITING)) continue;
        P.td_x = P.hx; P.td_v = P.hv; P.td_d = P.hdist; P.td_vr = P.hvr;
        P.td_rs_star = P.td_x.x * H.spin.x + P.td_x.y * H.spin.y + P.td_x.z * H.spin.z;
        P.td_rs_planet = P.td_x.x * P.spin.x + P.td_x.y * P.spin.y + P.td_x.z * P.spin.z;
        P.td_acc = v3zero<R>(); P.td_dLdt = v3zero<R>();
    }
}
// rotational_flattening/common.rs:93-120
template <class R> inline void flattening_initialize(System<R> &S) {
    Body<R> &H = S.p[S.host];
    H.rf_x = H.hx; H.rf_d = H.hdist;
    H.rf_rs_star = R(0.); H.rf_rs_planet = R(0.);
    H.rf_acc = v3zero<R>(); H.rf_dLdt = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_FLAT_ORBITING)) continue;
        P.rf_x = P.hx; P.rf_d = P.hdist;
        P.rf_rs_star = P.rf_x.x * H.spin.x + P.rf_x.y * H.spin.y + P.rf_x.z * H.spin.z;
        P.rf_rs_planet = P.rf_x.x * P.spin.x + P.rf_x.y * P.spin.y + P.rf_x.z * P.spin.z;
        P.rf_acc = v3zero<R>(); P.rf_dLdt = v3zero<R>();
    }
}
// general_relativity.rs:86-109
template <class R> inline void gr_initialize(System<R> &S) {
    if (S.gr_impl == POS_GR_NEWHALL1983 || S.gr_impl == POS_GR_DISABLED) return;
    Body<R> &H = S.p[S.host];
    H.gr_factor = R(0.);
    H.gr_acc = v3zero<R>(); H.gr_dLdt = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_GR_ORBITING)) continue;
        P.gr_factor = H.mg * P.mg / pw2(H.mg + P.mg);
        P.gr_acc = v3zero<R>(); P.gr_dLdt = v3zero<R>();
    }
}

// tools.rs:251-285
template <class R> inline void perihelion_distance_and_eccentricity(const R &gm, const V3<R> &x, const V3<R> &u, R &q, R &e) {
    const R hx = x.y * u.z - x.z * u.y;
    const R hy = x.z * u.x - x.x * u.z;
    const R hz = x.x * u.y - x.y * u.x;
    const R h2 = hx * hx + hy * hy + hz * hz;   // powf(2.) is the exact square
    const R v2 = u.x * u.x + u.y * u.y + u.z * u.z;
    const R r = sqrt_r(x.x * x.x + x.y * x.y + x.z * x.z);
    const R s = h2 / gm;
    const R temp = R(1.) + s * (v2 / gm - R(2.) / r);
    e = val(temp) <= 0. ? R(0.) : sqrt_r(temp);
    q = s / (R(1.) + e);
}

// constant_time_lag.rs:20-194: the dissipation factors of each host-body pair.
// A host with an evolving-dissipation model gets, per orbiting body whose
// mean motion excites its dynamical tide, the pair value (20-94); an
// orbiting body with such a model gets its own pair value when its spin is
// excited by the orbit (96-165).  The map of these values: the host's
// entries are set or removed in every evaluation, and removed again when an
// evolving orbiting body is not excited (the reference removes the host's
// key there, 157-160); the orbiting body's entries are never removed.  A
// missing entry means the body's own scaled factor (188-194).
template <class R>
inline bool excited_pair_factor(const R &gm, const V3<R> &x, const V3<R> &u, const R &norm_spin, const R &radius,
                                const R &lag, const R &scale, const R &factor, R &value) {
    R q, e;
    perihelion_distance_and_eccentricity(gm, x, u, q, e);
    const R mean_motion = sqrt_r(gm) * pow_r(q / (R(1.0) - e), -1.5);
    R half = abs_r(norm_spin - mean_motion);
    if (!(val(half) < val(norm_spin))) return false;
    if (val(half) < SMOOTHING_FACTOR_DYN_TIDE_COROTATION) half = R(SMOOTHING_FACTOR_DYN_TIDE_COROTATION);
    const R inverse_of_half = R(1.) / half;
    value = scale * (R(2.0 * K2) / (R(3.0) * pw5(radius)) * lag * inverse_of_half + factor);
    return true;
}

template <class R> inline void pair_dependent_dissipation_factors(System<R> &S) {
    const Body<R> &H = S.p[S.host];
    const bool host_dyn = S.ev[S.host].dynamical();
    bool hp_set[MAXB];
    R hp[MAXB];
    for (int i = 0; i < S.n; ++i) hp_set[i] = false;
    if (host_dyn) {   // planet-dependent factors of the host (20-94)
        const R host_norm_spin = sqrt_r(H.spin2);
        for (int i = 0; i < S.n; ++i) {
            if (i == S.host) continue;
            const Body<R> &P = S.p[i];
            if (!(P.roles & POS_ROLE_TIDES_ORBITING)) continue;
            hp_set[i] = excited_pair_factor(H.mg + P.mg, P.td_x, P.td_v, host_norm_spin, H.radius, H.lag, H.sigma_scale,
                                            H.sigma_factor, hp[i]);
        }
    }
    for (int i = 0; i < S.n; ++i) {   // host-dependent factors of the orbiting bodies (96-165)
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!S.ev[i].dynamical() || !(P.roles & POS_ROLE_TIDES_ORBITING)) continue;
        R v;
        if (excited_pair_factor(H.mg + P.mg, P.td_x, P.td_v, sqrt_r(P.spin2), P.radius, P.lag, P.sigma_scale,
                                P.sigma_factor, v)) {
            P.ph = v;
            P.ph_set = true;
        } else {
            hp_set[i] = false;
        }
    }
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        P.sigma_host = (host_dyn && hp_set[i]) ? hp[i] : H.sigma;
        P.sigma_pair = (S.ev[i].dynamical() && P.ph_set) ? P.ph : P.sigma;
    }
}

// constant_time_lag.rs:206-264
template <class R> inline void tides_orthogonal(System<R> &S) {
    Body<R> &H = S.p[S.host];
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_TIDES_ORBITING)) continue;
        R d7 = pw7(P.td_d);
        P.td_orth_star = R(4.5) * pw2(P.m) * pw10(H.radius) * P.sigma_host / d7;
    }
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_TIDES_ORBITING)) continue;
        R d7 = pw7(P.td_d);
        P.td_orth_planet = R(4.5) * pw2(H.m) * pw10(P.radius) * P.sigma_pair / d7;
    }
}
// oblate_spheroid.rs:8-49
template <class R> inline void flattening_orthogonal(System<R> &S) {
    Body<R> &H = S.p[S.host];
    const R host_k2f = H.k2f;   // host is a flattening central body (checked at load)
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_FLAT_ORBITING)) continue;
        P.rf_fac_star = P.m * host_k2f * H.spin2 * pw5(H.radius) / R(6.);
        P.rf_orth_star = R(-6.) * P.rf_fac_star * P.rf_rs_star / (H.spin2 * pw5(P.rf_d));
    }
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_FLAT_ORBITING)) continue;
        P.rf_fac_planet = H.m * P.k2f * P.spin2 * pw5(P.radius) / R(6.);
        P.rf_orth_planet = R(-6.) * P.rf_fac_planet * P.rf_rs_planet / (P.spin2 * pw5(P.rf_d));
    }
}
// constant_time_lag.rs:266-307
template <class R> inline void tides_radial(System<R> &S) {
    Body<R> &H = S.p[S.host];
    const R host_mass_2 = H.m * H.m;
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_TIDES_ORBITING)) continue;
        const R pm2 = P.m * P.m;
        R conservative = R(-3.0 * K2) / pw7(P.td_d) * (pm2 * pw5(H.radius) * H.k2 + host_mass_2 * pw5(P.radius) * P.k2);
        R factor1 = R(-13.5) * P.td_vr / pw8(P.td_d);
        R term1 = pm2 * pw10(H.radius) * P.sigma_host;
        R term2 = host_mass_2 * pw10(P.radius) * P.sigma_pair;
        P.td_diss_pm = factor1 * term2;
        R dissipative = P.td_diss_pm + factor1 * term1;
        P.td_radial = conservative + dissipative;
    }
}
// constant_time_lag.rs:309-332 and tides/common.rs:312-345
template <class R> inline void tides_acceleration(System<R> &S) {
    Body<R> &H = S.p[S.host];
    const R factor2 = R(1.) / H.m;
    V3<R> sum = v3zero<R>();
    const V3<R> hs = H.spin;
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_TIDES_ORBITING)) continue;
        const R d = P.td_d;
        const V3<R> &x = P.td_x, &u = P.td_v;
        const R factor3 = P.td_radial + (P.td_orth_star + P.td_orth_planet) * P.td_vr / d;
        V3<R> F;
        F.x = factor3 * x.x / d
            + P.td_orth_star / d * (hs.y * x.z - hs.z * x.y - u.x)
            + P.td_orth_planet / d * (P.spin.y * x.z - P.spin.z * x.y - u.x);
        F.y = factor3 * x.y / d
            + P.td_orth_star / d * (hs.z * x.x - hs.x * x.z - u.y)
            + P.td_orth_planet / d * (P.spin.z * x.x - P.spin.x * x.z - u.y);
        F.z = factor3 * x.z / d
            + P.td_orth_star / d * (hs.x * x.y - hs.y * x.x - u.z)
            + P.td_orth_planet / d * (P.spin.x * x.y - P.spin.y * x.x - u.z);
        const R factor1 = R(1.) / P.m;
        sum.x = sum.x + F.x; sum.y = sum.y + F.y; sum.z = sum.z + F.z;
        P.td_acc.x = factor1 * F.x; P.td_acc.y = factor1 * F.y; P.td_acc.z = factor1 * F.z;
    }
    H.td_acc.x = R(-1.0) * factor2 * sum.x;
    H.td_acc.y = R(-1.0) * factor2 * sum.y;
    H.td_acc.z = R(-1.0) * factor2 * sum.z;
}
// oblate_spheroid.rs:51-60
template <class R> inline void flattening_radial(System<R> &S) {
    Body<R> &H = S.p[S.host];
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_FLAT_ORBITING)) continue;
        P.rf_radial = R(-3.) / pw5(P.rf_d) * (P.rf_fac_planet + P.rf_fac_star)
            + R(15.) / pw7(P.rf_d) * (P.rf_fac_star * P.rf_rs_star * P.rf_rs_star / H.spin2
                                      + P.rf_fac_planet * P.rf_rs_planet * P.rf_rs_planet / P.spin2);
    }
}
// oblate_spheroid.rs:83-97 and rotational_flattening/common.rs:203-237
template <class R> inline void flattening_acceleration(System<R> &S) {
    Body<R> &H = S.p[S.host];
    const R factor2 = R(1.) / H.m;
    V3<R> sum = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_FLAT_ORBITING)) continue;
        V3<R> F;
        F.x = P.rf_radial * P.rf_x.x + P.rf_orth_planet * P.spin.x + P.rf_orth_star * H.spin.x;
        F.y = P.rf_radial * P.rf_x.y + P.rf_orth_planet * P.spin.y + P.rf_orth_star * H.spin.y;
        F.z = P.rf_radial * P.rf_x.z + P.rf_orth_planet * P.spin.z + P.rf_orth_star * H.spin.z;
        sum.x = sum.x + F.x; sum.y = sum.y + F.y; sum.z = sum.z + F.z;
        const R factor1 = R(1.) / P.m;
        P.rf_acc.x = factor1 * F.x; P.rf_acc.y = factor1 * F.y; P.rf_acc.z = factor1 * F.z;
    }
    H.rf_acc.x = R(-1.0) * factor2 * sum.x;
    H.rf_acc.y = R(-1.0) * factor2 * sum.y;
    H.rf_acc.z = R(-1.0) * factor2 * sum.z;
}
// constant_time_lag.rs:171-204 and tides/common.rs:223-261
template <class R> inline V3<R> ctl_torque(const Body<R> &P, const V3<R> &ref_spin, const R &rs, const R &orth) {
    const R d = P.td_d;
    const V3<R> &x = P.td_x, &u = P.td_v;
    V3<R> N;
    N.x = orth * (d * ref_spin.x - rs * x.x / d - R(1.0) / d * (x.y * u.z - x.z * u.y));
    N.y = orth * (d * ref_spin.y - rs * x.y / d - R(1.0) / d * (x.z * u.x - x.x * u.z));
    N.z = orth * (d * ref_spin.z - rs * x.z / d - R(1.0) / d * (x.x * u.y - x.y * u.x));
    return N;
}
template <class R> inline void tides_torques(System<R> &S) {
    Body<R> &H = S.p[S.host];
    const R factor = R(-1.0);
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_TIDES_ORBITING)) continue;
        V3<R> N = ctl_torque(P, P.spin, P.td_rs_planet, P.td_orth_planet);
        P.td_dLdt.x = factor * N.x; P.td_dLdt.y = factor * N.y; P.td_dLdt.z = factor * N.z;
    }
    // the host's torque runs over every other body: the test of
    // tides/common.rs:244-255 is on the host's effect, not the body's
    V3<R> sum = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        V3<R> N = ctl_torque(P, H.spin, P.td_rs_star, P.td_orth_star);
        sum.x = sum.x + factor * N.x; sum.y = sum.y + factor * N.y; sum.z = sum.z + factor * N.z;
    }
    H.td_dLdt = sum;
}
// oblate_spheroid.rs:62-81 and rotational_flattening/common.rs:165-201
template <class R> inline V3<R> flat_torque(const Body<R> &P, const V3<R> &s, const R &orth) {
    const V3<R> &x = P.rf_x;
    V3<R> N;
    N.x = orth * (x.y * s.z - x.z * s.y);
    N.y = orth * (x.z * s.x - x.x * s.z);
    N.z = orth * (x.x * s.y - x.y * s.x);
    return N;
}
template <class R> inline void flattening_torques(System<R> &S) {
    Body<R> &H = S.p[S.host];
    const R factor = R(-1.0);
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_FLAT_ORBITING)) continue;
        V3<R> N = flat_torque(P, P.spin, P.rf_orth_planet);
        P.rf_dLdt.x = factor * N.x; P.rf_dLdt.y = factor * N.y; P.rf_dLdt.z = factor * N.z;
    }
    V3<R> sum = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        V3<R> N = flat_torque(P, H.spin, P.rf_orth_star);
        sum.x = sum.x + factor * N.x; sum.y = sum.y + factor * N.y; sum.z = sum.z + factor * N.z;
    }
    H.rf_dLdt = sum;
}

// general_relativity.rs:187-239
template <class R> inline void kidder_first_order(System<R> &S) {
    Body<R> &H = S.p[S.host];
    V3<R> sum = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_GR_ORBITING)) continue;
        const R mg = H.mg + P.mg;
        const R d = P.hdist, f = P.gr_factor;
        const R distance_2 = pw2(d);
        const R radial_velocity_2 = pw2(P.hvr);
        const R radial = -mg / (distance_2 * R(SPEED_OF_LIGHT_2))
            * ((R(1.0) + R(3.0) * f) * P.hnv2
               - R(2.0) * (R(2.0) + f) * mg / d
               - R(1.5) * f * radial_velocity_2);
        const R orthogonal = mg / (distance_2 * R(SPEED_OF_LIGHT_2)) * R(2.0) * (R(2.0) - f) * P.hvr * P.hnv;
        V3<R> a;
        a.x = radial * P.hx.x / d + orthogonal * P.hv.x / P.hnv;
        a.y = radial * P.hx.y / d + orthogonal * P.hv.y / P.hnv;
        a.z = radial * P.hx.z / d + orthogonal * P.hv.z / P.hnv;
        sum.x = sum.x + P.m / H.m * a.x;
        sum.y = sum.y + P.m / H.m * a.y;
        sum.z = sum.z + P.m / H.m * a.z;
        P.gr_acc = a;
    }
    H.gr_acc.x = R(-1.0) * sum.x; H.gr_acc.y = R(-1.0) * sum.y; H.gr_acc.z = R(-1.0) * sum.z;
}
// general_relativity.rs:241-298
template <class R> inline void kidder_second_order(System<R> &S) {
    Body<R> &H = S.p[S.host];
    V3<R> sum = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_GR_ORBITING)) continue;
        const R mg = H.mg + P.mg;
        const R d = P.hdist, f = P.gr_factor;
        const R distance_2 = pw2(d);
        const R nv2 = P.hnv2;
        const R nv4 = pw2(nv2);
        const R vr2 = pw2(P.hvr);
        const R vr4 = pw2(vr2);
        const R f2 = pw2(f);
        const R radial = -mg / (distance_2 * R(SPEED_OF_LIGHT_2))
            * (R(3.0 / 4.0) * (R(12.0) + R(29.0) * f) * (pw2(mg) / distance_2)
               + f * (R(3.0) - R(4.0) * f) * nv4
               + R(15.0 / 8.0) * f * (R(1.0) - R(3.0) * f) * vr4
               - R(3.0 / 2.0) * f * (R(3.0) - R(4.0) * f) * vr2 * nv2
               - R(0.5) * f * (R(13.0) - R(4.0) * f) * (mg / d) * nv2
               - (R(2.0) + R(25.0) * f + R(2.0) * f2) * (mg / d) * vr2);
        const R orthogonal = -mg / (distance_2 * R(SPEED_OF_LIGHT_2))
            * R(-0.5) * P.hvr
            * (f * (R(15.0) + R(4.0) * f) * nv2
               - (R(4.0) + R(41.0) * f + R(8.0) * f2) * (mg / d)
               - R(3.0) * f * (R(3.0) + R(2.0) * f) * vr2);
        V3<R> a;
        a.x = radial * P.hx.x / d + orthogonal * P.hv.x;
        a.y = radial * P.hx.y / d + orthogonal * P.hv.y;
        a.z = radial * P.hx.z / d + orthogonal * P.hv.z;
        sum.x = sum.x + P.m / H.m * a.x;
        sum.y = sum.y + P.m / H.m * a.y;
        sum.z = sum.z + P.m / H.m * a.z;
        P.gr_acc.x = P.gr_acc.x + a.x; P.gr_acc.y = P.gr_acc.y + a.y; P.gr_acc.z = P.gr_acc.z + a.z;
    }
    H.gr_acc.x = H.gr_acc.x + R(-1.0) * sum.x;
    H.gr_acc.y = H.gr_acc.y + R(-1.0) * sum.y;
    H.gr_acc.z = H.gr_acc.z + R(-1.0) * sum.z;
}
// general_relativity.rs:300-456
template <class R> inline void kidder_spin_orbit(System<R> &S) {
    Body<R> &H = S.p[S.host];
    const V3<R> Ls{H.inertia * H.spin.x, H.inertia * H.spin.y, H.inertia * H.spin.z};
    V3<R> sum = v3zero<R>();
    H.gr_dLdt = v3zero<R>();
    const R factor_a = R(G) / R(SPEED_OF_LIGHT_2);
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_GR_ORBITING)) continue;
        const R star_planet_mass = H.m + P.m;
        const R star_planet_diff_mass = H.m - P.m;
        const R mass_factor = star_planet_diff_mass / star_planet_mass;
        const V3<R> Lp{P.inertia * P.spin.x, P.inertia * P.spin.y, P.inertia * P.spin.z};
        const R d = P.hdist;
        const V3<R> n{P.hx.x / d, P.hx.y / d, P.hx.z / d};
        const V3<R> &v = P.hv;
        V3<R> msf;
        msf.x = mass_factor * star_planet_mass * (Lp.x / P.m - Ls.x / H.m);
        msf.y = mass_factor * star_planet_mass * (Lp.y / P.m - Ls.y / H.m);
        msf.z = mass_factor * star_planet_mass * (Lp.z / P.m - Ls.z / H.m);
        V3<R> e1;
        e1.x = R(6.) * n.x * ((n.y * v.z - n.z * v.y) * (R(2.) * (Ls.x + Lp.x) + msf.x));
        e1.y = R(6.) * n.y * ((n.z * v.x - n.x * v.z) * (R(2.) * (Ls.y + Lp.y) + msf.y));
        e1.z = R(6.) * n.z * ((n.x * v.y - n.y * v.x) * (R(2.) * (Ls.z + Lp.z) + msf.z));
        V3<R> e7{R(7.) * (Ls.x + Lp.x) + R(3.) * msf.x,
                 R(7.) * (Ls.y + Lp.y) + R(3.) * msf.y,
                 R(7.) * (Ls.z + Lp.z) + R(3.) * msf.z};
        V3<R> e2{v.y * e7.z - v.z * e7.y, v.z * e7.x - v.x * e7.z, v.x * e7.y - v.y * e7.x};
        V3<R> e3s{R(3.) * (Ls.x + Lp.x) + msf.x, R(3.) * (Ls.y + Lp.y) + msf.y, R(3.) * (Ls.z + Lp.z) + msf.z};
        V3<R> e3;
        e3.x = R(3.) * P.hvr * (n.y * e3s.z - n.z * e3s.y);
        e3.y = R(3.) * P.hvr * (n.z * e3s.x - n.x * e3s.z);
        e3.z = R(3.) * P.hvr * (n.x * e3s.y - n.y * e3s.x);
        V3<R> a{factor_a * (e1.x - e2.x + e3.x), factor_a * (e1.y - e2.y + e3.y), factor_a * (e1.z - e2.z + e3.z)};
        sum.x = sum.x + P.m / H.m * a.x;
        sum.y = sum.y + P.m / H.m * a.y;
        sum.z = sum.z + P.m / H.m * a.z;
        P.gr_acc.x = P.gr_acc.x + a.x; P.gr_acc.y = P.gr_acc.y + a.y; P.gr_acc.z = P.gr_acc.z + a.z;

        // Kidder 1995 eq. 2.4a (host) and 2.4b (planet)
        const R mu = (H.m * P.m) / star_planet_mass;
        const V3<R> Lo{mu * (P.hx.y * v.z - P.hx.z * v.y), mu * (P.hx.z * v.x - P.hx.x * v.z), mu * (P.hx.x * v.y - P.hx.y * v.x)};
        R fm = R(2.) + R(3. / 2.) * P.m / H.m;
        V3<R> E1{fm * (Lo.y * Ls.z - Lo.z * Ls.y), fm * (Lo.z * Ls.x - Lo.x * Ls.z), fm * (Lo.x * Ls.y - Lo.y * Ls.x)};
        V3<R> E2{Lp.y * Ls.z - Lp.z * Ls.y, Lp.z * Ls.x - Lp.x * Ls.z, Lp.x * Ls.y - Lp.y * Ls.x};
        R sp = n.x * Lp.x + n.y * Lp.y + n.z * Lp.z;
        V3<R> E3{R(3.) * sp * (n.y * Ls.z - n.z * Ls.y), R(3.) * sp * (n.z * Ls.x - n.x * Ls.z), R(3.) * sp * (n.x * Ls.y - n.y * Ls.x)};
        H.gr_dLdt.x = H.gr_dLdt.x + factor_a * (E1.x - E2.x + E3.x);
        H.gr_dLdt.y = H.gr_dLdt.y + factor_a * (E1.y - E2.y + E3.y);
        H.gr_dLdt.z = H.gr_dLdt.z + factor_a * (E1.z - E2.z + E3.z);
        fm = R(2.) + R(3. / 2.) * H.m / P.m;
        V3<R> F1{fm * (Lo.y * Lp.z - Lo.z * Lp.y), fm * (Lo.z * Lp.x - Lo.x * Lp.z), fm * (Lo.x * Lp.y - Lo.y * Lp.x)};
        V3<R> F2{Ls.y * Lp.z - Ls.z * Lp.y, Ls.z * Lp.x - Ls.x * Lp.z, Ls.x * Lp.y - Ls.y * Lp.x};
        R sh = n.x * Ls.x + n.y * Ls.y + n.z * Ls.z;
        V3<R> F3{R(3.) * sh * (n.y * Lp.z - n.z * Lp.y), R(3.) * sh * (n.z * Lp.x - n.x * Lp.z), R(3.) * sh * (n.x * Lp.y - n.y * Lp.x)};
        P.gr_dLdt.x = factor_a * (F1.x - F2.x + F3.x);
        P.gr_dLdt.y = factor_a * (F1.y - F2.y + F3.y);
        P.gr_dLdt.z = factor_a * (F1.z - F2.z + F3.z);
    }
    H.gr_acc.x = H.gr_acc.x + R(-1.0) * sum.x;
    H.gr_acc.y = H.gr_acc.y + R(-1.0) * sum.y;
    H.gr_acc.z = H.gr_acc.z + R(-1.0) * sum.z;
}

// wind.rs:61-91
// general_relativity.rs:641-681: the gravitational accelerations of the last
// gravity evaluation plus the host pair terms the splitting ignored.  The
// pair separation mixes the host's inertial position with the body's
// heliocentric one, as the reference does.  Order: host, then the other
// bodies in index order.
template <class R> inline void gr_newtonian_accelerations(const System<R> &S, V3<R> acc[]) {
    const Body<R> &H = S.p[S.host];
    acc[0] = H.a_grav;
    int c = 1;
    for (int i = 0; i < S.n; ++i)
        if (i != S.host) acc[c++] = S.p[i].a_grav;
    const int ignore = S.coord == POS_COORD_JACOBI ? IGNORE_WHFAST_ONE : IGNORE_WHFAST_TWO;
    const int np = ignore == IGNORE_WHFAST_ONE ? 1 : S.n - 1;
    c = 1;
    for (int i = 0; i < S.n && c <= np; ++i) {
        if (i == S.host) continue;
        const Body<R> &P = S.p[i];
        const R dx = H.x.x - P.hx.x, dy = H.x.y - P.hx.y, dz = H.x.z - P.hx.z;
        const R r2 = pw2(dx) + pw2(dy) + pw2(dz);
        const R r = sqrt_r(r2);
        const R prefac = R(G) / (r2 * r);
        const R pms = prefac * H.m, pmp = prefac * P.m;
        acc[0].x = acc[0].x - pmp * dx; acc[0].y = acc[0].y - pmp * dy; acc[0].z = acc[0].z - pmp * dz;
        acc[c].x = acc[c].x + pms * dx; acc[c].y = acc[c].y + pms * dy; acc[c].z = acc[c].z + pms * dz;
        ++c;
    }
}

// Anderson et al. 1975 via Jacobi coordinates (general_relativity.rs:461-639)
template <class R> inline void anderson1975(System<R> &S) {
    Body<R> &H = S.p[S.host];
    V3<R> an[MAXB];
    gr_newtonian_accelerations(S, an);
    V3<R> jx[MAXB], jv[MAXB], ja[MAXB];
    int idx[MAXB], np = 0;
    R eta = H.m;
    V3<R> s{eta * H.x.x, eta * H.x.y, eta * H.x.z};
    V3<R> sv{eta * H.v.x, eta * H.v.y, eta * H.v.z};
    V3<R> sa{eta * an[0].x, eta * an[0].y, eta * an[0].z};
    int c = 0;
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        ++c;
        const Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_GR_ORBITING)) continue;
        const R ei = R(1.) / eta;
        eta = eta + P.m;
        const R pme = eta * ei;
        V3<R> x{P.x.x - s.x * ei, P.x.y - s.y * ei, P.x.z - s.z * ei};
        V3<R> v{P.v.x - sv.x * ei, P.v.y - sv.y * ei, P.v.z - sv.z * ei};
        V3<R> a{an[c].x - sa.x * ei, an[c].y - sa.y * ei, an[c].z - sa.z * ei};
        s.x = s.x * pme + P.m * x.x; s.y = s.y * pme + P.m * x.y; s.z = s.z * pme + P.m * x.z;
        sv.x = sv.x * pme + P.m * v.x; sv.y = sv.y * pme + P.m * v.y; sv.z = sv.z * pme + P.m * v.z;
        sa.x = sa.x * pme + P.m * a.x; sa.y = sa.y * pme + P.m * a.y; sa.z = sa.z * pme + P.m * a.z;
        jx[np] = x; jv[np] = v; ja[np] = a; idx[np] = i; ++np;
    }
    const R mtot = eta;
    const R mu = H.mg;
    const R c2 = R(SPEED_OF_LIGHT_2);
    for (int q = 0; q < np; ++q) {
        V3<R> vi = jv[q];
        R vi2 = pw2(jv[q].x) + pw2(jv[q].y) + pw2(jv[q].z);
        const R ri = sqrt_r(pw2(jx[q].x) + pw2(jx[q].y) + pw2(jx[q].z));
        R fa = (R(0.5) * vi2 + R(3.) * mu / ri) / c2;
        for (int it = 0; it < 10; ++it) {
            const V3<R> old = vi;
            vi.x = jv[q].x / (R(1.) - fa); vi.y = jv[q].y / (R(1.) - fa); vi.z = jv[q].z / (R(1.) - fa);
            vi2 = vi.x * vi.x + vi.y * vi.y + vi.z * vi.z;
            fa = (R(0.5) * vi2 + R(3.) * mu / ri) / c2;
            const R dvx = vi.x - old.x, dvy = vi.y - old.y, dvz = vi.z - old.z;
            if (val((dvx * dvx + dvy * dvy + dvz * dvz) / vi2) < DBL_EPS_2) break;
        }
        const R ri3 = ri * ri * ri;
        const R fb = (mu / ri - R(1.5) * vi2) * mu / ri3 / c2;
        const R rdr = jx[q].x * jv[q].x + jx[q].y * jv[q].y + jx[q].z * jv[q].z;
        const V3<R> vd{ja[q].x + fb * jx[q].x, ja[q].y + fb * jx[q].y, ja[q].z + fb * jx[q].z};
        const R vdv = vi.x * vd.x + vi.y * vd.y + vi.z * vd.z;
        const R fd = (vdv - R(3.) * mu / ri3 * rdr) / c2;
        ja[q].x = fb * (R(1.) - fa) * jx[q].x - fa * ja[q].x - fd * vi.x;
        ja[q].y = fb * (R(1.) - fa) * jx[q].y - fa * ja[q].y - fd * vi.y;
        ja[q].z = fb * (R(1.) - fa) * jx[q].z - fa * ja[q].z - fd * vi.z;
    }
    // back to inertial accelerations, the Jacobi star acceleration being zero
    eta = mtot;
    sa = V3<R>{eta * R(0.), eta * R(0.), eta * R(0.)};
    for (int q = np - 1; q >= 0; --q) {
        Body<R> &P = S.p[idx[q]];
        const R ei = R(1.) / eta;
        sa.x = (sa.x - P.m * ja[q].x) * ei; sa.y = (sa.y - P.m * ja[q].y) * ei; sa.z = (sa.z - P.m * ja[q].z) * ei;
        P.gr_acc = V3<R>{ja[q].x + sa.x, ja[q].y + sa.y, ja[q].z + sa.z};
        eta = eta - P.m;
        sa.x = sa.x * eta; sa.y = sa.y * eta; sa.z = sa.z * eta;
    }
    const R mi = R(1.) / eta;
    H.gr_acc = V3<R>{sa.x * mi, sa.y * mi, sa.z * mi};
}

// Newhall et al. 1983, all pairs (general_relativity.rs:683-895).  Bodies
// in the order host, then the others by index.
template <class R> inline void newhall1983(System<R> &S) {
    const int n = S.n;
    int ord[MAXB];
    ord[0] = S.host;
    for (int i = 0, c = 1; i < n; ++i)
        if (i != S.host) ord[c++] = i;
    V3<R> an[MAXB];
    gr_newtonian_accelerations(S, an);
    V3<R> drs[MAXB][MAXB];
    R rs[MAXB][MAXB];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            drs[i][j] = v3zero<R>();
            rs[i][j] = R(0.);
        }
    auto gr_on = [&](int i) { return (S.p[ord[i]].roles & (POS_ROLE_GR_CENTRAL | POS_ROLE_GR_ORBITING)) != 0; };
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            if (j == i || !(gr_on(i) || gr_on(j))) continue;
            const Body<R> &Pi = S.p[ord[i]], &Pj = S.p[ord[j]];
            drs[i][j] = V3<R>{Pi.x.x - Pj.x.x, Pi.x.y - Pj.x.y, Pi.x.z - Pj.x.z};
            rs[i][j] = sqrt_r(pw2(drs[i][j].x) + pw2(drs[i][j].y) + pw2(drs[i][j].z));
        }
    const R c2 = R(SPEED_OF_LIGHT_2);
    V3<R> ac[MAXB];
    for (int i = 0; i < n; ++i) {
        const Body<R> &Pi = S.p[ord[i]];
        V3<R> a = v3zero<R>();
        for (int j = 0; j < n; ++j) {
            if (j == i || !(gr_on(i) || gr_on(j))) continue;
            const Body<R> &Pj = S.p[ord[j]];
            const R dx = drs[i][j].x, dy = drs[i][j].y, dz = drs[i][j].z;
            const R rij2 = pw2(rs[i][j]);
            const R rij3 = rij2 * rs[i][j];
            R a1 = R(0.);
            for (int k = 0; k < n; ++k)
                if (k != i) a1 = a1 + (R(4.) / c2) * R(G) * S.p[ord[k]].m / rs[i][k];
            R a2 = R(0.);
            for (int l = 0; l < n; ++l)
                if (l != j) a2 = a2 + (R(1.) / c2) * R(G) * S.p[ord[l]].m / rs[l][j];
            const R vi2 = pw2(Pi.v.x) + pw2(Pi.v.y) + pw2(Pi.v.z);
            const R a3 = -vi2 / c2;
            const R vj2 = pw2(Pj.v.x) + pw2(Pj.v.y) + pw2(Pj.v.z);
            const R a4 = R(-2.) * vj2 / c2;
            const R a5 = (R(4.) / c2) * (Pi.v.x * Pj.v.x + Pi.v.y * Pj.v.y + Pi.v.z * Pj.v.z);
            const R a60 = dx * Pj.v.x + dy * Pj.v.y + dz * Pj.v.z;
            const R a6 = (R(3.) / (R(2.) * c2)) * pw2(a60) / rij2;
            const R f1 = a1 + a2 + a3 + a4 + a5 + a6;
            const R gm = R(G) * Pj.m;
            a.x = a.x + gm * dx * f1 / rij3;
            a.y = a.y + gm * dy * f1 / rij3;
            a.z = a.z + gm * dz * f1 / rij3;
            const R dvx = Pi.v.x - Pj.v.x, dvy = Pi.v.y - Pj.v.y, dvz = Pi.v.z - Pj.v.z;
            const R f2 = dx * (R(4.) * Pi.v.x - R(3.) * Pj.v.x) + dy * (R(4.) * Pi.v.y - R(3.) * Pj.v.y)
                       + dz * (R(4.) * Pi.v.z - R(3.) * Pj.v.z);
            a.x = a.x + gm * f2 * dvx / rij3 / c2;
            a.y = a.y + gm * f2 * dvy / rij3 / c2;
            a.z = a.z + gm * f2 * dvz / rij3 / c2;
        }
        ac[i] = a;
    }
    V3<R> anew[MAXB], aold[MAXB];
    for (int i = 0; i < n; ++i) anew[i] = v3zero<R>();
    const double dev_limit = 1.0e-30;
    for (int k = 0; k < 10; ++k) {
        for (int i = 0; i < n; ++i) aold[i] = anew[i];
        for (int i = 0; i < n; ++i) {
            V3<R> nc = v3zero<R>();
            for (int j = 0; j < n; ++j) {
                if (j == i || !(gr_on(i) || gr_on(j))) continue;
                const Body<R> &Pj = S.p[ord[j]];
                const R dx = drs[i][j].x, dy = drs[i][j].y, dz = drs[i][j].z;
                const R rij = rs[i][j];
                const R rij2 = pw2(rij);
                const R rij3 = rij2 * rij;
                const V3<R> s{an[j].x + aold[j].x, an[j].y + aold[j].y, an[j].z + aold[j].z};
                const R proj = dx * s.x + dy * s.y + dz * s.z;
                const R gm = R(G) * Pj.m;
                nc.x = nc.x + (gm * dx / rij3) * proj / (R(2.) * c2) + (R(7.) / (R(2.) * c2)) * R(G) * Pj.m * s.x / rij;
                nc.y = nc.y + (gm * dy / rij3) * proj / (R(2.) * c2) + (R(7.) / (R(2.) * c2)) * R(G) * Pj.m * s.y / rij;
                nc.z = nc.z + (gm * dz / rij3) * proj / (R(2.) * c2) + (R(7.) / (R(2.) * c2)) * R(G) * Pj.m * s.z / rij;
            }
            anew[i] = V3<R>{ac[i].x + nc.x, ac[i].y + nc.y, ac[i].z + nc.z};
        }
        double maxdev = 0., dx = 0., dy = 0., dz = 0.;
        for (int i = 0; i < n; ++i) {
            if (!gr_on(i)) continue;
            if (std::fabs(val(anew[i].x)) < dev_limit) dx = std::fabs(val(anew[i].x) - val(aold[i].x)) / val(anew[i].x);
            if (std::fabs(val(anew[i].y)) < dev_limit) dy = std::fabs(val(anew[i].y) - val(aold[i].y)) / val(anew[i].y);
            if (std::fabs(val(anew[i].z)) < dev_limit) dz = std::fabs(val(anew[i].z) - val(aold[i].z)) / val(anew[i].z);
            if (dx > maxdev) maxdev = dx;
            if (dy > maxdev) maxdev = dy;
            if (dz > maxdev) maxdev = dz;
        }
        if (maxdev < dev_limit) break;
    }
    for (int i = 1; i < n; ++i) {
        Body<R> &P = S.p[ord[i]];
        if (P.roles & POS_ROLE_GR_ORBITING) P.gr_acc = anew[i];
    }
    S.p[S.host].gr_acc = anew[0];
}

template <class R> inline void wind_factor(System<R> &S) {
    for (int i = 0; i < S.n; ++i) {
        Body<R> &P = S.p[i];
        if (!(P.roles & POS_ROLE_WIND)) continue;
        const R threshold = sqrt_r(P.spin2);
        const R shape = sqrt_r(P.radius / R(R_SUN) * R(1.) / P.m);
        if (val(threshold) >= val(P.wind_sat)) {
            P.wind_dLdt.x = R(-1.) * P.wind_k * P.spin.x * P.wind_sat2 * shape;
            P.wind_dLdt.y = R(-1.) * P.wind_k * P.spin.y * P.wind_sat2 * shape;
            P.wind_dLdt.z = R(-1.) * P.wind_k * P.spin.z * P.wind_sat2 * shape;
        } else {
            P.wind_dLdt.x = R(-1.) * P.wind_k * P.spin.x * P.spin2 * shape;
            P.wind_dLdt.y = R(-1.) * P.wind_k * P.spin.y * P.spin2 * shape;
            P.wind_dLdt.z = R(-1.) * P.wind_k * P.spin.z * P.spin2 * shape;
        }
    }
}

// disk.rs:155-194: the host's DiskProperties methods
template <class R>
inline R disk_migration_timescale(const double *dp, const R &time, const R &host_mass, const R &host_mass_g,
                                  const R &mass, const R &distance, const R &sma) {
    const double inner = dp[0], outer = dp[1], lifetime = dp[2], alpha = dp[3], sdn = dp[4], mmw = dp[5];
    // gas radial velocity (disk.rs:222-226) from the viscosity (212-220)
    const R keplerian_frequency = sqrt_r(host_mass_g / pw3(distance));
    const R temperature = R(280.0) * pow_r(distance, -0.5) * host_mass;
    const R speed_of_sound_2 = R(BOLTZMANN_CONSTANT) * temperature / (R(mmw) * R(MASS_HYDROGEN_ATOM));
    const R viscosity = R(alpha) * speed_of_sound_2 / keplerian_frequency;
    const R gas_radial_velocity = R(-3.0) * viscosity / (R(2.0) * sma);
    // surface density (disk.rs:196-205); powi(-1) is 1/x
    const R initial = R(sdn) * (R(1.) / distance) * exp_r(R(-1.0) * distance / R(outer)) *
                      (R(1.0) - sqrt_r(R(inner) / distance));
    R surface_density = R(0.0);
    if (val(distance) > inner) surface_density = initial * exp_r(-time / R(lifetime));
    const R ratio = R(2.0) * surface_density * pw2(sma) / mass;
    const R a_dot = gas_radial_velocity * (std::isnan(val(ratio)) || val(ratio) > 1.0 ? R(1.0) : ratio);
    return -sma / a_dot;
}

// disk.rs:86-99 (initialize) and 228-268 (Alibert et al. 2013 forces)
template <class R> inline void disk_acceleration(System<R> &S, const R &t) {
    Body<R> &H = S.p[S.host];
    const R factor2 = R(1.) / H.m;
    V3<R> sum = v3zero<R>();
    for (int i = 0; i < S.n; ++i) S.p[i].disk_acc = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        if (i == S.host) continue;
        Body<R> &P = S.p[i];
        if (P.roles & POS_ROLE_DISK_ORBITING) {
            const R d = P.hdist;
            const R sma = (P.mg + H.mg) * d / (R(2.0) * (P.mg + H.mg) - d * P.hnv2);
            P.disk_migration = disk_migration_timescale(S.disk_props, t, H.m, H.mg, P.m, d, sma);
            const R factor1 = R(1.) / P.m;
            const R factor_migration = -P.m / P.disk_migration;
            const R ecc_timescale = R(0.1) * P.disk_migration;
            const R factor_ecc = -P.m * R(2.0) / ecc_timescale;
            const R inc_timescale = ecc_timescale;
            const R factor_inc = -P.m * R(2.0) / inc_timescale;
            const R sp = R(1.0) / (d * d) * (P.hx.x * P.hv.x + P.hx.y * P.hv.y + P.hx.z * P.hv.z);
            const V3<R> Fm{factor_migration * P.hv.x, factor_migration * P.hv.y, factor_migration * P.hv.z};
            const V3<R> Fe{factor_ecc * sp * P.hx.x, factor_ecc * sp * P.hx.y, factor_ecc * sp * P.hx.z};
            const V3<R> Fi{R(0.0), R(0.0), factor_inc * P.hv.z};
            const V3<R> F{Fm.x + Fe.x + Fi.x, Fm.y + Fe.y + Fi.y, Fm.z + Fe.z + Fi.z};
            sum.x = sum.x + F.x; sum.y = sum.y + F.y; sum.z = sum.z + F.z;
            P.disk_acc = V3<R>{factor1 * F.x, factor1 * F.y, factor1 * F.z};
        }
        H.disk_acc = V3<R>{R(-1.0) * factor2 * sum.x, R(-1.0) * factor2 * sum.y, R(-1.0) * factor2 * sum.z};
    }
}

// universe.rs:353-426 + 428-538 + 540-614
template <class R>
inline void calculate_additional_effects(System<R> &S, const R &t, bool evolution, bool dangular_momentum_dt, bool accelerations) {
    const bool init_tides = (dangular_momentum_dt && S.tides) || (accelerations && S.tides);
    const bool init_rf = (dangular_momentum_dt && S.rf) || (accelerations && S.rf);
    const bool init_wind = dangular_momentum_dt && S.wind;
    const bool init_gr = (dangular_momentum_dt && S.gr && S.gr_impl == POS_GR_KIDDER1995) || (accelerations && S.gr);
    if (init_wind)
        for (int i = 0; i < S.n; ++i)
            if (S.p[i].roles & POS_ROLE_WIND) S.p[i].wind_dLdt = v3zero<R>();
    if (init_tides) tides_initialize(S);
    if (init_rf) flattening_initialize(S);
    if (init_gr) gr_initialize(S);
    if (accelerations)
        for (int i = 0; i < S.n; ++i) S.p[i].a_add = v3zero<R>();

    calculate_spin_and_evolving_quantities(S, t, evolution);
    if (dangular_momentum_dt && S.wind) wind_factor(S);

    if (S.tides || S.rf) {
        if ((dangular_momentum_dt && (S.tides || S.rf)) || (accelerations && (S.tides || S.disk || S.rf || S.gr))) {
            if (S.tides) pair_dependent_dissipation_factors(S);
            if (S.tides) tides_orthogonal(S);
            if (S.rf) flattening_orthogonal(S);
            if (accelerations && (S.tides || S.rf)) {
                if (S.tides) { tides_radial(S); tides_acceleration(S); }
                if (S.rf) { flattening_radial(S); flattening_acceleration(S); }
            }
            if (dangular_momentum_dt && (S.tides || S.rf)) {
                if (S.tides) tides_torques(S);
                if (S.rf) flattening_torques(S);
            }
        }
    }
    if (accelerations && S.gr && S.gr_impl == POS_GR_KIDDER1995) {
        kidder_first_order(S);
        kidder_second_order(S);
        kidder_spin_orbit(S);
    }
    if (accelerations && S.gr && S.gr_impl == POS_GR_ANDERSON1975) anderson1975(S);
    if (accelerations && S.gr && S.gr_impl == POS_GR_NEWHALL1983) newhall1983(S);
    if (accelerations && S.disk) disk_acceleration(S, t);
    if (dangular_momentum_dt && (S.tides || S.rf || (S.gr && S.gr_impl == POS_GR_KIDDER1995) || S.wind)) {
        for (int i = 0; i < S.n; ++i) {
            Body<R> &P = S.p[i];
            P.dLdt.x = P.td_dLdt.x + P.rf_dLdt.x + P.gr_dLdt.x + P.wind_dLdt.x;
            P.dLdt.y = P.td_dLdt.y + P.rf_dLdt.y + P.gr_dLdt.y + P.wind_dLdt.y;
            P.dLdt.z = P.td_dLdt.z + P.rf_dLdt.z + P.gr_dLdt.z + P.wind_dLdt.z;
        }
    }
    if (accelerations) {
        for (int i = 0; i < S.n; ++i) {
            Body<R> &P = S.p[i];
            if (S.tides) { P.a_add.x = P.a_add.x + P.td_acc.x; P.a_add.y = P.a_add.y + P.td_acc.y; P.a_add.z = P.a_add.z + P.td_acc.z; }
            if (S.disk) { P.a_add.x = P.a_add.x + P.disk_acc.x; P.a_add.y = P.a_add.y + P.disk_acc.y; P.a_add.z = P.a_add.z + P.disk_acc.z; }
            if (S.rf) { P.a_add.x = P.a_add.x + P.rf_acc.x; P.a_add.y = P.a_add.y + P.rf_acc.y; P.a_add.z = P.a_add.z + P.rf_acc.z; }
            if (S.gr) { P.a_add.x = P.a_add.x + P.gr_acc.x; P.a_add.y = P.a_add.y + P.gr_acc.y; P.a_add.z = P.a_add.z + P.gr_acc.z; }
        }
    }
}

// ---------------------------------------------------------------------------
// Newtonian gravity (universe.rs:198-303)
// ---------------------------------------------------------------------------

template <class R> inline void gravity_calculate_acceleration(System<R> &S, int ignore) {
    const int mm = S.host;
    V3<R> acc[MAXB];
    for (int i = 0; i < S.n; ++i) acc[i] = v3zero<R>();
    for (int i = 0; i < S.n; ++i) {
        for (int j = 0; j < S.n; ++j) {
            if (i == j) continue;
            const R dx = S.p[i].x.x - S.p[j].x.x;
            const R dy = S.p[i].x.y - S.p[j].x.y;
            const R dz = S.p[i].x.z - S.p[j].x.z;
            const R distance_2 = dx * dx + dy * dy + dz * dz;
            if (i < j) {
                const double rr = S.roche[i * S.n + j];
                if (val(distance_2) <= rr * rr) S.status |= POS_STATUS_ROCHE;
                if (val(distance_2) <= val(pw2(S.p[i].radius + S.p[j].radius))) S.status |= POS_STATUS_COLLISION;
                if (MAX_DISTANCE_2 > 0. && i == mm && val(distance_2) > MAX_DISTANCE_2) S.status |= POS_STATUS_EJECTION;
            }
            if (ignore == IGNORE_WHFAST_ONE &&
                ((i == mm && mm == 0 && j == 1) || (i == mm && mm > 0 && j == 0) ||
                 (j == mm && mm == 0 && i == 1) || (j == mm && mm > 0 && i == 0)))
                continue;
            if (ignore == IGNORE_WHFAST_TWO && (i == mm || j == mm)) continue;
            const R distance = sqrt_r(distance_2);
            const R prefact = R(-G) / (distance * distance * distance) * S.p[j].m;
            acc[i].x = acc[i].x + prefact * dx;
            acc[i].y = acc[i].y + prefact * dy;
            acc[i].z = acc[i].z + prefact * dz;
        }
    }
    for (int i = 0; i < S.n; ++i) S.p[i].a_grav = acc[i];
}

// ---------------------------------------------------------------------------
// WHFast operators (whfast.rs:469-1155)
// ---------------------------------------------------------------------------
// Stumpff c0..c3 with argument reduction (whfast.rs:844-876)
template <class R> inline void stumpff_cs3(R z, R cs[4], uint32_t *status) {
    static const double invfactorial[14] = {1., 1., 1. / 2., 1. / 6., 1. / 24., 1. / 120., 1. / 720., 1. / 5040.,
                                           1. / 40320., 1. / 362880., 1. / 3628800., 1. / 39916800.,
                                           1. / 479001600., 1. / 6227020800.};
    int n = 0;
    while (rfabs(val(z)) > 0.1) {
        if (n >= STUMPFF_REDUCTION_CAP) { *status |= POS_STATUS_ITERATION_CAP; break; }
        z = z / R(4.);
        n += 1;
    }
    const int nmax = 13;
    R c_odd = R(invfactorial[nmax]);
    R c_even = R(invfactorial[nmax - 1]);
    for (int np = nmax - 2; np >= 3; np -= 2) {
        c_odd = R(invfactorial[np]) - z * c_odd;
        c_even = R(invfactorial[np - 1]) - z * c_even;
    }
    cs[3] = c_odd;
    cs[2] = c_even;
    cs[1] = R(invfactorial[1]) - z * c_odd;
    cs[0] = R(invfactorial[0]) - z * c_even;
    while (n > 0) {
        cs[3] = (cs[2] + cs[0] * cs[3]) * R(0.25);
        cs[2] = cs[1] * cs[1] * R(0.5);
        cs[1] = cs[0] * cs[1];
        cs[0] = R(2.) * cs[0] * cs[0] - R(1.);
        n -= 1;
    }
}
// whfast.rs:835-842
template <class R> inline void stiefel_gs3(const R &beta, const R &x, R gs[4], uint32_t *status) {
    const R x2 = pw2(x);
    stumpff_cs3(beta * x2, gs, status);
    gs[1] = gs[1] * x;
    gs[2] = gs[2] * x2;
    gs[3] = gs[3] * (x2 * x);
}

// Tangent-linear runs only (R = Dual<K>): the derivative of the bisection's
// root.  The bisection (whfast.rs:776-808) finds the root of the universal
// Kepler equation s(x) = r0 x + eta0 G2(x) + zeta0 G3(x) - dt to 1e-15, so
// the step's map follows that root; differentiating the bisection iterates
// instead gives the derivatives of the bracket's end points.  The root's
// derivative is dx = -s_p / s_x (implicit function theorem, with
// s_x = r0 + eta0 G1 + zeta0 G2), and the Stiefel functions of the last
// evaluation take their derivatives from it.  Values are not touched.
template <class R> struct IsDual { static constexpr bool value = false; };
template <int K> struct IsDual<Dual<K>> { static constexpr bool value = true; };
template <class R>
inline void bisection_root_derivative(const R &beta, const R &r0, const R &eta0, const R &zeta0, const R &dt,
                                      R &x_eval, R &x, R gs[4], uint32_t *status) {
    if constexpr (IsDual<R>::value) {
        const R xc(val(x_eval));   // the last evaluated point, no derivative
        R g[4];
        uint32_t st = 0;
        stiefel_gs3(beta, xc, g, &st);
        const R sres = r0 * xc + eta0 * g[2] + zeta0 * g[3] - dt;
        const double sx = val(r0) + val(eta0) * val(g[1]) + val(zeta0) * val(g[2]);
        R xd(val(x_eval));
        for (int k = 0; k < (int)(sizeof(xd.d) / sizeof(xd.d[0])); ++k) xd.d[k] = -sres.d[k] / sx;
        x_eval = xd;
        stiefel_gs3(beta, x_eval, gs, status);   // same values as the last evaluation
        R xf(val(x));
        for (int k = 0; k < (int)(sizeof(xf.d) / sizeof(xf.d[0])); ++k) xf.d[k] = xd.d[k];
        x = xf;
    } else {
        (void)beta; (void)r0; (void)eta0; (void)zeta0; (void)dt; (void)x_eval; (void)x; (void)gs; (void)status;
    }
}

// whfast.rs:676-833
template <class R> inline void kepler_individual_step(System<R> &S, int i, const R &mass_g, const R &dt) {
    Alt<R> &A = S.alt[i];
    const V3<R> x0 = A.x, v0 = A.v;
    const R r0 = sqrt_r(pw2(x0.x) + pw2(x0.y) + pw2(x0.z));
    const R r0i = R(1.) / r0;
    const R v2 = pw2(v0.x) + pw2(v0.y) + pw2(v0.z);
    const R beta = R(2.) * mass_g * r0i - v2;
    const R eta0 = x0.x * v0.x + x0.y * v0.y + x0.z * v0.z;
    const R zeta0 = mass_g - beta * r0;
    R x, gs[4];
    R invperiod = R(0.);
    R x_per_period;
    const bool elliptic = val(beta) > 0.;
    if (elliptic) {
        const R sqrt_beta = sqrt_r(beta);
        invperiod = sqrt_beta * beta / (R(2.) * R(PI) * mass_g);
        x_per_period = R(2. * PI) / sqrt_beta;
        if (rfabs(val(dt)) * val(invperiod) > 1. && S.timestep_warning == 0) {
            S.status |= POS_STATUS_TIMESTEP_WARNING;
            S.timestep_warning += 1;
        }
        const R dtr0i = dt * r0i;
        x = dtr0i * (R(1.) - dtr0i * eta0 * R(0.5) * r0i);
    } else {
        x = R(0.);
        x_per_period = R(std::numeric_limits<double>::quiet_NaN());
    }
    int converged = 0;
    R old_x = x;
    stiefel_gs3(beta, x, gs, &S.status);
    R eta0_gs1_zeta0_gs2 = eta0 * gs[1] + zeta0 * gs[2];
    R ri = R(1.) / (r0 + eta0_gs1_zeta0_gs2);
    x = ri * (x * eta0_gs1_zeta0_gs2 - eta0 * gs[2] - zeta0 * gs[3] + dt);

    if (!risnan(val(x_per_period)) && rfabs(val(x) - val(old_x)) > 0.01 * val(x_per_period)) {
        // Laguerre-Conway quartic solver
        x = beta * dt / mass_g;
        double prev_x[WHFAST_NMAX_QUART + 1];
        for (int k = 0; k <= WHFAST_NMAX_QUART; ++k) prev_x[k] = 0.;
        for (int n_lag = 1; n_lag < WHFAST_NMAX_QUART; ++n_lag) {
            stiefel_gs3(beta, x, gs, &S.status);
            const R f = r0 * x + eta0 * gs[2] + zeta0 * gs[3] - dt;
            const R fp = r0 + eta0 * gs[1] + zeta0 * gs[2];
            const R fpp = eta0 * gs[0] + zeta0 * gs[1];
            const R denom = fp + sqrt_r(abs_r(R(16.) * fp * fp - R(20.) * f * fpp));
            x = (x * denom - R(5.) * f) / denom;
            bool hit = false;
            for (int k = 1; k < n_lag; ++k)
                if (val(x) == prev_x[k]) { hit = true; break; }
            if (hit) { converged = 1; break; }
            prev_x[n_lag] = val(x);
        }
        eta0_gs1_zeta0_gs2 = eta0 * gs[1] + zeta0 * gs[2];
        ri = R(1.) / (r0 + eta0_gs1_zeta0_gs2);
    } else {
        // Newton's method
        R old_x2;
        for (int k = 1; k < WHFAST_NMAX_NEWT; ++k) {
            old_x2 = old_x;
            old_x = x;
            stiefel_gs3(beta, x, gs, &S.status);
            eta0_gs1_zeta0_gs2 = eta0 * gs[1] + zeta0 * gs[2];
            ri = R(1.) / (r0 + eta0_gs1_zeta0_gs2);
            x = ri * (x * eta0_gs1_zeta0_gs2 - eta0 * gs[2] - zeta0 * gs[3] + dt);
            if (val(x) == val(old_x) || val(x) == val(old_x2)) { converged = 1; break; }
        }
    }
    if (converged == 0) {
        S.status |= POS_STATUS_KEPLER_BISECTION;
        R x_min, x_max;
        if (elliptic) {
            x_min = x_per_period * R(rfloor(val(dt * invperiod)));
            x_max = x_min + x_per_period;
        } else {
            const R h2 = r0 * r0 * v2 - eta0 * eta0;
            const R q = h2 / mass_g / (R(1.) + sqrt_r(R(1.) - h2 * beta / (mass_g * mass_g)));
            const R vq = sqrt_r(h2) / q;
            x_min = R(1.) / (vq + r0 / dt);
            x_max = dt / q;
        }
        x = (x_max + x_min) / R(2.);
        int it = 0;
        R x_eval = x;
        for (;;) {
            if (++it > BISECTION_CAP) { S.status |= POS_STATUS_ITERATION_CAP; break; }
            x_eval = x;
            stiefel_gs3(beta, x, gs, &S.status);
            const R s = r0 * x + eta0 * gs[2] + zeta0 * gs[3] - dt;
            if (val(s) >= 0.) x_max = x; else x_min = x;
            x = (x_max + x_min) / R(2.);
            if (rfabs(val(x_max) - val(x_min)) / val(x_max) <= 1e-15) break;
        }
        bisection_root_derivative(beta, r0, eta0, zeta0, dt, x_eval, x, gs, &S.status);
        eta0_gs1_zeta0_gs2 = eta0 * gs[1] + zeta0 * gs[2];
        ri = R(1.) / (r0 + eta0_gs1_zeta0_gs2);
    }
    if (risnan(val(ri))) {
        ri = R(0.); gs[1] = R(0.); gs[2] = R(0.); gs[3] = R(0.);
    }
    const R f = -mass_g * gs[2] * r0i;
    const R g = dt - mass_g * gs[3];
    const R fd = -mass_g * gs[1] * r0i * ri;
    const R gd = -mass_g * gs[2] * ri;
    A.x.x = A.x.x + (f * x0.x + g * v0.x);
    A.x.y = A.x.y + (f * x0.y + g * v0.y);
    A.x.z = A.x.z + (f * x0.z + g * v0.z);
    A.v.x = A.v.x + (fd * x0.x + gd * v0.x);
    A.v.y = A.v.y + (fd * x0.y + gd * v0.y);
    A.v.z = A.v.z + (fd * x0.z + gd * v0.z);
}

template <class R> inline void kepler_steps(System<R> &S, const R &dt) {
    const int h = S.host;
    if (S.coord == POS_COORD_JACOBI) {
        R star_planets_mass_g = S.p[h].mg;
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            star_planets_mass_g = star_planets_mass_g + S.alt[i].mg;
            kepler_individual_step(S, i, star_planets_mass_g, dt);
        }
    } else if (S.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) {
        const R star_mass_g = S.p[h].mg;
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            kepler_individual_step(S, i, star_mass_g, dt);
        }
    } else {
        const R star_mass_g = S.p[h].mg;
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            kepler_individual_step(S, i, star_mass_g + S.alt[i].mg, dt);
        }
    }
    Alt<R> &AH = S.alt[h];
    AH.x.x = AH.x.x + dt * AH.v.x;
    AH.x.y = AH.x.y + dt * AH.v.y;
    AH.x.z = AH.x.z + dt * AH.v.z;
}

template <class R> inline void jump_step(System<R> &S, const R &dt) {
    const int h = S.host;
    if (S.coord == POS_COORD_JACOBI) return;
    const R m0 = S.p[h].m;
    V3<R> p = v3zero<R>();
    if (S.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) {
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            const R m = S.p[i].m;
            p.x = p.x + m * S.alt[i].v.x; p.y = p.y + m * S.alt[i].v.y; p.z = p.z + m * S.alt[i].v.z;
        }
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            S.alt[i].x.x = S.alt[i].x.x + dt * p.x / m0;
            S.alt[i].x.y = S.alt[i].x.y + dt * p.y / m0;
            S.alt[i].x.z = S.alt[i].x.z + dt * p.z / m0;
        }
    } else {
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            const R m = S.p[i].m;
            const R factor = m0 + m;
            p.x = p.x + m * S.alt[i].v.x / factor;
            p.y = p.y + m * S.alt[i].v.y / factor;
            p.z = p.z + m * S.alt[i].v.z / factor;
        }
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            const R m = S.p[i].m;
            const R factor = m0 + m;
            S.alt[i].x.x = S.alt[i].x.x + dt * (p.x - (m * S.alt[i].v.x / factor));
            S.alt[i].x.y = S.alt[i].x.y + dt * (p.y - (m * S.alt[i].v.y / factor));
            S.alt[i].x.z = S.alt[i].x.z + dt * (p.z - (m * S.alt[i].v.z / factor));
        }
    }
}

// whfast.rs:935-963
template <class R> inline void inertial_to_jacobi_acc(System<R> &S) {
    const int h = S.host;
    R eta = S.p[h].m;
    V3<R> s{eta * S.p[h].a_grav.x, eta * S.p[h].a_grav.y, eta * S.p[h].a_grav.z};
    for (int i = 0; i < S.n; ++i) {
        if (i == h) continue;
        const Body<R> &P = S.p[i];
        Alt<R> &A = S.alt[i];
        const R ei = R(1.) / eta;
        eta = eta + P.m;
        const R pme = eta * ei;
        A.a.x = P.a_grav.x - s.x * ei; A.a.y = P.a_grav.y - s.y * ei; A.a.z = P.a_grav.z - s.z * ei;
        s.x = s.x * pme + P.m * A.a.x; s.y = s.y * pme + P.m * A.a.y; s.z = s.z * pme + P.m * A.a.z;
    }
    const R mtot_i = R(1.) / eta;
    S.alt[h].a.x = s.x * mtot_i; S.alt[h].a.y = s.y * mtot_i; S.alt[h].a.z = s.z * mtot_i;
}

template <class R> inline void interaction_step(System<R> &S, const R &dt) {
    const int h = S.host;
    if (S.coord == POS_COORD_JACOBI) {
        inertial_to_jacobi_acc(S);
        const double softening = 1e-12;
        const R m0 = S.p[h].m;
        R eta = m0;
        int k = 0;
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            Alt<R> &A = S.alt[i];
            eta = eta + A.m;
            A.v.x = A.v.x + dt * A.a.x; A.v.y = A.v.y + dt * A.a.y; A.v.z = A.v.z + dt * A.a.z;
            if (k > 0) {
                const R rj2i = R(1.) / (pw2(A.x.x) + pw2(A.x.y) + pw2(A.x.z) + R(softening));
                const R rji = sqrt_r(rj2i);
                const R rj3im = rji * rj2i * R(G) * eta;
                const R prefac = dt * rj3im;
                A.v.x = A.v.x + prefac * A.x.x; A.v.y = A.v.y + prefac * A.x.y; A.v.z = A.v.z + prefac * A.x.z;
            }
            ++k;
        }
    } else if (S.coord == POS_COORD_DEMOCRATIC_HELIOCENTRIC) {
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            Alt<R> &A = S.alt[i];
            const Body<R> &P = S.p[i];
            A.v.x = A.v.x + dt * P.a_grav.x; A.v.y = A.v.y + dt * P.a_grav.y; A.v.z = A.v.z + dt * P.a_grav.z;
        }
    } else {
        const R m0 = S.p[h].m;
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            Alt<R> &A = S.alt[i];
            const Body<R> &P = S.p[i];
            const R factor = m0 + P.m;
            A.v.x = A.v.x + dt * factor * P.a_grav.x / m0;
            A.v.y = A.v.y + dt * factor * P.a_grav.y / m0;
            A.v.z = A.v.z + dt * factor * P.a_grav.z / m0;
        }
    }
}

// whfast.rs:889-933 (Jacobi) and 973-1023 (DH, WHDS)
template <class R> inline void inertial_to_alternative_posvel(System<R> &S) {
    const int h = S.host;
    Body<R> &H = S.p[h];
    Alt<R> &AH = S.alt[h];
    if (S.coord == POS_COORD_JACOBI) {
        const R m0 = H.m;
        R eta = m0;
        V3<R> s{eta * H.x.x, eta * H.x.y, eta * H.x.z};
        V3<R> sv{eta * H.v.x, eta * H.v.y, eta * H.v.z};
        for (int i = 0; i < S.n; ++i) {
            if (i == h) continue;
            const Body<R> &P = S.p[i];
            Alt<R> &A = S.alt[i];
            const R ei = R(1.) / eta;
            eta = eta + P.m;
            const R pme = eta * ei;
            A.m = P.m; A.mg = P.mg;
            A.x.x = P.x.x - s.x * ei; A.x.y = P.x.y - s.y * ei; A.x.z = P.x.z - s.z * ei;
            A.v.x = P.v.x - sv.x * ei; A.v.y = P.v.y - sv.y * ei; A.v.z = P.v.z - sv.z * ei;
            s.x = s.x * pme + P.m * A.x.x; s.y = s.y * pme + P.m * A.x.y; s.z = s.z * pme + P.m * A.x.z;
            sv.x = sv.x * pme + P.m * A.v.x; sv.y = sv.y * pme + P.m * A.v.y; sv.z = sv.z * pme + P.m * A.v.z;
        }
        const R mtot_i = R(1.) / eta;
        AH.m = eta;
        AH.x.x = s.x * mtot_i; AH.x.y = s.y * mtot_i; AH.x.z = s.z * mtot_i;
        AH.v.x = sv.x * mtot_i; AH.v.y = sv.y * mtot_i; AH.v.z = sv.z * mtot_i;
        return;
    }
    AH.x = v3zero<R>(); AH.v = v3zero<R>(); AH.m = R(0.); AH.mg = R(0.);
    for (int k = -1; k < S.n; ++k) {   // the host first, then the others in order
        const int i = (k < 0) ? h : k;
        if (k >= 0 && k == h) continue;
        const Body<R> &P = S.p[i];
        AH.x.x = AH.x.x + P.x.x * P.m; AH.x.y = AH.x.y + P.x.y * P.m; AH.x.z = AH.x.z + P.x.z * P.m;
        AH.v.x = AH.v.x + P.v.x * P.m; AH.v.y = AH.v.y + P.v.y * P.m; AH.v.z = AH.v.z + P.v.z * P.m;
        AH.m = AH.m + P.m;
        AH.mg = AH.mg + P.mg;
    }
    const R mtot = AH.m;
    AH.x.x = AH.x.x / mtot; AH.x.y = AH.x.y / mtot; AH.x.z = AH.x.z / mtot;
    AH.v.x = AH.v.x / mtot; AH.v.y = AH.v.y / mtot; AH.v.z = AH.v.z / mtot;
    for (int i = 0; i < S.n; ++i) {
        if (i == h) continue;
        const Body<R> &P = S.p[i];
        Alt<R> &A = S.alt[i];
        A.x.x = P.x.x - H.x.x; A.x.y = P.x.y - H.x.y; A.x.z = P.x.z - H.x.z;
        A.v.x = P.v.x - AH.v.x; A.v.y = P.v.y - AH.v.y; A.v.z = P.v.z - AH.v.z;
        A.m = P.m; A.mg = P.mg;
        if (S.coord == POS_COORD_WHDS) {
            const R factor = (H.m + P.m) / H.m;
            A.v.x = A.v.x * factor; A.v.y = A.v.y * factor; A.v.z = A.v.z * factor;
        }
    }
}

// whfast.rs:1034-1155
template <class R> inline void alternative_to_inertial_posvel(System<R> &S) {
    const int h = S.host;
    Body<R> &H = S.p[h];
    Alt<R> &AH = S.alt[h];
    if (S.coord == POS_COORD_JACOBI) {
        const R m0 = AH.m;
        R eta = m0;
        V3<R> s{eta * AH.x.x, eta * AH.x.y, eta * AH.x.z};
        V3<R> sv{eta * AH.v.x, eta * AH.v.y, eta * AH.v.z};
        for (int i = S.n - 1; i >= 0; --i) {
            if (i == h) continue;
            Body<R> &P = S.p[i];
            const Alt<R> &A = S.alt[i];
            const R ei = R(1.) / eta;
            s.x = (s.x - P.m * A.x.x) * ei; s.y = (s.y - P.m * A.x.y) * ei; s.z = (s.z - P.m * A.x.z) * ei;
            sv.x = (sv.x - P.m * A.v.x) * ei; sv.y = (sv.y - P.m * A.v.y) * ei; sv.z = (sv.z - P.m * A.v.z) * ei;
            P.x.x = A.x.x + s.x; P.x.y = A.x.y + s.y; P.x.z = A.x.z + s.z;
            P.v.x = A.v.x + sv.x; P.v.y = A.v.y + sv.y; P.v.z = A.v.z + sv.z;
            eta = eta - P.m;
            s.x = s.x * eta; s.y = s.y * eta; s.z = s.z * eta;
            sv.x = sv.x * eta; sv.y = sv.y * eta; sv.z = sv.z * eta;
        }
        const R mtot_i = R(1.) / eta;
        H.x.x = s.x * mtot_i; H.x.y = s.y * mtot_i; H.x.z = s.z * mtot_i;
        H.v.x = sv.x * mtot_i; H.v.y = sv.y * mtot_i; H.v.z = sv.z * mtot_i;
        return;
    }
    // positions
    const R mtot = AH.m;
    V3<R> new_star_position = AH.x;
    for (int i = 0; i < S.n; ++i) {
        if (i == h) continue;
        const Alt<R> &A = S.alt[i];
        const R m = S.p[i].m;
        new_star_position.x = new_star_position.x - A.x.x * m / mtot;
        new_star_position.y = new_star_position.y - A.x.y * m / mtot;
        new_star_position.z = new_star_position.z - A.x.z * m / mtot;
    }
    H.x = new_star_position;
    for (int i = 0; i < S.n; ++i) {
        if (i == h) continue;
        const Alt<R> &A = S.alt[i];
        Body<R> &P = S.p[i];
        P.x.x = A.x.x + H.x.x; P.x.y = A.x.y + H.x.y; P.x.z = A.x.z + H.x.z;
    }
    // velocities
    const R m0 = H.m;
    V3<R> new_star_velocity = AH.v;
    for (int i = 0; i < S.n; ++i) {
        if (i == h) continue;
        const Alt<R> &A = S.alt[i];
        Body<R> &P = S.p[i];
        R factor;
        if (S.coord == POS_COORD_WHDS) {
            const R f = (m0 + P.m) / m0;
            P.v.x = A.v.x / f + AH.v.x; P.v.y = A.v.y / f + AH.v.y; P.v.z = A.v.z / f + AH.v.z;
            factor = P.m / (m0 + P.m);
        } else {
            P.v.x = A.v.x + AH.v.x; P.v.y = A.v.y + AH.v.y; P.v.z = A.v.z + AH.v.z;
            factor = P.m / m0;
        }
        new_star_velocity.x = new_star_velocity.x - A.v.x * factor;
        new_star_velocity.y = new_star_velocity.y - A.v.y * factor;
        new_star_velocity.z = new_star_velocity.z - A.v.z * factor;
    }
    H.v = new_star_velocity;
}

// ---------------------------------------------------------------------------
// Implicit midpoint for velocity-dependent forces (whfast.rs:322-466)
// ---------------------------------------------------------------------------
template <class R> inline void integrate_velocity_dependent_forces(System<R> &S, const R &dt, bool integrate_spin, bool evolution) {
    const int n = S.n;
    V3<R> v_orig[MAXB], L_orig[MAXB], v_final[MAXB], L_final[MAXB], v_prev[MAXB], L_prev[MAXB];
    V3<R> dv[MAXB], dL[MAXB];
    for (int i = 0; i < n; ++i) {
        v_orig[i] = S.p[i].v; L_orig[i] = S.p[i].L;
        v_final[i] = v_orig[i]; L_final[i] = L_orig[i];
        dv[i] = v3zero<R>(); dL[i] = v3zero<R>();
    }
    bool converged = false;
    for (int it = 0; it < IMPLICIT_MIDPOINT_MAX_ITER; ++it) {
        for (int i = 0; i < n; ++i) { v_prev[i] = v_final[i]; L_prev[i] = L_final[i]; }
        inertial_to_heliocentric(S);
        calculate_additional_effects(S, S.t, evolution && it == 0, integrate_spin, true);
        S.midpoint_iterations += 1;
        for (int i = 0; i < n; ++i) {
            const Body<R> &P = S.p[i];
            dv[i].x = dt * P.a_add.x - S.err_v[i].x;
            dv[i].y = dt * P.a_add.y - S.err_v[i].y;
            dv[i].z = dt * P.a_add.z - S.err_v[i].z;
            v_final[i].x = v_orig[i].x + dv[i].x;
            v_final[i].y = v_orig[i].y + dv[i].y;
            v_final[i].z = v_orig[i].z + dv[i].z;
            if (integrate_spin) {
                dL[i].x = dt * P.dLdt.x - S.err_L[i].x;
                dL[i].y = dt * P.dLdt.y - S.err_L[i].y;
                dL[i].z = dt * P.dLdt.z - S.err_L[i].z;
                L_final[i].x = L_orig[i].x + dL[i].x;
                L_final[i].y = L_orig[i].y + dL[i].y;
                L_final[i].z = L_orig[i].z + dL[i].z;
            }
        }
        if (it >= IMPLICIT_MIDPOINT_MIN_ITER - 1) {
            R fv2 = R(0.), dv2 = R(0.), fL2 = R(0.), dL2 = R(0.);
            for (int i = 0; i < n; ++i) {
                const R dvx = v_final[i].x - v_prev[i].x, dvy = v_final[i].y - v_prev[i].y, dvz = v_final[i].z - v_prev[i].z;
                dv2 = dv2 + (pw2(dvx) + pw2(dvy) + pw2(dvz));
                fv2 = fv2 + (pw2(v_final[i].x) + pw2(v_final[i].y) + pw2(v_final[i].z));
                if (integrate_spin) {
                    const R dsx = L_final[i].x - L_prev[i].x, dsy = L_final[i].y - L_prev[i].y, dsz = L_final[i].z - L_prev[i].z;
                    dL2 = dL2 + (pw2(dsx) + pw2(dsy) + pw2(dsz));
                    fL2 = fL2 + (pw2(L_final[i].x) + pw2(L_final[i].y) + pw2(L_final[i].z));
                }
            }
            const bool ok = (integrate_spin && val(dv2 / fv2) < DBL_EPS_2 && val(dL2 / fL2) < DBL_EPS_2) ||
                            (!integrate_spin && val(dv2 / fv2) < DBL_EPS_2);
            if (ok) { converged = true; break; }
        }
        for (int i = 0; i < n; ++i) {
            S.p[i].v.x = R(0.5) * (v_orig[i].x + v_final[i].x);
            S.p[i].v.y = R(0.5) * (v_orig[i].y + v_final[i].y);
            S.p[i].v.z = R(0.5) * (v_orig[i].z + v_final[i].z);
            if (integrate_spin) {
                S.p[i].L.x = R(0.5) * (L_orig[i].x + L_final[i].x);
                S.p[i].L.y = R(0.5) * (L_orig[i].y + L_final[i].y);
                S.p[i].L.z = R(0.5) * (L_orig[i].z + L_final[i].z);
            }
        }
    }
    if (!converged) S.status |= POS_STATUS_MIDPOINT_NOT_CONVERGED;
    for (int i = 0; i < n; ++i) {
        Body<R> &P = S.p[i];
        P.v = v_final[i];
        S.err_v[i].x = (P.v.x - v_orig[i].x) - dv[i].x;
        S.err_v[i].y = (P.v.y - v_orig[i].y) - dv[i].y;
        S.err_v[i].z = (P.v.z - v_orig[i].z) - dv[i].z;
        if (integrate_spin) {
            P.L = L_final[i];
            S.err_L[i].x = (P.L.x - L_orig[i].x) - dL[i].x;
            S.err_L[i].y = (P.L.y - L_orig[i].y) - dL[i].y;
            S.err_L[i].z = (P.L.z - L_orig[i].z) - dL[i].z;
        }
    }
}

// ---------------------------------------------------------------------------
// WHFast::iterate (whfast.rs:235-305).  Returns false once the system is
// completed or frozen.
// ---------------------------------------------------------------------------
template <class R> inline bool iterate(System<R> &S) {
    if (S.status & (POS_STATUS_COMPLETED | POS_STATUS_FATAL)) return false;
    const bool first_snapshot_trigger = S.last_hist < 0.;
    const bool historic_snapshot_time_trigger = S.last_hist + S.hist_period <= val(S.t);
    if (first_snapshot_trigger || historic_snapshot_time_trigger) {
        inertial_to_heliocentric(S);
        calculate_spin_and_evolving_quantities(S, S.t, true);
        if (!first_snapshot_trigger) S.last_hist += S.hist_period;
        else S.last_hist = 0.;
    }
    const int ignore = (S.coord == POS_COORD_JACOBI) ? IGNORE_WHFAST_ONE : IGNORE_WHFAST_TWO;
    const bool kidder_dL = S.gr && S.gr_impl == POS_GR_KIDDER1995;
    const bool integrate_spin = S.tides || S.rf || S.evolution || kidder_dL;

    integrate_velocity_dependent_forces(S, S.half_dt, integrate_spin, true);
    inertial_to_alternative_posvel(S);
    kepler_steps(S, S.half_dt);
    jump_step(S, S.half_dt);
    alternative_to_inertial_posvel(S);
    gravity_calculate_acceleration(S, ignore);
    interaction_step(S, S.dt);
    jump_step(S, S.half_dt);
    kepler_steps(S, S.half_dt);
    alternative_to_inertial_posvel(S);
    integrate_velocity_dependent_forces(S, S.half_dt, integrate_spin, false);
    S.t = S.t + S.dt;
    S.iteration += 1;
    if (val(S.t + S.dt) > val(S.time_limit)) S.status |= POS_STATUS_COMPLETED;
    return !(S.status & (POS_STATUS_COMPLETED | POS_STATUS_FATAL));
}

}  // namespace oracle
