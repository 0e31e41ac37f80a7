// oracle_capi.cpp -- extern "C" entry points of the CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py through ctypes.  It consumes the same
// pos_ensemble_desc / pos_state structures as the product library so that
// both sides read identical inputs.
#include <atomic>
#include <limits>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "oracle_types.hpp"
#include "posidonius_oracle.hpp"

namespace {

thread_local std::string g_err;

template <class R> void store_state(const oracle::System<R> &S, int b0, int s, pos_state *o) {
    using oracle::val;
    for (int i = 0; i < S.n; ++i) {
        const int b = b0 + i;
        const auto &P = S.p[i];
        auto put3 = [&](double *dst, const oracle::V3<R> &v) {
            if (!dst) return;
            dst[3 * b] = val(v.x); dst[3 * b + 1] = val(v.y); dst[3 * b + 2] = val(v.z);
        };
        put3(o->position, P.x);
        put3(o->velocity, P.v);
        put3(o->acceleration, P.a_grav);
        put3(o->angular_momentum, P.L);
        put3(o->spin, P.spin);
        put3(o->velocity_error, S.err_v[i]);
        put3(o->angular_momentum_error, S.err_L[i]);
        if (o->radius) o->radius[b] = val(P.radius);
        if (o->radius_of_gyration_2) o->radius_of_gyration_2[b] = val(P.rg2);
        if (o->moment_of_inertia) o->moment_of_inertia[b] = val(P.inertia);
        if (o->evolver_left_index) o->evolver_left_index[b] = (int32_t)S.ev[i].left_index;
        if (o->pair_dissipation_factor)
            o->pair_dissipation_factor[b] = P.ph_set ? val(P.ph) : std::numeric_limits<double>::quiet_NaN();
    }
    if (o->roche_radius)
        for (int i = 0; i < S.n; ++i)
            for (int j = 0; j < oracle::MAXB; ++j)
                o->roche_radius[(size_t)(b0 + i) * POS_MAX_BODIES + j] = (j < S.n && j != i) ? S.roche[i * S.n + j] : 0.;
    if (o->timestep_warning) o->timestep_warning[s] = S.timestep_warning;
    if (o->current_time) o->current_time[s] = val(S.t);
    if (o->last_historic_snapshot_time) o->last_historic_snapshot_time[s] = S.last_hist;
    if (o->current_iteration) o->current_iteration[s] = S.iteration;
    if (o->status) o->status[s] = S.status;
}

}  // namespace

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

// One universal-variable Kepler drift (kepler_individual_step, whfast.rs:676-876)
// of a single body relative to a central mass with G M = mass_g, over dt; the
// status bits it raises (timestep warning, bisection fallback, iteration cap)
// are returned in *status.  Used by the tests to check the drift against the
// closed-form two-body solution.
int oracle_kepler_drift(double mass_g, double dt, const double *x, const double *v, double *out_x, double *out_v,
                        uint32_t *status) {
    auto *S = new oracle::System<double>();
    S->alt[1].x = oracle::V3<double>{x[0], x[1], x[2]};
    S->alt[1].v = oracle::V3<double>{v[0], v[1], v[2]};
    oracle::kepler_individual_step(*S, 1, mass_g, dt);
    const oracle::V3<double> X = S->alt[1].x, W = S->alt[1].v;
    out_x[0] = X.x; out_x[1] = X.y; out_x[2] = X.z;
    out_v[0] = W.x; out_v[1] = W.y; out_v[2] = W.z;
    *status = S->status;
    delete S;
    return 0;
}

// Integrate every system for up to n_steps calls of WHFast::iterate (fewer if
// it completes or freezes) on n_threads host threads, one system at a time
// per thread.  If out_midpoint_iterations is non-NULL it receives, per system,
// the number of additional-effects evaluations executed.  Returns 0 or -1.
int oracle_run(const pos_ensemble_desc *d, int64_t n_steps, int32_t n_threads, pos_state *out,
               int64_t *out_steps_done, int64_t *out_midpoint_iterations) {
    const int ns = d->n_systems;
    std::atomic<int> next{0};
    std::atomic<int> failed{-1};
    std::string first_err;
    auto worker = [&]() {
        oracle::System<double> *S = new oracle::System<double>();
        for (;;) {
            const int s = next.fetch_add(1);
            if (s >= ns) break;
            std::string err;
            if (!oracle::load_system(d, s, *S, &err)) {
                int expected = -1;
                if (failed.compare_exchange_strong(expected, s)) first_err = err;
                continue;
            }
            oracle::start_system(*S, d, s);
            int64_t k = 0;
            while (k < n_steps) {
                ++k;
                if (!oracle::iterate(*S)) break;
            }
            store_state(*S, d->body_offset[s], s, out);
            if (out_steps_done) out_steps_done[s] = k;
            if (out_midpoint_iterations) out_midpoint_iterations[s] = S->midpoint_iterations;
        }
        delete S;
    };
    const int nt = n_threads < 1 ? 1 : n_threads;
    if (nt == 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int i = 0; i < nt; ++i) pool.emplace_back(worker);
        for (auto &t : pool) t.join();
    }
    if (failed.load() >= 0) {
        g_err = "system " + std::to_string(failed.load()) + ": " + first_err;
        return -1;
    }
    return 0;
}

// Operator-level evaluations at the current state (no time advance).
// Universe::gravity_calculate_acceleration with the WHFast ignore rule of
// the system's coordinates (whfast.rs:263-267).
int oracle_gravity(const pos_ensemble_desc *d, double *out_acc, uint32_t *out_status) {
    oracle::System<double> *S = new oracle::System<double>();
    for (int s = 0; s < d->n_systems; ++s) {
        std::string err;
        if (!oracle::load_system(d, s, *S, &err)) { g_err = err; delete S; return -1; }
        oracle::start_system(*S, d, s);
        const int ignore = S->coord == POS_COORD_JACOBI ? oracle::IGNORE_WHFAST_ONE : oracle::IGNORE_WHFAST_TWO;
        oracle::gravity_calculate_acceleration(*S, ignore);
        const int b0 = d->body_offset[s];
        for (int i = 0; i < S->n; ++i) {
            out_acc[3 * (b0 + i)] = S->p[i].a_grav.x;
            out_acc[3 * (b0 + i) + 1] = S->p[i].a_grav.y;
            out_acc[3 * (b0 + i) + 2] = S->p[i].a_grav.z;
        }
        if (out_status) out_status[s] |= S->status;
    }
    delete S;
    return 0;
}

// inertial_to_heliocentric + calculate_additional_effects(t, evolution=false,
// dangular_momentum_dt=integrate_spin, accelerations=true).
int oracle_additional_effects(const pos_ensemble_desc *d, double *out_acc, double *out_dLdt) {
    oracle::System<double> *S = new oracle::System<double>();
    for (int s = 0; s < d->n_systems; ++s) {
        std::string err;
        if (!oracle::load_system(d, s, *S, &err)) { g_err = err; delete S; return -1; }
        oracle::start_system(*S, d, s);
        const bool integrate_spin = S->tides || S->rf || S->evolution || (S->gr && S->gr_impl == POS_GR_KIDDER1995);
        oracle::inertial_to_heliocentric(*S);
        oracle::calculate_additional_effects(*S, S->t, false, integrate_spin, true);
        const int b0 = d->body_offset[s];
        for (int i = 0; i < S->n; ++i) {
            const int b = b0 + i;
            out_acc[3 * b] = S->p[i].a_add.x; out_acc[3 * b + 1] = S->p[i].a_add.y; out_acc[3 * b + 2] = S->p[i].a_add.z;
            out_dLdt[3 * b] = S->p[i].dLdt.x; out_dLdt[3 * b + 1] = S->p[i].dLdt.y; out_dLdt[3 * b + 2] = S->p[i].dLdt.z;
        }
    }
    delete S;
    return 0;
}

// Algorithmic FP64 operation count (+ - * / sqrt) per system for n_steps
// steps, from the FlopCount instantiation of the same restated step.
int oracle_flops(const pos_ensemble_desc *d, int64_t n_steps, double *out_flops, int64_t *out_steps_done) {
    oracle::System<oracle::FlopCount> *S = new oracle::System<oracle::FlopCount>();
    for (int s = 0; s < d->n_systems; ++s) {
        std::string err;
        if (!oracle::load_system(d, s, *S, &err)) { g_err = err; delete S; return -1; }
        oracle::start_system(*S, d, s);
        oracle::FlopCount::count = 0;
        int64_t k = 0;
        while (k < n_steps) {
            ++k;
            if (!oracle::iterate(*S)) break;
        }
        out_flops[s] = (double)oracle::FlopCount::count;
        if (out_steps_done) out_steps_done[s] = k;
    }
    delete S;
    return 0;
}

}  // extern "C"

extern "C" {

// Tangent-linear (variational) integration: the Dual<ORACLE_TANGENT_DIRS>
// instantiation of the restated step.  Direction k seeds the initial
// position, velocity and spin angular momentum with d_x/d_v/d_L[k] ([3*nb]
// each, may be NULL = 0) and the parameters sigma (scale * dissipation
// factor), tidal k2 and fluid k2 with d_sigma/d_k2/d_k2f[k] ([nb], may be
// NULL).  out_* receive, per direction, the tangents of the final position,
// velocity and angular momentum ([n_dirs][3*nb]); *out (optional) the final
// values.  Runs every system for n_steps steps on n_threads threads.
constexpr int ORACLE_TANGENT_DIRS = 8;
int oracle_tangent_dirs(void) { return ORACLE_TANGENT_DIRS; }

int oracle_tangent(const pos_ensemble_desc *d, int64_t n_steps, int32_t n_threads, int32_t n_dirs,
                   const double *d_x, const double *d_v, const double *d_L, const double *d_sigma,
                   const double *d_k2, const double *d_k2f, pos_state *out, double *out_x, double *out_v,
                   double *out_L) {
    using R = oracle::Dual<ORACLE_TANGENT_DIRS>;
    if (n_dirs < 1 || n_dirs > ORACLE_TANGENT_DIRS) { g_err = "n_dirs out of range"; return -1; }
    const int ns = d->n_systems;
    const int nb = d->body_offset[ns];
    std::atomic<int> next{0};
    std::atomic<int> failed{-1};
    std::string first_err;
    auto worker = [&]() {
        oracle::System<R> *S = new oracle::System<R>();
        for (;;) {
            const int s = next.fetch_add(1);
            if (s >= ns) break;
            std::string err;
            if (!oracle::load_system(d, s, *S, &err)) {
                int expected = -1;
                if (failed.compare_exchange_strong(expected, s)) first_err = err;
                continue;
            }
            const int b0 = d->body_offset[s];
            for (int i = 0; i < S->n; ++i) {
                const int b = b0 + i;
                auto &P = S->p[i];
                for (int k = 0; k < n_dirs; ++k) {
                    const size_t o3 = (size_t)k * 3 * nb + 3 * b, o1 = (size_t)k * nb + b;
                    if (d_x) { P.x.x.d[k] = d_x[o3]; P.x.y.d[k] = d_x[o3 + 1]; P.x.z.d[k] = d_x[o3 + 2]; }
                    if (d_v) { P.v.x.d[k] = d_v[o3]; P.v.y.d[k] = d_v[o3 + 1]; P.v.z.d[k] = d_v[o3 + 2]; }
                    if (d_L) { P.L.x.d[k] = d_L[o3]; P.L.y.d[k] = d_L[o3 + 1]; P.L.z.d[k] = d_L[o3 + 2]; }
                    if (d_sigma) P.sigma.d[k] = d_sigma[o1];
                    if (d_k2) P.k2.d[k] = d_k2[o1];
                    if (d_k2f) P.k2f.d[k] = d_k2f[o1];
                }
            }
            oracle::start_system(*S, d, s);
            int64_t k = 0;
            while (k < n_steps) {
                ++k;
                if (!oracle::iterate(*S)) break;
            }
            if (out) store_state(*S, b0, s, out);
            for (int i = 0; i < S->n; ++i) {
                const int b = b0 + i;
                const auto &P = S->p[i];
                for (int q = 0; q < n_dirs; ++q) {
                    const size_t o3 = (size_t)q * 3 * nb + 3 * b;
                    if (out_x) { out_x[o3] = P.x.x.d[q]; out_x[o3 + 1] = P.x.y.d[q]; out_x[o3 + 2] = P.x.z.d[q]; }
                    if (out_v) { out_v[o3] = P.v.x.d[q]; out_v[o3 + 1] = P.v.y.d[q]; out_v[o3 + 2] = P.v.z.d[q]; }
                    if (out_L) { out_L[o3] = P.L.x.d[q]; out_L[o3 + 1] = P.L.y.d[q]; out_L[o3 + 2] = P.L.z.d[q]; }
                }
            }
        }
        delete S;
    };
    const int nt = n_threads < 1 ? 1 : n_threads;
    std::vector<std::thread> pool;
    for (int i = 0; i < nt; ++i) pool.emplace_back(worker);
    for (auto &t : pool) t.join();
    if (failed.load() >= 0) {
        g_err = "system " + std::to_string(failed.load()) + ": " + first_err;
        return -1;
    }
    return 0;
}

}  // extern "C"
