// posidonius_hip.hip -- kernels and the C ABI (include/posidonius_hip.h).
//
// Device layout: systems are reordered into "device order" (sorted by body
// count, then by input index) so that the systems of one wavefront are
// contiguous in memory.  Per-body vectors are stored component-major
// ([3][n_bodies]), per-system scalars in plain arrays.  The host keeps the
// permutation and applies it in pos_get_state / pos_set_state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "whfast_step.hpp"
#include "adjoint_kernels.hpp"

#ifndef POS_WAVES_PER_EU
#define POS_WAVES_PER_EU 2   // occupancy target of the step kernel (waves per SIMD)
#endif

using namespace posdev;

// ===========================================================================
// Kernels
// ===========================================================================
namespace {

template <int V = 0>
__device__ inline void load_lane(const DevEnsemble &E, WaveLds &lds, Lane &B, int wave) {
    const int lane = threadIdx.x;
    const LaneGroup lg = lane_group(E, wave, lane);
    const int k = lg.k;
    const bool gvalid = lg.gvalid;
    const int s = lg.sys;
    B.lane = lane;
    B.gbase = lg.gbase;
    B.k = k;
    B.sys = s;
    B.n = gvalid ? (E.body_offset[s + 1] - E.body_offset[s]) : 0;
    const bool valid = gvalid && k < B.n;
    const int b0 = E.body_offset[s];
    B.host = gvalid ? E.host[s] : 0;   // keeps gbase + host inside this wave's LDS rows
    B.body = b0 + (valid ? k : 0);
    B.coord = E.coord[s];
    const uint32_t eff = E.effects[s];
    const int b = B.body;
    const uint32_t roles = E.roles[b];
    B.evo = E.evo_type[b];
    uint32_t f = 0;
    if (valid) f |= F_VALID;
    if (valid && k == B.host) f |= F_HOST;
    if (eff & POS_EFFECT_TIDES) f |= F_TIDES;
    if (eff & POS_EFFECT_ROTATIONAL_FLATTENING) f |= F_RF;
    if (eff & POS_EFFECT_GENERAL_RELATIVITY) f |= F_GR;
    if (eff & POS_EFFECT_WIND) f |= F_WIND;
    if (eff & POS_EFFECT_EVOLUTION) f |= F_EVOLUTION;
    // spins are integrated with tides, flattening, evolution or Kidder GR
    // (whfast.rs:259-262); the other GR variants only add accelerations
    if ((eff & (POS_EFFECT_TIDES | POS_EFFECT_ROTATIONAL_FLATTENING | POS_EFFECT_EVOLUTION)) ||
        ((eff & POS_EFFECT_GENERAL_RELATIVITY) && eff_gr_variant(eff) == GRV_KIDDER))
        f |= F_SPIN;
    if (roles & POS_ROLE_WIND) f |= F_WIND_ROLE;
    if (eff & POS_EFFECT_DISK) f |= F_DISK;
    if (roles & POS_ROLE_DISK_ORBITING) f |= F_DISK_ROLE;
    if (evo_dynamical(B.evo)) f |= F_DYN;
    if (gvalid && evo_dynamical(E.evo_type[b0 + B.host])) f |= F_DYN_HOST;
    if ((B.evo == POS_EVO_LECONTE_CHABRIER2013 || B.evo == POS_EVO_LECONTE_CHABRIER2013_DYN) &&
        (roles & (POS_ROLE_TIDES_CENTRAL | POS_ROLE_TIDES_ORBITING)))
        f |= F_LC_LOVE;
    if (valid && k != B.host) {
        if ((f & F_TIDES) && (roles & POS_ROLE_TIDES_ORBITING)) f |= F_TIDES_ROLE;
        if ((f & F_RF) && (roles & POS_ROLE_FLAT_ORBITING)) f |= F_RF_ROLE;
        if ((f & F_GR) && (roles & POS_ROLE_GR_ORBITING)) f |= F_GR_ROLE;
    }
    if constexpr ((V & VAR_FAST) != 0) {   // the launch holds only systems of the fixed effect set
        const bool orb = valid && k != B.host;   // every other body has every orbiting role there
        f = (valid ? F_VALID : 0u) | ((valid && k == B.host) ? F_HOST : 0u) | F_TIDES | F_RF | F_GR | F_SPIN |
            (orb ? (F_TIDES_ROLE | F_RF_ROLE | F_GR_ROLE) : 0u);
        B.coord = (V & VAR_JACOBI) ? POS_COORD_JACOBI : POS_COORD_DEMOCRATIC_HELIOCENTRIC;
    }
    B.flags = f;
    B.dt = E.dt[s];
    B.hist_period = E.hist_period[s];
    B.time_limit = E.time_limit[s];
    B.roche_max2 = -1.;   // no pair: no distance is below it
    for (int j = k + 1; valid && j < B.n; ++j) {
        const double rr = E.roche[(size_t)b0 * POS_MAX_BODIES + (size_t)k * POS_MAX_BODIES + j];
        B.roche_max2 = fmax(B.roche_max2, rr * rr);
    }
    B.t = E.t[s];
    B.last_hist = E.last_hist[s];
    B.iteration = E.iteration[s];
    B.status = E.status[s];
    if (!valid) B.status |= POS_STATUS_COMPLETED;   // idle lanes never iterate
    B.tw = E.timestep_warning[s];
    B.seen_status = B.status;   // per system: what the group agrees on at load
    B.seen_tw = B.tw;
    B.left_index = E.left_index[b];
    B.hcount = E.hist ? E.hist_count[s] : 0;
    const int nb = E.n_bodies;
    B.m = E.mass[b];
    B.inv_m = 1. / B.m;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        B.x[q] = E.x[q * nb + b];
        B.v[q] = E.v[q * nb + b];
        B.L[q] = E.L[q * nb + b];
        B.ev[q] = E.err_v[q * nb + b];
        B.eL[q] = E.err_L[q * nb + b];
    }
    // persistent LDS area
    lds.m[lane] = valid ? B.m : 0.;
    {
        const double hm = E.mass[b0 + B.host];
        const double hmg = hm * K2, pmg = B.m * K2;
        const bool orbiting = valid && k != B.host;
        lds.kf[lane] = orbiting ? hmg * pmg / p2(hmg + pmg) : 0.;
        lds.inv_M[lane] = orbiting ? 1. / (hm + B.m) : 0.;
    }
    lds.inv_m[lane] = B.inv_m;
    lds.sigma[lane] = E.sigma[b];
    lds.k2[lane] = E.k2[b];
    lds.k2f[lane] = E.k2f[b];
    lds.R[lane] = E.radius[b];
    lds.I[lane] = E.inertia[b];
    double sp[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        sp[q] = E.spin[q * nb + b];
        lds.s[q][lane] = sp[q];
        lds.ag[q][lane] = E.acc[q * nb + b];
    }
    if (lane < NX) lds.xb[lane][WAVE] = 0.;   // the exchange area's zero column
    if (lane == 0) lds.m[WAVE] = 0.;
#pragma unroll
    for (int q = 0; q < 3; ++q) { B.ax[q] = 0.; B.av[q] = 0.; B.comx[q] = 0.; B.comv[q] = 0.; }
    B.com_m = 0.;
    __syncthreads();
}

__device__ inline void store_lane(const DevEnsemble &E, const WaveLds &lds, const Lane &B) {
    if (!B.valid()) return;
    // The indices pass through an empty asm so that the field addresses are
    // formed here again rather than kept from load_lane across the step loop
    // (where they would be spilled: 94 spilled registers against 35).
    int b = B.body, s = B.sys;
    asm volatile("" : "+v"(b), "+v"(s));
    const int nb = E.n_bodies;
    const int lane = B.lane;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        E.x[q * nb + b] = B.x[q];
        E.v[q * nb + b] = B.v[q];
        E.L[q * nb + b] = B.L[q];
        E.spin[q * nb + b] = lds.s[q][lane];
        E.err_v[q * nb + b] = B.ev[q];
        E.err_L[q * nb + b] = B.eL[q];
        E.acc[q * nb + b] = lds.ag[q][lane];
    }
    E.radius[b] = lds.R[lane];
    E.inertia[b] = lds.I[lane];
    E.left_index[b] = B.left_index;
    if (B.k == 0) {
        E.t[s] = B.t;
        E.last_hist[s] = B.last_hist;
        E.iteration[s] = B.iteration;
        E.status[s] = B.status;
        E.timestep_warning[s] = B.tw;
        if (E.hist) E.hist_count[s] = B.hcount;
    }
}

constexpr uint32_t FROZEN = POS_STATUS_COMPLETED | POS_STATUS_FATAL;

// K steps of every system of waves [wave_begin, wave_end), all of kernel
// variant V (GR implementation, with or without the extended effects); the
// grid may be smaller than the number of waves, in which case each workgroup
// walks the table with a grid stride (persistent launch).
template <int V>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(POS_WAVES_PER_EU)))
void whfast_steps_kernel(DevEnsemble E, int n_steps, int wave_begin, int wave_end) {
    __shared__ WaveLds lds;
    for (int wave = wave_begin + blockIdx.x; wave < wave_end; wave += gridDim.x) {
        Lane B;
        load_lane<V>(E, lds, B, wave);
        const TapeRef none{nullptr, 0};
        for (int step = 0; step < n_steps; ++step) {
            // uniform within the group, not the wavefront: the other groups keep
            // calling wsync(), which is correct only because the workgroup is
            // this one wavefront (__launch_bounds__(64), launched with 64 threads)
            if (B.status & FROZEN) break;
            iterate<false, V>(E, lds, B, none, 0);
        }
        store_lane(E, lds, B);
        __syncthreads();
    }
}

// The same K steps over waves [wave_begin, wave_end) of variant V (the
// adjoint admits the generic Kidder variant and the fixed-effect-set one),
// also writing the step tape of the adjoint (one row per step plus the final
// state).
template <int V>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(POS_WAVES_PER_EU)))
void whfast_record_kernel(DevEnsemble E, int n_steps, TapeRef T, int wave_begin, int wave_end) {
    __shared__ WaveLds lds;
    for (int wave = wave_begin + blockIdx.x; wave < wave_end; wave += gridDim.x) {
        Lane B;
        load_lane<V>(E, lds, B, wave);
        for (int step = 0; step < n_steps; ++step) {
            // uniform within the group, not the wavefront: the other groups keep
            // calling wsync(), which is correct only because the workgroup is
            // this one wavefront (__launch_bounds__(64), launched with 64 threads)
            if (B.status & FROZEN) break;
            iterate<true, V>(E, lds, B, T, step);
        }
        store_lane(E, lds, B);
        __syncthreads();
    }
}

// Reverse sweep over one recorded segment, last step first.  ALT: the
// ensemble holds systems in Jacobi or WHDS coordinates; EVO: systems with
// evolution tables.  Without them the reverse steps of the other coordinate
// types and the per-step tape reads of the radii are compiled out, which
// keeps the heliocentric sweep's register allocation as it is alone.
template <bool ALT, bool EVO>
__global__ __launch_bounds__(64) void adjoint_segment_kernel(AdjArgs Ar) {
    __shared__ AdjLds L;
    for (int wave = blockIdx.x; wave < Ar.E.n_waves; wave += gridDim.x) {
        AdjLane A;
        adjoint_load(Ar, L, A, wave);
        for (int j = Ar.seg_len - 1; j >= 0; --j) {
            // uniform within the group, not the wavefront: correct only because
            // the workgroup is this one wavefront (see whfast_steps_kernel)
            if (Ar.seg_start + j >= A.done) continue;
            if (ALT && A.jacobi) adjoint_step_jacobi<EVO>(Ar, L, A, j);
            else if (ALT && A.whds) adjoint_step<EVO, true>(Ar, L, A, j);
            else adjoint_step<EVO, false>(Ar, L, A, j);
        }
        adjoint_store(Ar, A);
        __syncthreads();
    }
}

// initialize_physical_values (whfast.rs:226-233) for systems at t == 0:
// evolution at t, then spin = L / I.
__global__ __launch_bounds__(64) void init_kernel(DevEnsemble E) {
    __shared__ WaveLds lds;
    const int wave = blockIdx.x;
    if (wave >= E.n_waves) return;
    Lane B;
    load_lane(E, lds, B, wave);
    if (B.t == 0.) {
        if (B.flags & F_EVOLUTION) evolve_body(E, lds, B, B.t);
        own_spin(lds, B);
        if (B.flags & F_EVOLUTION) evolve_spin_dependent(E, lds, B, B.t);
        merge_status(lds, B);
        store_lane(E, lds, B);
    }
}

// Operator-level evaluations (no state change).
__global__ __launch_bounds__(64) void gravity_kernel(DevEnsemble E, double *out_acc, uint32_t *out_status) {
    __shared__ WaveLds lds;
    const int wave = blockIdx.x;
    if (wave >= E.n_waves) return;
    Lane B;
    load_lane(E, lds, B, wave);
    gravity(E, lds, B);
    merge_status(lds, B);
    const int nb = E.n_bodies;
    if (B.valid()) {
#pragma unroll
        for (int q = 0; q < 3; ++q) out_acc[q * nb + B.body] = lds.ag[q][B.lane];
        if (B.k == 0) out_status[B.sys] = B.status;
    }
}

template <int V>
__global__ __launch_bounds__(64) void effects_kernel(DevEnsemble E, double *out_acc, double *out_dLdt, int wave_begin,
                                                     int wave_end) {
    __shared__ WaveLds lds;
    const int wave = wave_begin + blockIdx.x;
    if (wave >= wave_end) return;
    Lane B;
    load_lane<V>(E, lds, B, wave);
    double a[3], dL[3];
    PairGeo unused;
    additional_effects<V>(E, lds, B, false, a, dL, unused);
    const int nb = E.n_bodies;
    if (B.valid()) {
#pragma unroll
        for (int q = 0; q < 3; ++q) { out_acc[q * nb + B.body] = a[q]; out_dLdt[q * nb + B.body] = dL[q]; }
    }
}

// Per-body field from device order to the caller's order: component q of
// device body d (stored q * n + d) goes to caller body perm[d], comps values
// per body.  perm is a permutation of 0..n-1, so no two threads write the
// same element.
__global__ __launch_bounds__(256) void unpermute_kernel(const double *__restrict__ src, const int *__restrict__ perm, int n,
                                                        int comps, double *__restrict__ dst) {
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= n) return;
    const size_t c = (size_t)perm[d] * comps;
    for (int q = 0; q < comps; ++q) dst[c + q] = src[(size_t)q * n + d];
}

// The inverse of unpermute_kernel: caller order (comps per body) to device
// order (component q of device body d at q * n + d).
__global__ __launch_bounds__(256) void permute_kernel(const double *__restrict__ src, const int *__restrict__ perm, int n,
                                                      int comps, double *__restrict__ dst) {
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= n) return;
    const size_t c = (size_t)perm[d] * comps;
    for (int q = 0; q < comps; ++q) dst[(size_t)q * n + d] = src[c + q];
}

}  // namespace

// Scratch of pos_additional_effects_noalloc: the device-order accelerations
// and spin derivatives of one evaluation, component-major, before they are put
// into the caller's order.  Host code never passes this array's name to a
// kernel (on the host it is a stand-in, not a device address); it uses the
// address hipGetSymbolAddress gives per device (noalloc_scratch()).
static __device__ __align__(256) double scratch[2 * 3 * (size_t)POS_NOALLOC_MAX_BODIES];

// ===========================================================================
// Host side
// ===========================================================================
namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                 \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(POS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T> struct DBuf {
    T *p = nullptr;
    size_t n = 0;
};

}  // namespace

struct pos_ens {
    int device = 0;
    hipStream_t stream = nullptr;
    pos_options opts{};
    int n_systems = 0, n_bodies = 0;
    // device order <-> input order
    std::vector<int> sys_perm;          // device system d -> input system
    std::vector<int> body_perm;         // device body db -> input body
    std::vector<int> body_offset_dev;   // device CSR
    std::vector<int> body_offset_in;    // input CSR
    std::vector<int> n_per_sys_dev;
    // launch tables per mapping
    std::vector<int> wave_first, wave_count, wave_group, wave_sys;
    int variant_waves[N_VARIANTS][2] = {};   // [kernel variant] -> wave range [begin, end)
    std::vector<int> variant_dev;           // kernel variant of each device system
    int *d_wave_first = nullptr, *d_wave_count = nullptr, *d_wave_group = nullptr;   // capacity n_systems
    int *d_wave_sys = nullptr;                                                        // capacity n_systems
    double *d_op_a = nullptr, *d_op_l = nullptr;   // operator outputs, [3][n_bodies]
    uint32_t *d_op_st = nullptr;                    // operator status, [n_systems]
    double *d_unperm = nullptr;                     // per-body field in caller order, [n_bodies][3]
    int *d_body_perm = nullptr;                     // body_perm on the device
    int steps_per_launch = 1;
    int grid = 0;
    // device allocations
    std::vector<void *> allocs;
    DevEnsemble E{};
    // graph of graph_launches chunk launches
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_launches = 0;
    // per system / per body kinds in device order (adjoint support checks)
    std::vector<uint32_t> eff_dev, roles_dev;
    std::vector<int> coord_dev, evo_dev, host_dev;
    std::vector<double> mass_dev, sigma_dev, dt_dev;   // constants the history records carry
    int hist_cap = 0;
    pos_launch_stats stats{};
    hipEvent_t op_ev[2] = {nullptr, nullptr};   // around the kernels of an operator call
    // discrete adjoint: checkpoints of the recorded run and the segment tape
    struct Adjoint {
        bool recorded = false;
        int64_t n_steps = 0;
        int seg = 0, n_seg = 0;
        size_t state_bytes = 0;
        char *ckpt = nullptr;              // (n_seg + 1) states: segment starts, then the final state
        double *tape[2] = {nullptr, nullptr};   // two segment tapes, (seg + 1) * TAPE_FIELDS * n_bodies each
        hipStream_t sweep = nullptr;            // stream of the reverse sweeps
        hipEvent_t recorded_ev[2] = {nullptr, nullptr}, swept_ev[2] = {nullptr, nullptr};
        long long *iter0 = nullptr, *steps_done = nullptr;
        double *lx = nullptr, *lv = nullptr, *lL = nullptr, *gs = nullptr, *gk2 = nullptr, *gkf = nullptr;
        std::vector<void *> allocs;
    } adj;
};

namespace {

template <class T> int dalloc(pos_ens *h, T **p, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return fail(POS_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = (T *)q;
    return POS_OK;
}

template <class T> int upload(pos_ens *h, T *dst, const std::vector<T> &src) {
    if (src.empty()) return POS_OK;
    HIP_TRY(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, h->stream));
    return POS_OK;
}

template <class T> int download(pos_ens *h, std::vector<T> &dst, const T *src, size_t n) {
    dst.resize(n);
    if (n == 0) return POS_OK;
    HIP_TRY(hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, h->stream));
    return POS_OK;
}


// Copy a per-body field (comps components, device order) to the caller's
// host array in the caller's body order: reordered on the device, then one
// copy.  Synchronous.
int body_staging(pos_ens *h) {
    int rc;
    if (!h->d_unperm) {
        const int nb = h->n_bodies;
        if ((rc = dalloc(h, &h->d_unperm, 3 * (size_t)nb)) || (rc = dalloc(h, &h->d_body_perm, nb)) ||
            (rc = upload(h, h->d_body_perm, h->body_perm)))
            return rc;
    }
    return POS_OK;
}

int download_bodies(pos_ens *h, double *dst, const double *src, int comps) {
    const int nb = h->n_bodies;
    int rc;
    if ((rc = body_staging(h))) return rc;
    if (comps < 1 || comps > 3) return fail(POS_E_INVALID, "internal error: field width");
    hipLaunchKernelGGL(unpermute_kernel, dim3((nb + 255) / 256), dim3(256), 0, h->stream, src, h->d_body_perm, nb, comps,
                       h->d_unperm);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(dst, h->d_unperm, (size_t)comps * nb * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return POS_OK;
}

// The inverse of download_bodies: the caller's host array (caller body
// order) copied once, then put into device order on the device.  Synchronous.
int upload_bodies(pos_ens *h, double *dst, const double *src, int comps) {
    const int nb = h->n_bodies;
    int rc;
    if ((rc = body_staging(h))) return rc;
    if (comps < 1 || comps > 3) return fail(POS_E_INVALID, "internal error: field width");
    HIP_TRY(hipMemcpyAsync(h->d_unperm, src, (size_t)comps * nb * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(permute_kernel, dim3((nb + 255) / 256), dim3(256), 0, h->stream, h->d_unperm, h->d_body_perm, nb,
                       comps, dst);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(h->stream));
    return POS_OK;
}

int next_pow2(int n) {
    int g = 1;
    while (g < n) g <<= 1;
    return g;
}

// Lanes of the group of a system of n bodies: the set group size (at least
// n), exactly n (group_size < 0), or the next power of two (0).
int group_lanes(int group_size, int n) {
    if (group_size > 0) return std::max(group_size, n);
    if (group_size < 0) return n;
    return next_pow2(n);
}

// Build the wave table for the current mapping; systems are in device order
// (sorted by body count).
int build_launch(pos_ens *h, const pos_options &o) {
    h->wave_first.clear();
    h->wave_count.clear();
    h->wave_group.clear();
    h->wave_sys.clear();
    const int ns = h->n_systems;
    if (o.mapping == POS_MAPPING_PACKED && o.group_size == -2) {
        // exactly N lanes per system, systems of one kernel variant packed
        // into 64-lane wavefronts best-fit decreasing: largest systems first,
        // each into the open wavefront with the least room that holds it
        // (the device order is by variant, then by size).  A wavefront costs
        // about what its largest system costs, so the small systems fill the
        // lanes the large ones leave idle.
        int s = 0;
        while (s < ns) {
            int e = s;
            while (e < ns && h->variant_dev[e] == h->variant_dev[s]) ++e;
            std::vector<std::vector<int>> bins;
            std::vector<int> by_room[65];   // open wavefronts by free lanes
            for (int d = e - 1; d >= s; --d) {
                const int n = std::min(h->n_per_sys_dev[d], 64);
                int r = n;
                while (r <= 64 && by_room[r].empty()) ++r;
                int b;
                if (r > 64) {
                    b = (int)bins.size();
                    bins.emplace_back();
                    r = 64;
                } else {
                    b = by_room[r].back();
                    by_room[r].pop_back();
                }
                bins[b].push_back(d);
                by_room[r - n].push_back(b);
            }
            // lightest wavefronts first here: the table is reversed below
            for (size_t b = bins.size(); b-- > 0;) {
                h->wave_first.push_back((int)h->wave_sys.size());
                h->wave_count.push_back((int)bins[b].size());
                h->wave_group.push_back(0);
                h->wave_sys.insert(h->wave_sys.end(), bins[b].begin(), bins[b].end());
            }
            s = e;
        }
        h->steps_per_launch = o.steps_per_launch > 0 ? o.steps_per_launch : 64;
    } else if (o.mapping == POS_MAPPING_PLAIN) {
        for (int s = 0; s < ns; ++s) {
            h->wave_first.push_back(s);
            h->wave_count.push_back(1);
            h->wave_group.push_back(64);
        }
        h->steps_per_launch = 1;
    } else {
        int s = 0;
        while (s < ns) {
            const int n = h->n_per_sys_dev[s];
            int G = group_lanes(o.group_size, n);
            if (G > 64) G = 64;
            const int per_wave = 64 / G;
            // systems with the same group size and GR variant share waves
            int e = s;
            while (e < ns) {
                const int ne = h->n_per_sys_dev[e];
                const int Ge = group_lanes(o.group_size, ne);
                if (std::min(Ge, 64) != G || h->variant_dev[e] != h->variant_dev[s]) break;
                ++e;
            }
            for (int f = s; f < e; f += per_wave) {
                h->wave_first.push_back(f);
                h->wave_count.push_back(std::min(per_wave, e - f));
                h->wave_group.push_back(G);
            }
            s = e;
        }
        h->steps_per_launch = o.steps_per_launch > 0 ? o.steps_per_launch : 64;
    }
    // Heaviest wavefronts (largest systems) first: workgroups are dispatched
    // in index order, so the light ones fill the tail of the launch.
    std::reverse(h->wave_first.begin(), h->wave_first.end());
    std::reverse(h->wave_count.begin(), h->wave_count.end());
    std::reverse(h->wave_group.begin(), h->wave_group.end());
    const int nw = (int)h->wave_first.size();
    // device order groups the variants, so each one covers a contiguous range
    for (int v = 0; v < N_VARIANTS; ++v) {
        h->variant_waves[v][0] = nw;
        h->variant_waves[v][1] = 0;
    }
    for (int w = 0; w < nw; ++w) {
        const int first_sys = h->wave_group[w] > 0 ? h->wave_first[w] : h->wave_sys[h->wave_first[w]];
        const int v = h->variant_dev[first_sys];
        h->variant_waves[v][0] = std::min(h->variant_waves[v][0], w);
        h->variant_waves[v][1] = std::max(h->variant_waves[v][1], w + 1);
    }
    int rc;
    if (!h->d_wave_first) {
        if ((rc = dalloc(h, &h->d_wave_first, ns)) || (rc = dalloc(h, &h->d_wave_count, ns)) ||
            (rc = dalloc(h, &h->d_wave_group, ns)) || (rc = dalloc(h, &h->d_wave_sys, ns)))
            return rc;
    }
    if ((rc = upload(h, h->d_wave_first, h->wave_first)) || (rc = upload(h, h->d_wave_count, h->wave_count)) ||
        (rc = upload(h, h->d_wave_group, h->wave_group)) ||
        (!h->wave_sys.empty() && (rc = upload(h, h->d_wave_sys, h->wave_sys))))
        return rc;
    h->E.wave_sys = h->d_wave_sys;
    h->E.wave_first = h->d_wave_first;
    h->E.wave_count = h->d_wave_count;
    h->E.wave_group = h->d_wave_group;
    h->E.n_waves = nw;
    h->grid = o.reserved[0] > 0 ? std::min(o.reserved[0], nw) : nw;   // reserved[0]: persistent grid size
    h->opts = o;
    h->opts.steps_per_launch = h->steps_per_launch;
    if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
    if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
    h->graph_launches = 0;
    return POS_OK;
}

// consider_effects after Universe::new switched off what no body carries
// (disable_unnecessary_effects, universe.rs:699-764): tides, flattening, GR and
// the disk without a central body, wind without a wind body, evolution without
// an evolving body; GR also when its central body's implementation is Disabled.
uint32_t effective_effects(const pos_ensemble_desc *d, int s) {
    uint32_t eff = d->effects[s];
    if (d->gr_implementation[s] == POS_GR_DISABLED) eff &= ~POS_EFFECT_GENERAL_RELATIVITY;
    uint32_t any = 0;
    bool evolving = false;
    for (int b = d->body_offset[s]; b < d->body_offset[s + 1]; ++b) {
        any |= d->roles[b];
        if (d->evolution_type && d->evolution_type[b] != POS_EVO_NON_EVOLVING) evolving = true;
    }
    if (!(any & POS_ROLE_TIDES_CENTRAL)) eff &= ~POS_EFFECT_TIDES;
    if (!(any & POS_ROLE_FLAT_CENTRAL)) eff &= ~POS_EFFECT_ROTATIONAL_FLATTENING;
    if (!(any & POS_ROLE_GR_CENTRAL)) eff &= ~POS_EFFECT_GENERAL_RELATIVITY;
    if (!(any & POS_ROLE_DISK_CENTRAL)) eff &= ~POS_EFFECT_DISK;
    if (!(any & POS_ROLE_WIND)) eff &= ~POS_EFFECT_WIND;
    if (!evolving) eff &= ~POS_EFFECT_EVOLUTION;
    return eff;
}

// Every body of system s but its host is an orbiting body of tides,
// flattening and general relativity.
bool every_body_orbits(const pos_ensemble_desc *d, int s, int host) {
    constexpr uint32_t ALL = POS_ROLE_TIDES_ORBITING | POS_ROLE_FLAT_ORBITING | POS_ROLE_GR_ORBITING;
    const int b0 = d->body_offset[s], n = d->body_offset[s + 1] - b0;
    for (int i = 0; i < n; ++i)
        if (i != host && (d->roles[b0 + i] & ALL) != ALL) return false;
    return true;
}

// Kernel variant a system's steps are compiled for: its GR implementation
// (GRV_KIDDER also when GR is off) plus VAR_EXT with the disk effect or an
// evolving-dissipation host.
int system_variant(const pos_ensemble_desc *d, int s, int host) {
    const uint32_t effects = effective_effects(d, s);
    const int implementation = d->gr_implementation[s];
    int v = GRV_KIDDER;
    if (effects & POS_EFFECT_GENERAL_RELATIVITY) {
        if (implementation == POS_GR_ANDERSON1975) v = GRV_ANDERSON;
        if (implementation == POS_GR_NEWHALL1983) v = GRV_NEWHALL;
    }
    bool dyn = false;   // an evolving-dissipation body (host or orbiting)
    if (d->evolution_type)
        for (int b = d->body_offset[s]; b < d->body_offset[s + 1]; ++b) dyn = dyn || evo_dynamical(d->evolution_type[b]);
    if ((effects & POS_EFFECT_DISK) || dyn) v |= VAR_EXT;
    constexpr uint32_t FAST_EFFECTS = POS_EFFECT_TIDES | POS_EFFECT_ROTATIONAL_FLATTENING | POS_EFFECT_GENERAL_RELATIVITY;
    if (v == GRV_KIDDER && effects == FAST_EFFECTS && implementation == POS_GR_KIDDER1995 && every_body_orbits(d, s, host)) {
        if (d->coordinates[s] == POS_COORD_DEMOCRATIC_HELIOCENTRIC) v |= VAR_FAST;
        if (d->coordinates[s] == POS_COORD_JACOBI) v |= VAR_FAST | VAR_JACOBI;
    }
    return v;
}

// most massive body of system s (strictly greater keeps the first maximum)
int most_massive(const pos_ensemble_desc *d, int s) {
    const int b0 = d->body_offset[s], n = d->body_offset[s + 1] - b0;
    int mm = 0;
    double mmax = 0.;
    for (int i = 0; i < n; ++i)
        if (d->mass[b0 + i] > mmax) { mmax = d->mass[b0 + i]; mm = i; }
    return mm;
}

int validate(const pos_ensemble_desc *d) {
    if (!d) return fail(POS_E_INVALID, "desc is NULL");
    if (d->n_systems <= 0) return fail(POS_E_INVALID, "n_systems must be positive");
    const void *required[] = {d->body_offset, d->coordinates, d->effects, d->gr_implementation, d->time_step,
                              d->time_limit, d->mass, d->radius, d->radius_of_gyration_2, d->position,
                              d->velocity, d->angular_momentum, d->roles, d->tides_dissipation_factor,
                              d->tides_dissipation_factor_scale, d->tides_love_number, d->flattening_love_number};
    for (const void *p : required)
        if (!p) return fail(POS_E_INVALID, "a required desc array is NULL");
    if (d->body_offset[0] != 0) return fail(POS_E_INVALID, "body_offset[0] must be 0");
    for (int s = 0; s < d->n_systems; ++s) {
        const int b0 = d->body_offset[s], n = d->body_offset[s + 1] - b0;
        const std::string where = "system " + std::to_string(s) + ": ";
        if (n < 2 || n > POS_MAX_BODIES)
            return fail(POS_E_LIMIT, where + "body count " + std::to_string(n) + " outside [2, " + std::to_string(POS_MAX_BODIES) + "]");
        const uint32_t eff = d->effects[s];
        if ((eff & POS_EFFECT_DISK) && !d->disk_properties)
            return fail(POS_E_INVALID, where + "the disk effect needs disk_properties");
        if (d->coordinates[s] < 0 || d->coordinates[s] > 2) return fail(POS_E_INVALID, where + "bad coordinates");
        if ((eff & POS_EFFECT_GENERAL_RELATIVITY) &&
            (d->gr_implementation[s] < POS_GR_KIDDER1995 || d->gr_implementation[s] > POS_GR_DISABLED))
            return fail(POS_E_INVALID, where + "bad general relativity implementation");
        if (!(d->time_step[s] > 0.)) return fail(POS_E_INVALID, where + "time_step must be positive");
        int mm = -1;
        double mmax = 0.;
        for (int i = 0; i < n; ++i)
            if (d->mass[b0 + i] > mmax) { mmax = d->mass[b0 + i]; mm = i; }
        if (mm < 0) return fail(POS_E_INVALID, where + "no body has a positive mass");
        // the central body of each effect (universe.rs:699-764, 900-1006): at
        // most one per effect; an effect without one is switched off
        // (effective_effects); the reference panics on the conditions below
        const uint32_t central[4] = {POS_ROLE_TIDES_CENTRAL, POS_ROLE_FLAT_CENTRAL, POS_ROLE_GR_CENTRAL, POS_ROLE_DISK_CENTRAL};
        const char *names[4] = {"tidal", "rotational flattening", "general relativity", "disk"};
        int host_of[4] = {-1, -1, -1, -1};
        for (int e = 0; e < 4; ++e)
            for (int i = 0; i < n; ++i)
                if (d->roles[b0 + i] & central[e]) {
                    if (host_of[e] >= 0) return fail(POS_E_INVALID, where + "only one central body is allowed for " + names[e] + " effects");
                    host_of[e] = i;
                }
        const uint32_t on = effective_effects(d, s);
        const bool tides = on & POS_EFFECT_TIDES, rf = on & POS_EFFECT_ROTATIONAL_FLATTENING;
        if (tides && rf && host_of[0] != host_of[1])
            return fail(POS_E_INVALID, where + "the central body for tidal & rotational flattening effects needs to be the same");
        if ((on & POS_EFFECT_GENERAL_RELATIVITY) && host_of[2] != mm)
            return fail(POS_E_INVALID, where + "the central body for general relativity should be the most massive one");
        // a tidal / flattening / disk host other than the most massive body
        // (hosts.most_massive.all false, universe.rs:386-419) is not built
        if ((tides || rf) && host_of[tides ? 0 : 1] != mm)
            return fail(POS_E_UNSUPPORTED, where + "the tidal and flattening host must be the most massive body");
        if ((on & POS_EFFECT_DISK) && host_of[3] != mm)
            return fail(POS_E_UNSUPPORTED, where + "the disk host must be the most massive body");
        // Anderson 1975 / Newhall 1983 read the GR coordinates of every body,
        // which the reference leaves stale for a body whose GR is Disabled
        if ((on & POS_EFFECT_GENERAL_RELATIVITY) && d->gr_implementation[s] != POS_GR_KIDDER1995)
            for (int i = 0; i < n; ++i)
                if (i != mm && !(d->roles[b0 + i] & POS_ROLE_GR_ORBITING))
                    return fail(POS_E_UNSUPPORTED, where + "Anderson1975 and Newhall1983 need every other body to be a "
                                                           "general relativity orbiting body");
        const double last = d->last_historic_snapshot_time ? d->last_historic_snapshot_time[s] : -1.;
        if (last >= 0. && d->heliocentric_velocity) {
            const double *hv = d->heliocentric_velocity + 3 * (b0 + mm);
            if (hv[0] != 0. || hv[1] != 0. || hv[2] != 0.)
                return fail(POS_E_UNSUPPORTED, where + "a resumed system must have a zero host heliocentric velocity");
        }
        if (d->evolution_type) {
            for (int i = 0; i < n; ++i) {
                const int et = d->evolution_type[b0 + i];
                if (et < POS_EVO_NON_EVOLVING || et > POS_EVO_GALLET_BOLMONT2017)
                    return fail(POS_E_INVALID, where + "bad evolution type");
                if (evo_dynamical(et) && !d->evolver_inverse_tidal_q_factor)
                    return fail(POS_E_INVALID, where + "evolving dissipation needs evolver_inverse_tidal_q_factor");
                if (et != POS_EVO_NON_EVOLVING &&
                    (!d->evolver_offset || d->evolver_offset[b0 + i + 1] - d->evolver_offset[b0 + i] <= 0))
                    return fail(POS_E_INVALID, where + "evolving body without a table");
                if (et != POS_EVO_NON_EVOLVING && d->evolver_left_index && d->evolver_left_index[b0 + i] < 0)
                    return fail(POS_E_INVALID, where + "negative evolver left_index");
            }
        }
    }
    return POS_OK;
}

double host_roche(double ma, double mb, double ra, double rb) {
    // universe.rs:186-193 (setup only, evaluated with the host libm)
    if (ma > mb) return (rb / 0.462) * std::pow(ma / mb, 1. / 3.);
    return (ra / 0.462) * std::pow(mb / ma, 1. / 3.);
}

void adjoint_release(pos_ens *h) {
    pos_ens::Adjoint &A = h->adj;
    for (void *p : A.allocs) (void)hipFree(p);
    for (int i = 0; i < 2; ++i) {
        if (A.recorded_ev[i]) (void)hipEventDestroy(A.recorded_ev[i]);
        if (A.swept_ev[i]) (void)hipEventDestroy(A.swept_ev[i]);
    }
    if (A.sweep) (void)hipStreamDestroy(A.sweep);
    A = pos_ens::Adjoint();
}

// The mutable device state of an ensemble (what a checkpoint must hold).
std::vector<std::pair<void *, size_t>> adjoint_state_fields(pos_ens *h) {
    const size_t nb = h->n_bodies, ns = h->n_systems;
    DevEnsemble &E = h->E;
    const size_t v3 = 3 * nb * sizeof(double), s1 = nb * sizeof(double);
    return {{E.x, v3}, {E.v, v3}, {E.L, v3}, {E.spin, v3}, {E.err_v, v3}, {E.err_L, v3}, {E.acc, v3},
            {E.radius, s1}, {E.rg2, s1}, {E.inertia, s1}, {E.left_index, nb * sizeof(int)},
            {E.t, ns * sizeof(double)}, {E.last_hist, ns * sizeof(double)}, {E.iteration, ns * sizeof(long long)},
            {E.status, ns * sizeof(uint32_t)}, {E.timestep_warning, ns * sizeof(int)}};
}

// Copy the mutable state into (save) or out of (restore) a checkpoint slot.
int adjoint_copy_state(pos_ens *h, char *slot, bool save) {
    size_t off = 0;
    for (const auto &f : adjoint_state_fields(h)) {
        if (save)
            HIP_TRY(hipMemcpyAsync(slot + off, f.first, f.second, hipMemcpyDeviceToDevice, h->stream));
        else
            HIP_TRY(hipMemcpyAsync(f.first, slot + off, f.second, hipMemcpyDeviceToDevice, h->stream));
        off += (f.second + 255) / 256 * 256;
    }
    return POS_OK;
}

constexpr int GRAPH_LAUNCHES = 4;   // kernel launches per captured graph

void drop_graph(pos_ens *h) {
    if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
    if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
    h->graph_launches = 0;
}

// The device view of the ensemble with historic snapshots switched off: steps
// that are not part of the caller's run (autotuner probes, the adjoint's
// segment recomputation) must not append records.
DevEnsemble without_history(const DevEnsemble &E) {
    DevEnsemble r = E;
    r.hist = nullptr;
    r.hist_count = nullptr;
    r.hist_cap = 0;
    return r;
}

// Detaches the history buffer from the handle for the lifetime of the guard
// (every launch made meanwhile, graph captures included, sees none) and
// reattaches it on release; a graph captured while detached is dropped then.
struct HistoryDetach {
    pos_ens *h;
    double *hist;
    int *count;
    int cap;
    bool active = true;
    explicit HistoryDetach(pos_ens *h_) : h(h_), hist(h_->E.hist), count(h_->E.hist_count), cap(h_->E.hist_cap) {
        h->E = without_history(h->E);
    }
    void release() {
        if (!active) return;
        active = false;
        h->E.hist = hist;
        h->E.hist_count = count;
        h->E.hist_cap = cap;
        if (hist) drop_graph(h);
    }
    ~HistoryDetach() { release(); }
};

// Two timing events, destroyed with the guard.
struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    hipError_t create() {
        for (auto &x : e) {
            const hipError_t r = hipEventCreate(&x);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
    ~EventPair() {
        for (auto &x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

template <int V> void launch_steps_variant(pos_ens *h, int grid, int n_steps, int w0, int w1) {
    hipLaunchKernelGGL(whfast_steps_kernel<V>, dim3(grid), dim3(64), 0, h->stream, h->E, n_steps, w0, w1);
}

// One launch per kernel variant present; the persistent grid size caps each.
// Returns the number of launches through *launches (optional).
int launch_steps(pos_ens *h, int n_steps, int *launches = nullptr) {
    if (launches) *launches = 0;
    for (int v = 0; v < N_VARIANTS; ++v) {
        const int w0 = h->variant_waves[v][0], w1 = h->variant_waves[v][1];
        if (w1 <= w0) continue;
        if (launches) ++*launches;
        const int grid = std::min(h->grid, w1 - w0);
        switch (v) {
            case 0: launch_steps_variant<0>(h, grid, n_steps, w0, w1); break;
            case 1: launch_steps_variant<1>(h, grid, n_steps, w0, w1); break;
            case 2: launch_steps_variant<2>(h, grid, n_steps, w0, w1); break;
            case 4: launch_steps_variant<4>(h, grid, n_steps, w0, w1); break;
            case 5: launch_steps_variant<5>(h, grid, n_steps, w0, w1); break;
            case 6: launch_steps_variant<6>(h, grid, n_steps, w0, w1); break;
            case 8: launch_steps_variant<8>(h, grid, n_steps, w0, w1); break;
            case 24: launch_steps_variant<24>(h, grid, n_steps, w0, w1); break;
            default: return fail(POS_E_INVALID, "internal error: unknown kernel variant");
        }
        HIP_TRY(hipGetLastError());
    }
    return POS_OK;
}

// GRAPH_LAUNCHES chunk launches of the current launch table, captured from
// the handle's stream (nothing executes) and instantiated once; rebuilt after
// anything that changes the kernels' arguments.
int ensure_graph(pos_ens *h) {
    if (h->graph_exec) return POS_OK;
    HIP_TRY(hipStreamSynchronize(h->stream));   // capture starts on an idle stream
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int crc = POS_OK, per = 0, total = 0;
    for (int i = 0; i < GRAPH_LAUNCHES && crc == POS_OK; ++i) {
        crc = launch_steps(h, h->steps_per_launch, &per);
        total += per;
    }
    HIP_TRY(hipStreamEndCapture(h->stream, &h->graph));
    if (crc) return crc;
    HIP_TRY(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
    h->graph_launches = total;
    return POS_OK;
}

// The device address of `scratch` on device `dev`, looked up once per device
// with that device current (each GPU holds its own copy of the array).
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[MAX_DEVICES] = {};

int noalloc_scratch(int dev, double **out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(POS_E_LIMIT, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(POS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double *)ptr;
    }
    *out = g_scratch_ptr[dev];
    return POS_OK;
}

// One launch of effects_kernel per kernel variant present, writing device-order
// results to d_a, d_l ([3][n_bodies] each).
int launch_effects(pos_ens *h, double *d_a, double *d_l) {
    for (int v = 0; v < N_VARIANTS; ++v) {
        const int w0 = h->variant_waves[v][0], w1 = h->variant_waves[v][1];
        if (w1 <= w0) continue;
        const dim3 g(w1 - w0);
        switch (v) {
            case 0: hipLaunchKernelGGL(effects_kernel<0>, g, dim3(64), 0, h->stream, h->E, d_a, d_l, w0, w1); break;
            case 1: hipLaunchKernelGGL(effects_kernel<1>, g, dim3(64), 0, h->stream, h->E, d_a, d_l, w0, w1); break;
            case 2: hipLaunchKernelGGL(effects_kernel<2>, g, dim3(64), 0, h->stream, h->E, d_a, d_l, w0, w1); break;
            case 4: hipLaunchKernelGGL(effects_kernel<4>, g, dim3(64), 0, h->stream, h->E, d_a, d_l, w0, w1); break;
            case 5: hipLaunchKernelGGL(effects_kernel<5>, g, dim3(64), 0, h->stream, h->E, d_a, d_l, w0, w1); break;
            case 6: hipLaunchKernelGGL(effects_kernel<6>, g, dim3(64), 0, h->stream, h->E, d_a, d_l, w0, w1); break;
            case 8: hipLaunchKernelGGL(effects_kernel<8>, g, dim3(64), 0, h->stream, h->E, d_a, d_l, w0, w1); break;
            case 24: hipLaunchKernelGGL(effects_kernel<24>, g, dim3(64), 0, h->stream, h->E, d_a, d_l, w0, w1); break;
            default: return fail(POS_E_INVALID, "internal error: unknown kernel variant");
        }
        HIP_TRY(hipGetLastError());
    }
    return POS_OK;
}

// Device time between the two operator events (both recorded and completed).
int record_operator_time(pos_ens *h) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->op_ev[0], h->op_ev[1]));
    h->stats.last_operator_kernel_ms = ms;
    return POS_OK;
}

}  // namespace

extern "C" {

int32_t pos_abi_version(void) { return POS_ABI_VERSION; }

const char *pos_last_error(void) { return g_last_error.c_str(); }

int32_t pos_create(const pos_ensemble_desc *d, const pos_options *opts, pos_ens **out) {
    if (!out) return fail(POS_E_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate(d);
    if (rc) return rc;
    pos_options o{};
    if (opts) o = *opts;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(POS_E_NO_DEVICE, "no HIP device");
    if (o.device < 0 || o.device >= ndev) return fail(POS_E_INVALID, "device ordinal out of range");
    HIP_TRY(hipSetDevice(o.device));

    pos_ens *h = new pos_ens();
    h->device = o.device;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(POS_E_HIP, "hipStreamCreate failed");
    }
    if (hipEventCreate(&h->op_ev[0]) != hipSuccess || hipEventCreate(&h->op_ev[1]) != hipSuccess) {
        pos_destroy(h);
        return fail(POS_E_HIP, "hipEventCreate failed");
    }
    const int ns = d->n_systems;
    const int nb = d->body_offset[ns];
    h->n_systems = ns;
    h->n_bodies = nb;
    h->body_offset_in.assign(d->body_offset, d->body_offset + ns + 1);

    // device order: by GR variant, then body count, then input order
    std::vector<int> variant_in(ns);
    for (int s = 0; s < ns; ++s) variant_in[s] = system_variant(d, s, most_massive(d, s));
    h->sys_perm.resize(ns);
    std::iota(h->sys_perm.begin(), h->sys_perm.end(), 0);
    std::stable_sort(h->sys_perm.begin(), h->sys_perm.end(), [&](int a, int b) {
        if (variant_in[a] != variant_in[b]) return variant_in[a] < variant_in[b];
        return (d->body_offset[a + 1] - d->body_offset[a]) < (d->body_offset[b + 1] - d->body_offset[b]);
    });
    h->variant_dev.resize(ns);
    for (int ds = 0; ds < ns; ++ds) h->variant_dev[ds] = variant_in[h->sys_perm[ds]];
    h->body_offset_dev.assign(ns + 1, 0);
    h->n_per_sys_dev.resize(ns);
    h->body_perm.resize(nb);
    for (int ds = 0; ds < ns; ++ds) {
        const int s = h->sys_perm[ds];
        const int b0 = d->body_offset[s], n = d->body_offset[s + 1] - b0;
        h->n_per_sys_dev[ds] = n;
        h->body_offset_dev[ds + 1] = h->body_offset_dev[ds] + n;
        for (int i = 0; i < n; ++i) h->body_perm[h->body_offset_dev[ds] + i] = b0 + i;
    }

    // ---- host-side arrays in device order ----
    std::vector<int> host_idx(ns), coord(ns), tw(ns, 0);
    std::vector<uint32_t> eff(ns), status(ns, 0);
    std::vector<double> dt(ns), tl(ns), t(ns), lh(ns), hp(ns);
    std::vector<long long> iter(ns);
    std::vector<double> disk(6 * (size_t)ns, 0.);
    for (int ds = 0; ds < ns; ++ds) {
        const int s = h->sys_perm[ds];
        const int b0 = d->body_offset[s], n = d->body_offset[s + 1] - b0;
        int mm = 0;
        double mmax = 0.;
        for (int i = 0; i < n; ++i)
            if (d->mass[b0 + i] > mmax) { mmax = d->mass[b0 + i]; mm = i; }
        host_idx[ds] = mm;
        coord[ds] = d->coordinates[s];
        const uint32_t e = effective_effects(d, s);
        eff[ds] = e | ((uint32_t)h->variant_dev[ds] << EFF_VAR_SHIFT);
        dt[ds] = d->time_step[s];
        tl[ds] = d->time_limit[s];
        t[ds] = d->current_time ? d->current_time[s] : 0.;
        lh[ds] = d->last_historic_snapshot_time ? d->last_historic_snapshot_time[s] : -1.;
        hp[ds] = d->historic_snapshot_period ? d->historic_snapshot_period[s] : std::numeric_limits<double>::infinity();
        iter[ds] = d->current_iteration ? (long long)d->current_iteration[s] : 0;
        if (d->disk_properties)
            for (int q = 0; q < 6; ++q) disk[6 * (size_t)ds + q] = d->disk_properties[6 * (size_t)s + q];
    }
    std::vector<double> x(3 * nb), v(3 * nb), L(3 * nb), sp(3 * nb), ev(3 * nb, 0.), eL(3 * nb, 0.), acc(3 * nb, 0.);
    std::vector<double> m(nb), R(nb), rg2(nb), I(nb), sigma(nb), k2(nb), k2f(nb), wk(nb, 0.), ws(nb, 0.);
    std::vector<double> sig_scale(nb), sig_factor(nb), zeros_nb(nb, 0.);
    std::vector<double> pair(nb, std::numeric_limits<double>::quiet_NaN());
    std::vector<uint32_t> roles(nb);
    std::vector<int> evo_type(nb, POS_EVO_NON_EVOLVING), evo_off(nb, 0), evo_len(nb, 0), left(nb, 0);
    std::vector<double> tab_t, tab_R, tab_rg2, tab_love, tab_itq;
    for (int db = 0; db < nb; ++db) {
        const int b = h->body_perm[db];
        m[db] = d->mass[b];
        R[db] = d->radius[b];
        rg2[db] = d->radius_of_gyration_2[b];
        I[db] = d->moment_of_inertia ? d->moment_of_inertia[b] : m[db] * rg2[db] * (R[db] * R[db]);
        sigma[db] = d->tides_dissipation_factor_scale[b] * d->tides_dissipation_factor[b];
        sig_scale[db] = d->tides_dissipation_factor_scale[b];
        sig_factor[db] = d->tides_dissipation_factor[b];
        if (d->pair_dissipation_factor) pair[db] = d->pair_dissipation_factor[b];
        k2[db] = d->tides_love_number[b];
        k2f[db] = d->flattening_love_number[b];
        if (d->wind_k_factor) wk[db] = d->wind_k_factor[b];
        if (d->wind_rotation_saturation) ws[db] = d->wind_rotation_saturation[b];
        roles[db] = d->roles[b];
        for (int q = 0; q < 3; ++q) {
            x[q * nb + db] = d->position[3 * b + q];
            v[q * nb + db] = d->velocity[3 * b + q];
            L[q * nb + db] = d->angular_momentum[3 * b + q];
            sp[q * nb + db] = d->spin ? d->spin[3 * b + q] : L[q * nb + db] / I[db];
            if (d->velocity_error) ev[q * nb + db] = d->velocity_error[3 * b + q];
            if (d->angular_momentum_error) eL[q * nb + db] = d->angular_momentum_error[3 * b + q];
        }
        if (d->evolution_type && d->evolution_type[b] != POS_EVO_NON_EVOLVING) {
            const int o0 = d->evolver_offset[b], len = d->evolver_offset[b + 1] - o0;
            evo_type[db] = d->evolution_type[b];
            evo_off[db] = (int)tab_t.size();
            evo_len[db] = len;
            left[db] = d->evolver_left_index ? d->evolver_left_index[b] : 0;
            tab_t.insert(tab_t.end(), d->evolver_time + o0, d->evolver_time + o0 + len);
            tab_R.insert(tab_R.end(), d->evolver_radius + o0, d->evolver_radius + o0 + len);
            tab_rg2.insert(tab_rg2.end(), d->evolver_radius_of_gyration_2 + o0, d->evolver_radius_of_gyration_2 + o0 + len);
            tab_love.insert(tab_love.end(), d->evolver_love_number + o0, d->evolver_love_number + o0 + len);
            if (d->evolver_inverse_tidal_q_factor)
                tab_itq.insert(tab_itq.end(), d->evolver_inverse_tidal_q_factor + o0, d->evolver_inverse_tidal_q_factor + o0 + len);
            else
                tab_itq.insert(tab_itq.end(), (size_t)len, 0.);
        }
    }

    // ---- device allocations ----
    DevEnsemble &E = h->E;
    E.n_systems = ns;
    E.n_bodies = nb;
    int *d_bo, *d_host, *d_coord, *d_tw, *d_evo_type, *d_evo_off, *d_evo_len, *d_left;
    uint32_t *d_eff, *d_status, *d_roles;
    double *d_dt, *d_tl, *d_t, *d_lh, *d_hp;
    long long *d_iter;
    double *d_x, *d_v, *d_L, *d_sp, *d_ev, *d_eL, *d_acc, *d_R, *d_rg2, *d_I, *d_m, *d_sigma, *d_k2, *d_k2f, *d_wk, *d_ws, *d_roche;
    double *d_tt, *d_tR, *d_trg2, *d_tlove, *d_titq, *d_sscale, *d_sfactor, *d_lag, *d_mig, *d_disk, *d_pair;
    if ((rc = dalloc(h, &d_bo, ns + 1)) || (rc = dalloc(h, &d_host, ns)) || (rc = dalloc(h, &d_coord, ns)) ||
        (rc = dalloc(h, &d_tw, ns)) || (rc = dalloc(h, &d_eff, ns)) || (rc = dalloc(h, &d_status, ns)) ||
        (rc = dalloc(h, &d_dt, ns)) || (rc = dalloc(h, &d_tl, ns)) || (rc = dalloc(h, &d_t, ns)) ||
        (rc = dalloc(h, &d_lh, ns)) || (rc = dalloc(h, &d_hp, ns)) || (rc = dalloc(h, &d_iter, ns)) ||
        (rc = dalloc(h, &d_x, 3 * nb)) || (rc = dalloc(h, &d_v, 3 * nb)) || (rc = dalloc(h, &d_L, 3 * nb)) ||
        (rc = dalloc(h, &d_sp, 3 * nb)) || (rc = dalloc(h, &d_ev, 3 * nb)) || (rc = dalloc(h, &d_eL, 3 * nb)) ||
        (rc = dalloc(h, &d_acc, 3 * nb)) || (rc = dalloc(h, &d_R, nb)) || (rc = dalloc(h, &d_rg2, nb)) ||
        (rc = dalloc(h, &d_I, nb)) || (rc = dalloc(h, &d_m, nb)) || (rc = dalloc(h, &d_sigma, nb)) ||
        (rc = dalloc(h, &d_k2, nb)) || (rc = dalloc(h, &d_k2f, nb)) || (rc = dalloc(h, &d_wk, nb)) ||
        (rc = dalloc(h, &d_ws, nb)) || (rc = dalloc(h, &d_roles, nb)) || (rc = dalloc(h, &d_roche, (size_t)nb * POS_MAX_BODIES)) ||
        (rc = dalloc(h, &d_evo_type, nb)) || (rc = dalloc(h, &d_evo_off, nb)) || (rc = dalloc(h, &d_evo_len, nb)) ||
        (rc = dalloc(h, &d_left, nb)) || (rc = dalloc(h, &d_tt, tab_t.size())) || (rc = dalloc(h, &d_tR, tab_t.size())) ||
        (rc = dalloc(h, &d_trg2, tab_t.size())) || (rc = dalloc(h, &d_tlove, tab_t.size())) ||
        (rc = dalloc(h, &d_titq, tab_t.size())) || (rc = dalloc(h, &d_sscale, nb)) || (rc = dalloc(h, &d_sfactor, nb)) ||
        (rc = dalloc(h, &d_lag, nb)) || (rc = dalloc(h, &d_mig, nb)) || (rc = dalloc(h, &d_disk, 6 * (size_t)ns)) ||
        (rc = dalloc(h, &d_pair, nb))) {
        pos_destroy(h);
        return rc;
    }
    E.body_offset = d_bo; E.host = d_host; E.coord = d_coord; E.effects = d_eff;
    E.dt = d_dt; E.time_limit = d_tl; E.hist_period = d_hp; E.t = d_t; E.last_hist = d_lh;
    E.iteration = d_iter; E.status = d_status; E.timestep_warning = d_tw;
    E.x = d_x; E.v = d_v; E.L = d_L; E.spin = d_sp; E.err_v = d_ev; E.err_L = d_eL; E.acc = d_acc;
    E.radius = d_R; E.rg2 = d_rg2; E.inertia = d_I;
    E.mass = d_m; E.sigma = d_sigma; E.k2 = d_k2; E.k2f = d_k2f; E.wind_k = d_wk; E.wind_sat = d_ws; E.roles = d_roles;
    E.roche = d_roche;
    E.evo_type = d_evo_type; E.evo_off = d_evo_off; E.evo_len = d_evo_len; E.left_index = d_left;
    E.tab_time = d_tt; E.tab_radius = d_tR; E.tab_rg2 = d_trg2; E.tab_love = d_tlove; E.tab_itq = d_titq;
    E.sigma_scale = d_sscale; E.sigma_factor = d_sfactor; E.lag = d_lag; E.disk_migration = d_mig; E.disk_props = d_disk;
    E.pair_ph = d_pair;

    if ((rc = upload(h, d_bo, h->body_offset_dev)) || (rc = upload(h, d_host, host_idx)) || (rc = upload(h, d_coord, coord)) ||
        (rc = upload(h, d_tw, tw)) || (rc = upload(h, d_eff, eff)) || (rc = upload(h, d_status, status)) ||
        (rc = upload(h, d_dt, dt)) || (rc = upload(h, d_tl, tl)) || (rc = upload(h, d_t, t)) || (rc = upload(h, d_lh, lh)) ||
        (rc = upload(h, d_hp, hp)) || (rc = upload(h, d_iter, iter)) || (rc = upload(h, d_x, x)) || (rc = upload(h, d_v, v)) ||
        (rc = upload(h, d_L, L)) || (rc = upload(h, d_sp, sp)) || (rc = upload(h, d_ev, ev)) || (rc = upload(h, d_eL, eL)) ||
        (rc = upload(h, d_acc, acc)) || (rc = upload(h, d_R, R)) || (rc = upload(h, d_rg2, rg2)) || (rc = upload(h, d_I, I)) ||
        (rc = upload(h, d_m, m)) || (rc = upload(h, d_sigma, sigma)) || (rc = upload(h, d_k2, k2)) || (rc = upload(h, d_k2f, k2f)) ||
        (rc = upload(h, d_wk, wk)) || (rc = upload(h, d_ws, ws)) || (rc = upload(h, d_roles, roles)) ||
        (rc = upload(h, d_evo_type, evo_type)) || (rc = upload(h, d_evo_off, evo_off)) || (rc = upload(h, d_evo_len, evo_len)) ||
        (rc = upload(h, d_left, left)) || (rc = upload(h, d_tt, tab_t)) || (rc = upload(h, d_tR, tab_R)) ||
        (rc = upload(h, d_trg2, tab_rg2)) || (rc = upload(h, d_tlove, tab_love)) || (rc = upload(h, d_titq, tab_itq)) ||
        (rc = upload(h, d_sscale, sig_scale)) || (rc = upload(h, d_sfactor, sig_factor)) ||
        (rc = upload(h, d_lag, zeros_nb)) || (rc = upload(h, d_mig, zeros_nb)) || (rc = upload(h, d_disk, disk)) ||
        (rc = upload(h, d_pair, pair))) {
        pos_destroy(h);
        return rc;
    }

    // initialize_physical_values on the device, one system per wave
    {
        pos_options init_o{};
        init_o.mapping = POS_MAPPING_PLAIN;
        if ((rc = build_launch(h, init_o))) { pos_destroy(h); return rc; }
        hipLaunchKernelGGL(init_kernel, dim3(h->E.n_waves), dim3(64), 0, h->stream, h->E);
        if (hipGetLastError() != hipSuccess) { pos_destroy(h); return fail(POS_E_HIP, "init kernel launch failed"); }
    }
    // Roche radii (universe.rs:177-196): initialize_physical_values computes
    // them once, at t = 0, from the radii evolved to t = 0; a resumed system
    // keeps the values its snapshot carries (or, without them, gets them from
    // its current radii)
    {
        std::vector<double> Rnow;
        if ((rc = download(h, Rnow, d_R, nb))) { pos_destroy(h); return rc; }
        if (hipStreamSynchronize(h->stream) != hipSuccess) { pos_destroy(h); return fail(POS_E_HIP, "synchronize failed"); }
        std::vector<double> roche((size_t)nb * POS_MAX_BODIES, 0.);
        for (int ds = 0; ds < ns; ++ds) {
            const int b0 = h->body_offset_dev[ds], n = h->n_per_sys_dev[ds];
            const bool resumed = t[ds] != 0. && d->roche_radius;
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    if (i == j) continue;
                    const double carried =
                        resumed ? d->roche_radius[(size_t)h->body_perm[b0 + i] * POS_MAX_BODIES + j] : std::nan("");
                    roche[(size_t)(b0 + i) * POS_MAX_BODIES + j] =
                        std::isnan(carried) ? host_roche(m[b0 + i], m[b0 + j], Rnow[b0 + i], Rnow[b0 + j]) : carried;
                }
        }
        if ((rc = upload(h, d_roche, roche))) { pos_destroy(h); return rc; }
    }
    if ((rc = build_launch(h, o))) { pos_destroy(h); return rc; }
    if (hipStreamSynchronize(h->stream) != hipSuccess) { pos_destroy(h); return fail(POS_E_HIP, "synchronize failed"); }
    // the no-allocation operator's device resources: the scratch array's
    // address on this device and the body permutation
    {
        double *scratch_ptr = nullptr;
        if ((rc = noalloc_scratch(h->device, &scratch_ptr)) || (rc = body_staging(h))) { pos_destroy(h); return rc; }
    }
    h->eff_dev = eff;
    h->roles_dev = roles;
    h->host_dev = host_idx;
    h->coord_dev = coord;
    h->evo_dev = evo_type;
    h->mass_dev = m;
    h->sigma_dev = sigma;
    h->dt_dev = dt;
    h->E.hist = nullptr;
    h->E.hist_count = nullptr;
    h->E.hist_cap = 0;
    // with graph replay, capture now rather than inside the first long pos_step
    if (h->opts.use_graph && (rc = ensure_graph(h))) { pos_destroy(h); return rc; }
    *out = h;
    return POS_OK;
}

int32_t pos_step(pos_ens *h, int64_t n_steps) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    if (n_steps < 0) return fail(POS_E_INVALID, "n_steps must be >= 0");
    HIP_TRY(hipSetDevice(h->device));
    const int K = h->steps_per_launch;
    int64_t full = n_steps / K;
    const int rest = (int)(n_steps - full * K);
    h->stats.steps += n_steps;
    if (h->opts.use_graph && full >= GRAPH_LAUNCHES) {
        int grc = ensure_graph(h);
        if (grc) return grc;
        while (full >= GRAPH_LAUNCHES) {
            HIP_TRY(hipGraphLaunch(h->graph_exec, h->stream));
            full -= GRAPH_LAUNCHES;
            h->stats.graph_replays += 1;
            h->stats.launches_per_replay = h->graph_launches;
        }
    }
    int rc, nl = 0;
    for (int64_t i = 0; i < full; ++i) {
        if ((rc = launch_steps(h, K, &nl))) return rc;
        h->stats.kernel_launches += nl;
    }
    if (rest > 0) {
        if ((rc = launch_steps(h, rest, &nl))) return rc;
        h->stats.kernel_launches += nl;
    }
    return POS_OK;
}

int32_t pos_synchronize(pos_ens *h) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return POS_OK;
}

int32_t pos_get_state(pos_ens *h, pos_state *o) {
    if (!h || !o) return fail(POS_E_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->device));
    const DevEnsemble &E = h->E;
    const int ns = h->n_systems, nb = h->n_bodies;
    std::vector<double> buf;
    int rc;
    auto vec3 = [&](double *dst, double *src) -> int { return dst ? download_bodies(h, dst, src, 3) : POS_OK; };
    auto vec1 = [&](double *dst, double *src) -> int { return dst ? download_bodies(h, dst, src, 1) : POS_OK; };
    if ((rc = vec3(o->position, E.x)) || (rc = vec3(o->velocity, E.v)) || (rc = vec3(o->acceleration, E.acc)) ||
        (rc = vec3(o->angular_momentum, E.L)) || (rc = vec3(o->spin, E.spin)) || (rc = vec3(o->velocity_error, E.err_v)) ||
        (rc = vec3(o->angular_momentum_error, E.err_L)) || (rc = vec1(o->radius, E.radius)) ||
        (rc = vec1(o->radius_of_gyration_2, E.rg2)) || (rc = vec1(o->moment_of_inertia, E.inertia)))
        return rc;
    auto sys_d = [&](double *dst, double *src) -> int {
        if (!dst) return POS_OK;
        int r = download(h, buf, src, ns);
        if (r) return r;
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (int ds = 0; ds < ns; ++ds) dst[h->sys_perm[ds]] = buf[ds];
        return POS_OK;
    };
    if ((rc = sys_d(o->current_time, E.t)) || (rc = sys_d(o->last_historic_snapshot_time, E.last_hist))) return rc;
    if (o->current_iteration) {
        std::vector<long long> it;
        if ((rc = download(h, it, E.iteration, ns))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (int ds = 0; ds < ns; ++ds) o->current_iteration[h->sys_perm[ds]] = it[ds];
    }
    if (o->status) {
        std::vector<uint32_t> st;
        if ((rc = download(h, st, E.status, ns))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (int ds = 0; ds < ns; ++ds) o->status[h->sys_perm[ds]] = st[ds];
    }
    if (o->evolver_left_index) {
        std::vector<int> li;
        if ((rc = download(h, li, E.left_index, nb))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (int db = 0; db < nb; ++db) o->evolver_left_index[h->body_perm[db]] = li[db];
    }
    if (o->timestep_warning) {
        std::vector<int> tw;
        if ((rc = download(h, tw, E.timestep_warning, ns))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (int ds = 0; ds < ns; ++ds) o->timestep_warning[h->sys_perm[ds]] = tw[ds];
    }
    if ((rc = vec1(o->pair_dissipation_factor, E.pair_ph))) return rc;
    if (o->roche_radius) {   // rows move with their bodies; columns index within the system
        std::vector<double> rr;
        if ((rc = download(h, rr, E.roche, (size_t)nb * POS_MAX_BODIES))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (int db = 0; db < nb; ++db)
            std::memcpy(o->roche_radius + (size_t)h->body_perm[db] * POS_MAX_BODIES, rr.data() + (size_t)db * POS_MAX_BODIES,
                        POS_MAX_BODIES * sizeof(double));
    }
    return POS_OK;
}

int32_t pos_set_state(pos_ens *h, const pos_state *in) {
    if (!h || !in) return fail(POS_E_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->device));
    const DevEnsemble &E = h->E;
    const int ns = h->n_systems, nb = h->n_bodies;
    std::vector<double> buf;
    int rc;
    auto vec3 = [&](const double *src, double *dst) -> int { return src ? upload_bodies(h, dst, src, 3) : POS_OK; };
    auto vec1 = [&](const double *src, double *dst) -> int { return src ? upload_bodies(h, dst, src, 1) : POS_OK; };
    if ((rc = vec3(in->position, E.x)) || (rc = vec3(in->velocity, E.v)) || (rc = vec3(in->acceleration, E.acc)) ||
        (rc = vec3(in->angular_momentum, E.L)) || (rc = vec3(in->spin, E.spin)) || (rc = vec3(in->velocity_error, E.err_v)) ||
        (rc = vec3(in->angular_momentum_error, E.err_L)) || (rc = vec1(in->radius, E.radius)) ||
        (rc = vec1(in->radius_of_gyration_2, E.rg2)) || (rc = vec1(in->moment_of_inertia, E.inertia)))
        return rc;
    auto sys_d = [&](const double *src, double *dst) -> int {
        if (!src) return POS_OK;
        buf.assign(ns, 0.);
        for (int ds = 0; ds < ns; ++ds) buf[ds] = src[h->sys_perm[ds]];
        int r = upload(h, dst, buf);
        if (r) return r;
        HIP_TRY(hipStreamSynchronize(h->stream));
        return POS_OK;
    };
    if ((rc = sys_d(in->current_time, E.t)) || (rc = sys_d(in->last_historic_snapshot_time, E.last_hist))) return rc;
    if (in->current_iteration) {
        std::vector<long long> it(ns);
        for (int ds = 0; ds < ns; ++ds) it[ds] = in->current_iteration[h->sys_perm[ds]];
        if ((rc = upload(h, E.iteration, it))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    if (in->status) {
        std::vector<uint32_t> st(ns);
        for (int ds = 0; ds < ns; ++ds) st[ds] = in->status[h->sys_perm[ds]];
        if ((rc = upload(h, E.status, st))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    if (in->evolver_left_index) {
        std::vector<int> li(nb);
        for (int db = 0; db < nb; ++db) li[db] = in->evolver_left_index[h->body_perm[db]];
        if ((rc = upload(h, E.left_index, li))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    if (in->timestep_warning) {
        std::vector<int> tw(ns);
        for (int ds = 0; ds < ns; ++ds) tw[ds] = in->timestep_warning[h->sys_perm[ds]];
        if ((rc = upload(h, E.timestep_warning, tw))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    if ((rc = vec1(in->pair_dissipation_factor, E.pair_ph))) return rc;
    if (in->roche_radius) {
        std::vector<double> rr((size_t)nb * POS_MAX_BODIES);
        for (int db = 0; db < nb; ++db)
            std::memcpy(rr.data() + (size_t)db * POS_MAX_BODIES, in->roche_radius + (size_t)h->body_perm[db] * POS_MAX_BODIES,
                        POS_MAX_BODIES * sizeof(double));
        if ((rc = upload(h, const_cast<double *>(E.roche), rr))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    return POS_OK;
}

int32_t pos_set_time_limit(pos_ens *h, double time_limit) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    if (!(time_limit > 0.)) return POS_OK;
    HIP_TRY(hipSetDevice(h->device));
    std::vector<double> tl(h->n_systems, time_limit);
    int rc = upload(h, const_cast<double *>(h->E.time_limit), tl);
    if (rc) return rc;
    // a system re-opens if the new limit lies beyond t + dt
    std::vector<uint32_t> st;
    std::vector<double> t, dt;
    if ((rc = download(h, st, h->E.status, h->n_systems)) || (rc = download(h, t, h->E.t, h->n_systems)) ||
        (rc = download(h, dt, h->E.dt, h->n_systems)))
        return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int s = 0; s < h->n_systems; ++s)
        if (!(t[s] + dt[s] > time_limit)) st[s] &= ~(uint32_t)POS_STATUS_COMPLETED;
    if ((rc = upload(h, h->E.status, st))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return POS_OK;
}

int32_t pos_set_historic_snapshot_period(pos_ens *h, double period) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    if (!(period > 0.)) return POS_OK;
    HIP_TRY(hipSetDevice(h->device));
    std::vector<double> hp(h->n_systems, period);
    int rc = upload(h, const_cast<double *>(h->E.hist_period), hp);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return POS_OK;
}

int32_t pos_get_counts(const pos_ens *h, int32_t *n_systems, int32_t *n_bodies) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    if (n_systems) *n_systems = h->n_systems;
    if (n_bodies) *n_bodies = h->n_bodies;
    return POS_OK;
}

int32_t pos_get_stream(const pos_ens *h, void **stream) {
    if (!h || !stream) return fail(POS_E_INVALID, "NULL argument");
    *stream = (void *)h->stream;
    return POS_OK;
}

int32_t pos_get_options(const pos_ens *h, pos_options *out) {
    if (!h || !out) return fail(POS_E_INVALID, "NULL argument");
    *out = h->opts;
    return POS_OK;
}

int32_t pos_gravity_acceleration(pos_ens *h, double *out_acc, uint32_t *out_status) {
    if (!h || !out_acc) return fail(POS_E_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->device));
    const int ns = h->n_systems, nb = h->n_bodies;
    int rc;
    // one evaluation of every system over the current launch table
    if (!h->d_op_a && ((rc = dalloc(h, &h->d_op_a, 3 * (size_t)nb)) || (rc = dalloc(h, &h->d_op_l, 3 * (size_t)nb)) ||
                       (rc = dalloc(h, &h->d_op_st, ns))))
        return rc;
    double *d_out = h->d_op_a;
    uint32_t *d_st = h->d_op_st;
    HIP_TRY(hipEventRecord(h->op_ev[0], h->stream));
    hipLaunchKernelGGL(gravity_kernel, dim3(h->E.n_waves), dim3(64), 0, h->stream, h->E, d_out, d_st);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(h->op_ev[1], h->stream));
    std::vector<uint32_t> st;
    if ((rc = download_bodies(h, out_acc, d_out, 3)) || (rc = download(h, st, d_st, ns))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if ((rc = record_operator_time(h))) return rc;
    if (out_status)
        for (int ds = 0; ds < ns; ++ds) out_status[h->sys_perm[ds]] |= st[ds];
    return POS_OK;
}

int32_t pos_additional_effects(pos_ens *h, double *out_acc, double *out_dLdt) {
    if (!h || !out_acc || !out_dLdt) return fail(POS_E_INVALID, "NULL argument");
    HIP_TRY(hipSetDevice(h->device));
    const int nb = h->n_bodies;
    int rc;
    // one evaluation of every system over the current launch table
    if (!h->d_op_a && ((rc = dalloc(h, &h->d_op_a, 3 * (size_t)nb)) || (rc = dalloc(h, &h->d_op_l, 3 * (size_t)nb)) ||
                       (rc = dalloc(h, &h->d_op_st, h->n_systems))))
        return rc;
    double *d_a = h->d_op_a, *d_l = h->d_op_l;
    HIP_TRY(hipEventRecord(h->op_ev[0], h->stream));
    if ((rc = launch_effects(h, d_a, d_l))) return rc;
    HIP_TRY(hipEventRecord(h->op_ev[1], h->stream));
    if ((rc = download_bodies(h, out_acc, d_a, 3)) || (rc = download_bodies(h, out_dLdt, d_l, 3))) return rc;
    return record_operator_time(h);
}

int32_t pos_additional_effects_noalloc(pos_ens *h, double *dev_acc, double *dev_dLdt) {
    if (!h || !dev_acc || !dev_dLdt) return fail(POS_E_INVALID, "NULL argument");
    const int nb = h->n_bodies;
    if (nb > POS_NOALLOC_MAX_BODIES)
        return fail(POS_E_LIMIT, "the ensemble has " + std::to_string(nb) + " bodies; the scratch holds " +
                                     std::to_string(POS_NOALLOC_MAX_BODIES));
    HIP_TRY(hipSetDevice(h->device));
    double *s = nullptr;   // set up at pos_create; found here without allocating
    int rc;
    if ((rc = noalloc_scratch(h->device, &s))) return rc;
    double *d_a = s, *d_l = s + 3 * (size_t)POS_NOALLOC_MAX_BODIES;
    // every body belongs to one lane of one evaluated wavefront, so the kernels
    // write every element the reordering reads: nothing needs clearing
    if ((rc = launch_effects(h, d_a, d_l))) return rc;
    const dim3 grid((nb + 255) / 256);
    hipLaunchKernelGGL(unpermute_kernel, grid, dim3(256), 0, h->stream, d_a, h->d_body_perm, nb, 3, dev_acc);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(unpermute_kernel, grid, dim3(256), 0, h->stream, d_l, h->d_body_perm, nb, 3, dev_dLdt);
    HIP_TRY(hipGetLastError());
    return POS_OK;
}

int32_t pos_get_launch_stats(const pos_ens *h, pos_launch_stats *out) {
    if (!h || !out) return fail(POS_E_INVALID, "NULL argument");
    *out = h->stats;
    return POS_OK;
}

int32_t pos_tune(pos_ens *h, int32_t probe_steps, pos_options *chosen) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    if (probe_steps <= 0) probe_steps = 32;
    HIP_TRY(hipSetDevice(h->device));
    // snapshot the mutable state so that probing leaves it untouched: one
    // device buffer, filled and restored with device-to-device copies on
    // the handle's stream (the GPU does not wait on host copies between the
    // probes, nor before the caller's first steps)
    const int ns = h->n_systems, nb = h->n_bodies;
    int rc;
    // the probes are not part of the caller's run: no historic records, and
    // the launch counters are restored at the end
    HistoryDetach detached(h);
    const pos_launch_stats stats0 = h->stats;
    const DevEnsemble &E = h->E;
    struct Field { void *p; size_t bytes; };
    const Field fields[] = {
        {E.x, 3 * (size_t)nb * sizeof(double)}, {E.v, 3 * (size_t)nb * sizeof(double)},
        {E.L, 3 * (size_t)nb * sizeof(double)}, {E.spin, 3 * (size_t)nb * sizeof(double)},
        {E.err_v, 3 * (size_t)nb * sizeof(double)}, {E.err_L, 3 * (size_t)nb * sizeof(double)},
        {E.acc, 3 * (size_t)nb * sizeof(double)}, {E.radius, (size_t)nb * sizeof(double)},
        {E.rg2, (size_t)nb * sizeof(double)}, {E.inertia, (size_t)nb * sizeof(double)},
        {E.t, (size_t)ns * sizeof(double)}, {E.last_hist, (size_t)ns * sizeof(double)},
        {E.iteration, (size_t)ns * sizeof(long long)}, {E.status, (size_t)ns * sizeof(uint32_t)},
        {E.timestep_warning, (size_t)ns * sizeof(int)}, {E.left_index, (size_t)nb * sizeof(int)},
        {E.lag, (size_t)nb * sizeof(double)}, {E.disk_migration, (size_t)nb * sizeof(double)},
        {E.pair_ph, (size_t)nb * sizeof(double)}};
    size_t total = 0;
    for (const Field &f : fields) total += (f.bytes + 255) / 256 * 256;
    struct Snapshot {   // freed on every return path
        char *p = nullptr;
        ~Snapshot() { if (p) (void)hipFree(p); }
    } snap;
    HIP_TRY(hipMalloc((void **)&snap.p, total));
    auto copy_all = [&](bool save) -> int {
        size_t off = 0;
        for (const Field &f : fields) {
            if (f.bytes && f.p) {
                if (save) HIP_TRY(hipMemcpyAsync(snap.p + off, f.p, f.bytes, hipMemcpyDeviceToDevice, h->stream));
                else HIP_TRY(hipMemcpyAsync(f.p, snap.p + off, f.bytes, hipMemcpyDeviceToDevice, h->stream));
            }
            off += (f.bytes + 255) / 256 * 256;
        }
        return POS_OK;
    };
    if ((rc = copy_all(true))) return rc;
    auto restore = [&]() -> int {
        int r;
        if ((r = copy_all(false))) return r;
        HIP_TRY(hipStreamSynchronize(h->stream));
        return POS_OK;
    };
    int max_n = 0;
    for (int n : h->n_per_sys_dev) max_n = std::max(max_n, n);
    // next power of two / exactly N lanes per size class
    std::vector<int> groups = {0, -1};
    {
        // exactly N lanes with sizes mixed in a wavefront, whenever the
        // systems differ in size: the packing needs fewer wavefronts, and
        // whether that pays depends on how the last round of wavefronts
        // falls on the SIMDs (one wavefront per SIMD runs faster than two),
        // which only a measurement shows
        std::map<int, int64_t> per_size;
        for (int n : h->n_per_sys_dev) ++per_size[n];
        if (per_size.size() > 1) groups.push_back(-2);
    }
    for (int g = max_n; g <= 16; ++g)
        if (64 % g != 0) groups.push_back(g);   // exact packing for sizes that do not divide 64
    std::vector<int> ks = {16, 64};
    const int keep_graph = h->opts.use_graph;
    pos_options best = h->opts;
    double best_t = std::numeric_limits<double>::infinity();
    EventPair ev;
    HIP_TRY(ev.create());
    const hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
    for (int g : groups) {
        for (int K : ks) {
            pos_options o{};
            o.device = h->device;
            o.mapping = POS_MAPPING_PACKED;
            o.group_size = g;
            o.steps_per_launch = K;
            o.use_graph = 0;
            if ((rc = build_launch(h, o))) return rc;
            const int steps = std::max(probe_steps, K);
            if ((rc = pos_step(h, K))) return rc;   // warm-up
            HIP_TRY(hipEventRecord(e0, h->stream));
            if ((rc = pos_step(h, steps))) return rc;
            HIP_TRY(hipEventRecord(e1, h->stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            const double per_step = ms / steps;
            if (per_step < best_t) { best_t = per_step; best = o; }
            if ((rc = restore())) return rc;
        }
    }
    // graph replay against direct launches for the chosen mapping: a graph
    // saves host work and launch gaps on short launches and measures the same
    // as direct launches on long ones, so it is kept unless direct launches
    // are faster by more than the timing spread (2%)
    if (keep_graph) {
        double t_mode[2] = {0., 0.};
        for (int mode = 0; mode < 2; ++mode) {
            pos_options o = best;
            o.use_graph = mode;
            if ((rc = build_launch(h, o))) return rc;
            const int steps = GRAPH_LAUNCHES * o.steps_per_launch;
            if ((rc = pos_step(h, steps))) return rc;   // warm-up (captures the graph)
            HIP_TRY(hipEventRecord(e0, h->stream));
            if ((rc = pos_step(h, steps))) return rc;
            HIP_TRY(hipEventRecord(e1, h->stream));
            HIP_TRY(hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            t_mode[mode] = ms;
            if ((rc = restore())) return rc;
        }
        best.use_graph = t_mode[1] <= 1.02 * t_mode[0] ? 1 : 0;
    } else {
        best.use_graph = 0;
    }
    detached.release();
    h->stats = stats0;
    if ((rc = build_launch(h, best))) return rc;
    if (best.use_graph && (rc = ensure_graph(h))) return rc;   // captured now, not in the caller's first steps
    if (chosen) *chosen = h->opts;
    return POS_OK;
}

int32_t pos_history_enable(pos_ens *h, int32_t capacity) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    if (capacity <= 0) return fail(POS_E_INVALID, "capacity must be positive");
    if (h->E.hist) return fail(POS_E_INVALID, "historic snapshots are already enabled");
    HIP_TRY(hipSetDevice(h->device));
    const size_t bytes = (size_t)capacity * HIST_FIELDS * h->n_bodies * sizeof(double);
    int rc;
    double *buf = nullptr;
    int *cnt = nullptr;
    if ((rc = dalloc(h, &buf, (size_t)capacity * HIST_FIELDS * h->n_bodies)) || (rc = dalloc(h, &cnt, h->n_systems)))
        return rc;
    (void)bytes;
    std::vector<int> zero(h->n_systems, 0);
    if ((rc = upload(h, cnt, zero))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->E.hist = buf;
    h->E.hist_count = cnt;
    h->E.hist_cap = capacity;
    h->hist_cap = capacity;
    // captured launches hold the old kernel arguments
    if (h->graph_exec) { (void)hipGraphExecDestroy(h->graph_exec); h->graph_exec = nullptr; }
    if (h->graph) { (void)hipGraphDestroy(h->graph); h->graph = nullptr; }
    h->graph_launches = 0;
    if (h->opts.use_graph && (rc = ensure_graph(h))) return rc;
    return POS_OK;
}

int32_t pos_history_drain(pos_ens *h, pos_history_record *out, int64_t max_records, int64_t *n_out) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    if (!n_out) return fail(POS_E_INVALID, "n_out is NULL");
    *n_out = 0;
    if (!h->E.hist) return fail(POS_E_INVALID, "historic snapshots are not enabled");
    HIP_TRY(hipSetDevice(h->device));
    const int ns = h->n_systems, nb = h->n_bodies;
    std::vector<int> cnt;
    int rc;
    if ((rc = download(h, cnt, h->E.hist_count, ns))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    int max_slot = 0;
    int64_t total = 0;
    for (int ds = 0; ds < ns; ++ds) {
        const int c = std::min(cnt[ds], h->hist_cap);
        max_slot = std::max(max_slot, c);
        total += (int64_t)c * h->n_per_sys_dev[ds];
    }
    if (total > max_records) return fail(POS_E_LIMIT, "the output holds fewer records than are pending (" +
                                                      std::to_string(total) + ")");
    std::vector<double> buf;
    if (max_slot > 0) {
        if ((rc = download(h, buf, h->E.hist, (size_t)max_slot * HIST_FIELDS * nb))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    // input order: systems, then records, then bodies
    std::vector<int> dev_of(ns);
    for (int ds = 0; ds < ns; ++ds) dev_of[h->sys_perm[ds]] = ds;
    int64_t k = 0;
    for (int s = 0; s < ns; ++s) {
        const int ds = dev_of[s];
        const int c = std::min(cnt[ds], h->hist_cap);
        const int b0 = h->body_offset_dev[ds], n = h->n_per_sys_dev[ds];
        for (int r = 0; r < c; ++r) {
            const double *rec = buf.data() + (size_t)r * HIST_FIELDS * nb;
            for (int i = 0; i < n; ++i) {
                const int db = b0 + i;
                pos_history_record &o = out[k++];
                auto f = [&](int field) { return rec[(size_t)field * nb + db]; };
                o.system = s;
                o.body = i;
                o.time = f(HF_TIME);
                o.time_step = h->dt_dev[ds];
                for (int q = 0; q < 3; ++q) {
                    o.position[q] = f(HF_X + q);
                    o.spin[q] = f(HF_SPIN + q);
                    o.velocity[q] = f(HF_V + q);
                }
                o.mass = h->mass_dev[db];
                o.radius = f(HF_RADIUS);
                o.radius_of_gyration_2 = f(HF_RG2);
                o.love_number = f(HF_LOVE);
                o.scaled_dissipation_factor = h->sigma_dev[db];
                o.lag_angle = f(HF_LAG);
                o.denergy_dt = f(HF_DEDT);
                o.migration_timescale = f(HF_MIGRATION);
            }
        }
    }
    std::vector<int> zero(ns, 0);
    if ((rc = upload(h, h->E.hist_count, zero))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    *n_out = k;
    return POS_OK;
}

int32_t pos_adjoint_record(pos_ens *h, int64_t n_steps, int32_t segment_steps) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    if (n_steps <= 0) return fail(POS_E_INVALID, "n_steps must be positive");
    if (segment_steps < 0) return fail(POS_E_INVALID, "segment_steps must be >= 0");
    for (int ds = 0; ds < h->n_systems; ++ds) {
        const std::string where = "system " + std::to_string(h->sys_perm[ds]) + ": ";

        if (h->eff_dev[ds] & POS_EFFECT_DISK)
            return fail(POS_E_UNSUPPORTED, where + "the adjoint does not support the disk");
        if ((h->variant_dev[ds] & 3) != GRV_KIDDER)
            return fail(POS_E_UNSUPPORTED, where + "the adjoint supports Kidder1995 general relativity only");
        if (h->variant_dev[ds] & VAR_EXT)
            return fail(POS_E_UNSUPPORTED, where + "the adjoint does not support evolving dissipation");
        const uint32_t eff = h->eff_dev[ds];
        for (int b = h->body_offset_dev[ds]; b < h->body_offset_dev[ds + 1]; ++b) {
            if (b - h->body_offset_dev[ds] == h->host_dev[ds]) continue;
            const uint32_t r = h->roles_dev[b];
            if (((eff & POS_EFFECT_TIDES) && !(r & POS_ROLE_TIDES_ORBITING)) ||
                ((eff & POS_EFFECT_ROTATIONAL_FLATTENING) && !(r & POS_ROLE_FLAT_ORBITING)) ||
                ((eff & POS_EFFECT_GENERAL_RELATIVITY) && !(r & POS_ROLE_GR_ORBITING)))
                return fail(POS_E_UNSUPPORTED, where + "the adjoint needs every body but the host to be an orbiting body "
                                                       "of each effect that is on");
        }
    }
    HIP_TRY(hipSetDevice(h->device));
    pos_ens::Adjoint &A = h->adj;
    HIP_TRY(hipStreamSynchronize(h->stream));
    adjoint_release(h);
    const int64_t seg64 = std::min<int64_t>(segment_steps > 0 ? segment_steps : 1000, n_steps);
    if (seg64 > (1 << 24)) return fail(POS_E_LIMIT, "segment_steps too large");
    const int seg = (int)seg64;
    const int64_t n_seg = (n_steps + seg - 1) / seg;
    const size_t nb = h->n_bodies, ns = h->n_systems;
    const std::vector<std::pair<void *, size_t>> fields = adjoint_state_fields(h);
    size_t state_bytes = 0;
    for (const auto &f : fields) state_bytes += (f.second + 255) / 256 * 256;
    const size_t ckpt_bytes = state_bytes * (size_t)(n_seg + 1);
    const size_t tape_bytes = (size_t)(seg + 1) * TAPE_FIELDS * nb * sizeof(double);
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    if (ckpt_bytes + 2 * tape_bytes + 64 * nb * sizeof(double) > free_b * 9 / 10)
        return fail(POS_E_LIMIT, "checkpoints (" + std::to_string(ckpt_bytes >> 20) + " MiB) and segment tape (" +
                                 std::to_string(tape_bytes >> 20) + " MiB) exceed free device memory; use fewer steps "
                                 "per segment or fewer systems");
    auto aalloc = [&](void **p, size_t bytes) -> int {
        if (hipMalloc(p, std::max<size_t>(bytes, 256)) != hipSuccess) return fail(POS_E_HIP, "hipMalloc failed (adjoint)");
        A.allocs.push_back(*p);
        return POS_OK;
    };
    int rc;
    if ((rc = aalloc((void **)&A.ckpt, ckpt_bytes)) || (rc = aalloc((void **)&A.tape[0], tape_bytes)) ||
        (rc = aalloc((void **)&A.tape[1], tape_bytes)) ||
        (rc = aalloc((void **)&A.iter0, ns * sizeof(long long))) ||
        (rc = aalloc((void **)&A.steps_done, ns * sizeof(long long))) ||
        (rc = aalloc((void **)&A.lx, 3 * nb * sizeof(double))) || (rc = aalloc((void **)&A.lv, 3 * nb * sizeof(double))) ||
        (rc = aalloc((void **)&A.lL, 3 * nb * sizeof(double))) || (rc = aalloc((void **)&A.gs, nb * sizeof(double))) ||
        (rc = aalloc((void **)&A.gk2, nb * sizeof(double))) || (rc = aalloc((void **)&A.gkf, nb * sizeof(double))))
        return rc;
    HIP_TRY(hipStreamCreateWithFlags(&A.sweep, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        HIP_TRY(hipEventCreateWithFlags(&A.recorded_ev[i], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&A.swept_ev[i], hipEventDisableTiming));
    }
    A.seg = seg;
    A.n_seg = (int)n_seg;
    A.n_steps = n_steps;
    A.state_bytes = state_bytes;
    HIP_TRY(hipMemcpyAsync(A.iter0, h->E.iteration, ns * sizeof(long long), hipMemcpyDeviceToDevice, h->stream));
    for (int64_t s = 0; s < n_seg; ++s) {
        if ((rc = adjoint_copy_state(h, A.ckpt + (size_t)s * state_bytes, true))) return rc;
        const int64_t len = std::min<int64_t>(seg, n_steps - s * seg);
        if ((rc = pos_step(h, len))) return rc;
    }
    if ((rc = adjoint_copy_state(h, A.ckpt + (size_t)n_seg * state_bytes, true))) return rc;
    std::vector<long long> i0(ns), i1(ns);
    if ((rc = download(h, i0, A.iter0, ns)) || (rc = download(h, i1, h->E.iteration, ns))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    std::vector<long long> done(ns);
    for (size_t s = 0; s < ns; ++s) done[s] = i1[s] - i0[s];
    if ((rc = upload(h, A.steps_done, done))) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    A.recorded = true;
    return POS_OK;
}

int32_t pos_adjoint(pos_ens *h, const pos_adjoint_seed *seed, pos_adjoint_grad *grad) {
    if (!h) return fail(POS_E_INVALID, "handle is NULL");
    if (!seed || !grad) return fail(POS_E_INVALID, "seed and grad must not be NULL");
    pos_ens::Adjoint &A = h->adj;
    if (!A.recorded) return fail(POS_E_INVALID, "no recorded run: call pos_adjoint_record first");
    HIP_TRY(hipSetDevice(h->device));
    const int nb = h->n_bodies;
    int rc;
    // seeds, input -> device order
    {
        std::vector<double> lx(3 * (size_t)nb, 0.), lv(3 * (size_t)nb, 0.), lL(3 * (size_t)nb, 0.), zero(nb, 0.);
        for (int db = 0; db < nb; ++db) {
            const int b = h->body_perm[db];
            for (int q = 0; q < 3; ++q) {
                if (seed->position) lx[(size_t)q * nb + db] = seed->position[3 * (size_t)b + q];
                if (seed->velocity) lv[(size_t)q * nb + db] = seed->velocity[3 * (size_t)b + q];
                if (seed->angular_momentum) lL[(size_t)q * nb + db] = seed->angular_momentum[3 * (size_t)b + q];
            }
        }
        if ((rc = upload(h, A.lx, lx)) || (rc = upload(h, A.lv, lv)) || (rc = upload(h, A.lL, lL)) ||
            (rc = upload(h, A.gs, zero)) || (rc = upload(h, A.gk2, zero)) || (rc = upload(h, A.gkf, zero)))
            return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));   // the host vectors go out of scope
    }
    AdjArgs Ar{};
    Ar.E = h->E;
    Ar.steps_done = A.steps_done;
    Ar.lx = A.lx; Ar.lv = A.lv; Ar.lL = A.lL;
    Ar.gs = A.gs; Ar.gk2 = A.gk2; Ar.gkf = A.gkf;
    // Segment s is recomputed from its checkpoint on the handle's stream
    // (writing tape s % 2) while segment s + 1 is swept in reverse on the
    // sweep stream: the sweep of s waits for the recomputation of s, and the
    // recomputation of s - 2 waits for the sweep of s, which frees its tape.
    auto seg_len = [&](int s) { return (int)std::min<int64_t>(A.seg, A.n_steps - (int64_t)s * A.seg); };
    auto recompute = [&](int s) -> int {
        const int buf = s & 1;
        if (s + 2 < A.n_seg) HIP_TRY(hipStreamWaitEvent(h->stream, A.swept_ev[buf], 0));
        int r = adjoint_copy_state(h, A.ckpt + (size_t)s * A.state_bytes, false);
        if (r) return r;
        for (int v = 0; v < N_VARIANTS; ++v) {   // pos_adjoint_record admits variants 0, 8 and 24 only
            const int w0 = h->variant_waves[v][0], w1 = h->variant_waves[v][1];
            if (w1 <= w0) continue;
            const dim3 grid(std::min(h->grid, w1 - w0));
            const TapeRef tape{A.tape[buf], nb};
            // the recomputed steps were recorded once already: no historic records
            const DevEnsemble Er = without_history(h->E);
            if (v == GRV_KIDDER)
                hipLaunchKernelGGL(whfast_record_kernel<GRV_KIDDER>, grid, dim3(64), 0, h->stream, Er, seg_len(s), tape, w0, w1);
            else if (v == (GRV_KIDDER | VAR_FAST))
                hipLaunchKernelGGL(whfast_record_kernel<GRV_KIDDER | VAR_FAST>, grid, dim3(64), 0, h->stream, Er, seg_len(s), tape,
                                   w0, w1);
            else if (v == (GRV_KIDDER | VAR_FAST | VAR_JACOBI))
                hipLaunchKernelGGL(whfast_record_kernel<GRV_KIDDER | VAR_FAST | VAR_JACOBI>, grid, dim3(64), 0, h->stream, Er,
                                   seg_len(s), tape, w0, w1);
            else
                return fail(POS_E_INVALID, "internal error: kernel variant without an adjoint");
            HIP_TRY(hipGetLastError());
        }
        HIP_TRY(hipEventRecord(A.recorded_ev[buf], h->stream));
        return POS_OK;
    };
    const bool alt = std::any_of(h->coord_dev.begin(), h->coord_dev.end(),
                                 [](int c) { return c != POS_COORD_DEMOCRATIC_HELIOCENTRIC; });
    const bool evo = std::any_of(h->eff_dev.begin(), h->eff_dev.end(),
                                 [](uint32_t e) { return (e & POS_EFFECT_EVOLUTION) != 0; });
    void (*sweep_kernel)(AdjArgs) = alt ? (evo ? adjoint_segment_kernel<true, true> : adjoint_segment_kernel<true, false>)
                                    : (evo ? adjoint_segment_kernel<false, true> : adjoint_segment_kernel<false, false>);
    // the sweep stream must also see the seeds uploaded on the handle's stream
    HIP_TRY(hipEventRecord(A.swept_ev[0], h->stream));
    HIP_TRY(hipStreamWaitEvent(A.sweep, A.swept_ev[0], 0));
    if ((rc = recompute(A.n_seg - 1))) return rc;
    for (int s = A.n_seg - 1; s >= 0; --s) {
        const int buf = s & 1;
        HIP_TRY(hipStreamWaitEvent(A.sweep, A.recorded_ev[buf], 0));
        Ar.T = TapeRef{A.tape[buf], nb};
        Ar.seg_start = (int64_t)s * A.seg;
        Ar.seg_len = seg_len(s);
        void *args[] = {&Ar};
        HIP_TRY(hipLaunchKernel((const void *)sweep_kernel, dim3(h->grid), dim3(64), args, 0, A.sweep));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(A.swept_ev[buf], A.sweep));
        if (s > 0 && (rc = recompute(s - 1))) return rc;
    }
    HIP_TRY(hipStreamWaitEvent(h->stream, A.swept_ev[0], 0));
    if ((rc = adjoint_copy_state(h, A.ckpt + (size_t)A.n_seg * A.state_bytes, false))) return rc;
    // results, device -> input order
    std::vector<double> lx, lv, lL, gs, gk2, gkf;
    if ((rc = download(h, lx, A.lx, 3 * (size_t)nb)) || (rc = download(h, lv, A.lv, 3 * (size_t)nb)) ||
        (rc = download(h, lL, A.lL, 3 * (size_t)nb)) || (rc = download(h, gs, A.gs, nb)) ||
        (rc = download(h, gk2, A.gk2, nb)) || (rc = download(h, gkf, A.gkf, nb)))
        return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int db = 0; db < nb; ++db) {
        const int b = h->body_perm[db];
        for (int q = 0; q < 3; ++q) {
            if (grad->position) grad->position[3 * (size_t)b + q] = lx[(size_t)q * nb + db];
            if (grad->velocity) grad->velocity[3 * (size_t)b + q] = lv[(size_t)q * nb + db];
            if (grad->angular_momentum) grad->angular_momentum[3 * (size_t)b + q] = lL[(size_t)q * nb + db];
        }
        if (grad->dissipation_factor) grad->dissipation_factor[b] = gs[db];
        if (grad->love_number) grad->love_number[b] = gk2[db];
        if (grad->fluid_love_number) grad->fluid_love_number[b] = gkf[db];
    }
    return POS_OK;
}

void pos_destroy(pos_ens *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    if (h->adj.sweep) (void)hipStreamSynchronize(h->adj.sweep);
    adjoint_release(h);
    for (void *p : h->allocs) (void)hipFree(p);
    for (hipEvent_t e : h->op_ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
