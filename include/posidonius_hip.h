/*
 * posidonius_hip.h -- C ABI of the MI355X WHFast ensemble integrator.
 *
 * The reference exposes its integration step through the Rust trait
 * `Integrator` (src/integrator/mod.rs:16-26), implemented by `WHFast`
 * (src/integrator/whfast.rs:98-305).  That trait is object-safe and
 * single-system; this ABI is the same contract lifted to an ensemble of
 * independent systems that live in device memory behind an opaque handle.
 * A future Rust `impl Integrator for WHFastHip` binds these symbols one to
 * one.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative POS_E_* code on
 *     failure; pos_last_error() returns a thread-local message for the last
 *     failure on the calling thread.
 *   - The caller owns every host buffer; the handle owns every device buffer
 *     and one HIP stream.  A handle is not thread-safe.
 *   - All arithmetic is IEEE binary64.  Units are the reference's Mercury
 *     units: AU, day, solar mass, G = K2 (src/constants.rs:59-61).
 *   - Physics failures never abort: the reference's panic!() paths
 *     (src/particles/universe.rs:218-238) and warnings
 *     (src/integrator/whfast.rs:389-391, 702-707) become per-system status
 *     bits, and a system with a POS_STATUS_FATAL bit set stops integrating.
 */
#ifndef POSIDONIUS_HIP_H
#define POSIDONIUS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POS_ABI_VERSION 4

/* Largest body count per system.  The reference caps a universe at
 * MAX_PARTICLES = 10 (src/constants.rs:3); the uneven-size ensemble needs 12.
 * A system is mapped to one lane group of at most 16 lanes. */
#define POS_MAX_BODIES 16

/* ---- return codes ------------------------------------------------------ */
#define POS_OK                0
#define POS_E_INVALID        -1   /* bad argument or unsupported case      */
#define POS_E_HIP            -2   /* a HIP runtime call failed             */
#define POS_E_NO_DEVICE      -3   /* no usable gfx950 device               */
#define POS_E_UNSUPPORTED    -4   /* feature not available on the device   */
#define POS_E_LIMIT          -5   /* size beyond a documented maximum      */

/* ---- WHFast alternative coordinates (whfast.rs:91-95) ------------------- */
#define POS_COORD_JACOBI                  0
#define POS_COORD_DEMOCRATIC_HELIOCENTRIC 1
#define POS_COORD_WHDS                    2

/* ---- general relativity implementation (general_relativity.rs:40-45) --- */
#define POS_GR_KIDDER1995    0
#define POS_GR_ANDERSON1975  1
#define POS_GR_NEWHALL1983   2
#define POS_GR_DISABLED      3

/* ---- consider_effects (universe.rs:40-47), one bitmask per system ------ *
 * Values as left by Universe::new after disable_unnecessary_effects
 * (universe.rs:699-764). */
#define POS_EFFECT_TIDES                (1u << 0)
#define POS_EFFECT_ROTATIONAL_FLATTENING (1u << 1)
#define POS_EFFECT_GENERAL_RELATIVITY   (1u << 2)
#define POS_EFFECT_DISK                 (1u << 3)
#define POS_EFFECT_WIND                 (1u << 4)
#define POS_EFFECT_EVOLUTION            (1u << 5)

/* ---- per-body effect roles (TidesEffect, RotationalFlatteningEffect,
 *      GeneralRelativityEffect, WindEffect enums) ------------------------- */
#define POS_ROLE_TIDES_CENTRAL   (1u << 0)
#define POS_ROLE_TIDES_ORBITING  (1u << 1)
#define POS_ROLE_FLAT_CENTRAL    (1u << 2)
#define POS_ROLE_FLAT_ORBITING   (1u << 3)
#define POS_ROLE_GR_CENTRAL      (1u << 4)
#define POS_ROLE_GR_ORBITING     (1u << 5)
#define POS_ROLE_WIND            (1u << 6)
#define POS_ROLE_DISK_CENTRAL    (1u << 7)   /* DiskEffect::CentralBody (disk.rs:44-49) */
#define POS_ROLE_DISK_ORBITING   (1u << 8)   /* DiskEffect::OrbitingBody                */

/* ---- evolution models (src/effects/evolution.rs:9-17) ------------------ */
#define POS_EVO_NON_EVOLVING              0
#define POS_EVO_LECONTE2011               1
#define POS_EVO_BARAFFE1998               2
#define POS_EVO_BARAFFE2015               3
#define POS_EVO_LECONTE_CHABRIER2013      4  /* without dynamical tides */
#define POS_EVO_LECONTE_CHABRIER2013_DYN  5  /* with dynamical tides    */
#define POS_EVO_BOLMONT_MATHIS2016        6
#define POS_EVO_GALLET_BOLMONT2017        7

/* ---- per-system status bits -------------------------------------------- */
#define POS_STATUS_COMPLETED          (1u << 0) /* t + dt > time_limit (whfast.rs:300-302) */
#define POS_STATUS_ROCHE              (1u << 1) /* universe.rs:225-228 */
#define POS_STATUS_COLLISION          (1u << 2) /* universe.rs:230-233 */
#define POS_STATUS_EJECTION           (1u << 3) /* universe.rs:234-237 */
#define POS_STATUS_ZERO_INERTIA       (1u << 4) /* particles/common.rs:5-7 */
#define POS_STATUS_MIDPOINT_NOT_CONVERGED (1u << 5) /* whfast.rs:389-391 (warning) */
#define POS_STATUS_TIMESTEP_WARNING   (1u << 6) /* whfast.rs:702-707 (warning) */
#define POS_STATUS_KEPLER_BISECTION   (1u << 7) /* whfast.rs:776-808 fallback used */
#define POS_STATUS_ITERATION_CAP      (1u << 8) /* a loop the reference leaves unbounded hit its cap */
#define POS_STATUS_HISTORY_OVERFLOW   (1u << 9) /* a historic snapshot found its buffer full (dropped) */
#define POS_STATUS_FATAL (POS_STATUS_ROCHE | POS_STATUS_COLLISION | POS_STATUS_EJECTION | \
                          POS_STATUS_ZERO_INERTIA | POS_STATUS_ITERATION_CAP)

/*
 * Ensemble description: what the reference's case.json holds for each
 * universe (Universe, Particle and WHFast fields, universe.rs:50-63,
 * particle.rs:16-52, whfast.rs:98-120), flattened to structure-of-arrays.
 * Bodies of system s are body_offset[s] .. body_offset[s+1]-1, in the order
 * of the reference's particle array.  Vector fields hold 3 doubles per body
 * (x, y, z).  Pointers marked "optional" may be NULL.
 */
typedef struct pos_ensemble_desc {
    int32_t n_systems;
    const int32_t *body_offset;              /* [n_systems + 1]                      */

    /* per system */
    const int32_t  *coordinates;             /* POS_COORD_*                          */
    const uint32_t *effects;                 /* POS_EFFECT_* bitmask                 */
    const int32_t  *gr_implementation;       /* POS_GR_*                             */
    const double   *time_step;               /* dt (days)                            */
    const double   *time_limit;              /* universe.time_limit                  */
    const double   *current_time;            /* optional: WHFast.current_time, 0     */
    const double   *historic_snapshot_period;/* optional: default +inf (never)       */
    const double   *last_historic_snapshot_time; /* optional: default -1 (fresh)     */
    const int64_t  *current_iteration;       /* optional: default 0                  */

    /* per body */
    const double   *mass;                    /* M_sun                                */
    const double   *radius;                  /* AU                                   */
    const double   *radius_of_gyration_2;
    const double   *moment_of_inertia;       /* optional: default m*rg2*R^2          */
    const double   *position;                /* [3*nb] inertial (barycentric), AU    */
    const double   *velocity;                /* [3*nb] inertial, AU/day              */
    const double   *spin;                    /* optional [3*nb]; default L/I         */
    const double   *angular_momentum;        /* [3*nb] spin angular momentum         */
    const double   *heliocentric_velocity;   /* optional [3*nb]; only the host's value
                                                is read (stale before the first step,
                                                universe.rs:335-337)                 */
    const uint32_t *roles;                   /* POS_ROLE_* bitmask                   */
    const double   *tides_dissipation_factor;       /* CTL sigma (ConstantTimeLagParameters) */
    const double   *tides_dissipation_factor_scale;
    const double   *tides_love_number;
    const double   *flattening_love_number;  /* OblateSpheroidParameters.love_number */
    const double   *wind_k_factor;           /* optional                             */
    const double   *wind_rotation_saturation;/* optional                             */

    /* integrator state (optional, zero when NULL): compensated-summation
     * errors, whfast.rs:117-119 */
    const double   *velocity_error;          /* [3*nb] */
    const double   *angular_momentum_error;  /* [3*nb] */

    /* evolution tables (optional; all bodies non-evolving when NULL).
     * Rows of body b are evolver_offset[b] .. evolver_offset[b+1]-1 of the
     * flat time/radius/rg2/love arrays; time is relative to initial_time as
     * in case.json (evolution.rs:426-427). */
    const int32_t  *evolution_type;          /* [nb] POS_EVO_*                       */
    const int32_t  *evolver_offset;          /* [nb + 1]                             */
    const double   *evolver_time;
    const double   *evolver_radius;
    const double   *evolver_radius_of_gyration_2;
    const double   *evolver_love_number;
    const int32_t  *evolver_left_index;      /* [nb] Evolver.left_index              */

    /* inverse tidal Q factor column of the same tables (optional; read for
     * BolmontMathis2016, GalletBolmont2017 and LeconteChabrier2013 with
     * dynamical tides, evolution.rs:487-513) */
    const double   *evolver_inverse_tidal_q_factor;
    /* the disk of each system's disk host (optional; required for systems
     * with POS_EFFECT_DISK): 6 doubles per system, the host's DiskProperties
     * (disk.rs:7-14): inner_edge_distance, outer_edge_distance (AU),
     * lifetime (days), alpha, surface_density_normalization,
     * mean_molecular_weight */
    const double   *disk_properties;
    /* the pair-dependent dissipation map entries (optional; NaN or NULL =
     * no entry): for an orbiting body with evolving dissipation, its own
     * factor in its pair with the tidal host, the entry
     * id * MAX_PARTICLES + host_id of universe.pair_dependent_scaled_
     * dissipation_factor (constant_time_lag.rs:96-165, 188-194) */
    const double   *pair_dissipation_factor;   /* [nb] */
    /* Roche radii of a resumed system (optional): the radius of body b's pair
     * with body j of its system at roche_radius[b * POS_MAX_BODIES + j]
     * (universe.roche_radiuses, computed once by initialize_physical_values at
     * t = 0 and carried by the snapshot, universe.rs:177-196, output.rs:187-193).
     * Systems whose current_time is 0 compute them from the radii at t = 0 and
     * ignore this array; a resumed system without it (or a NaN entry) gets
     * them from its current radii. */
    const double   *roche_radius;              /* [nb * POS_MAX_BODIES] */
} pos_ensemble_desc;

/*
 * Mutable state of an ensemble, read by pos_get_state and written by
 * pos_set_state.  Every pointer is optional; NULL fields are skipped.
 * Per-body arrays have the ensemble's total body count (x3 for vectors).
 */
typedef struct pos_state {
    double   *position;              /* [3*nb] inertial                       */
    double   *velocity;              /* [3*nb] inertial                       */
    double   *acceleration;          /* [3*nb] gravity, universe.rs:300-302   */
    double   *angular_momentum;      /* [3*nb]                                */
    double   *spin;                  /* [3*nb]                                */
    double   *velocity_error;        /* [3*nb]                                */
    double   *angular_momentum_error;/* [3*nb]                                */
    double   *radius;                /* [nb] (changes only for evolving bodies)*/
    double   *radius_of_gyration_2;  /* [nb]                                  */
    double   *moment_of_inertia;     /* [nb]                                  */
    double   *current_time;          /* [n_systems]                           */
    double   *last_historic_snapshot_time; /* [n_systems]                     */
    int64_t  *current_iteration;     /* [n_systems]                           */
    uint32_t *status;                /* [n_systems] POS_STATUS_* bits         */
    int32_t  *evolver_left_index;    /* [nb] Evolver.left_index (evolution.rs:26) */
    int32_t  *timestep_warning;      /* [n_systems] WHFast.timestep_warning   */
    double   *pair_dissipation_factor; /* [nb] map entry of an evolving-dissipation
                                          orbiting body, NaN = none (see desc) */
    double   *roche_radius;          /* [nb * POS_MAX_BODIES] the Roche radii in use,
                                        laid out as in the desc               */
} pos_state;

/* Launch mapping chosen by the autotuner or by the caller. */
#define POS_MAPPING_AUTO        0   /* autotuned (pos_tune) or heuristic    */
#define POS_MAPPING_PLAIN       1   /* one system per wavefront, one step per launch */
#define POS_MAPPING_PACKED      2   /* lane groups, K steps per launch      */

typedef struct pos_options {
    int32_t device;            /* HIP device ordinal                          */
    int32_t mapping;           /* POS_MAPPING_*                               */
    int32_t group_size;        /* lanes per system for PACKED (0 = smallest power of two >= N,
                                  -1 = exactly N, > 0 = that many, at least N,
                                  -2 = exactly N with systems of different sizes
                                  sharing a wavefront, best-fit decreasing) */
    int32_t steps_per_launch;  /* K for PACKED (0 = default 64)               */
    int32_t use_graph;         /* capture chunk sequences in a hipGraph       */
    int32_t reserved[3];
} pos_options;

typedef struct pos_ens pos_ens;   /* opaque */

/* ABI version this library was built with. */
int32_t pos_abi_version(void);

/* Thread-local message for the last failing call on this thread. */
const char *pos_last_error(void);

/* Build an ensemble on the device.  Replaces WHFast::new + Universe::new
 * (whfast.rs:143-165, universe.rs:73-175) and, for systems whose
 * current_time is 0, Integrator::initialize_physical_values
 * (whfast.rs:226-233): spin = L/I and Roche radii (universe.rs:177-196). */
int32_t pos_create(const pos_ensemble_desc *desc, const pos_options *opts, pos_ens **out);

/* Advance every system by n_steps calls of WHFast::iterate
 * (whfast.rs:235-305).  Systems whose status has POS_STATUS_COMPLETED or a
 * fatal bit are left unchanged.  Asynchronous with respect to the host;
 * pos_synchronize or pos_get_state waits for completion. */
int32_t pos_step(pos_ens *ens, int64_t n_steps);

/* Wait for all work queued on the handle's stream. */
int32_t pos_synchronize(pos_ens *ens);

/* Copy state device -> host (blocking) / host -> device. */
int32_t pos_get_state(pos_ens *ens, pos_state *out);
int32_t pos_set_state(pos_ens *ens, const pos_state *in);

/* Integrator::set_time_limit / set_snapshot_periods (whfast.rs:187-224),
 * applied to every system; a value <= 0 leaves the field unchanged. */
int32_t pos_set_time_limit(pos_ens *ens, double time_limit);
int32_t pos_set_historic_snapshot_period(pos_ens *ens, double period);

/* Sizes of the ensemble. */
int32_t pos_get_counts(const pos_ens *ens, int32_t *n_systems, int32_t *n_bodies);

/* The HIP stream (hipStream_t) the handle launches on, so that a caller can
 * order its own work or timing events against it.  No counterpart in the
 * reference (which has no device). */
int32_t pos_get_stream(const pos_ens *ens, void **stream);

/* The launch mapping currently in use (after pos_tune or the heuristic). */
int32_t pos_get_options(const pos_ens *ens, pos_options *out);

/* What the handle has executed since pos_create, for a caller that times a
 * region and must say what ran inside it (read before and after).  The
 * autotuner's probe launches are not counted.  No counterpart in the
 * reference (which has no device). */
typedef struct pos_launch_stats {
    int64_t steps;                   /* ensemble steps asked of pos_step          */
    int64_t kernel_launches;         /* step-kernel launches made directly         */
    int64_t graph_replays;           /* hipGraph replays                           */
    int64_t launches_per_replay;     /* step-kernel launches one replay holds      */
    double  last_operator_kernel_ms; /* device time of the kernels of the last
                                        pos_gravity_acceleration or
                                        pos_additional_effects call (HIP events on
                                        the handle's stream), without the copy of
                                        its results                                */
} pos_launch_stats;
int32_t pos_get_launch_stats(const pos_ens *ens, pos_launch_stats *out);

/* Autotuner: times candidate (group size, steps per launch) mappings on a
 * copy of the ensemble for `probe_steps` steps each, keeps the fastest, and
 * restores the state; when the handle was created with use_graph, it then
 * times graph replay against direct launches for that mapping and keeps the
 * faster.  Writes the chosen mapping to *chosen (optional). */
int32_t pos_tune(pos_ens *ens, int32_t probe_steps, pos_options *chosen);

/* Operator-level entry points: the two sub-boundaries the reference
 * benchmarks directly (benches/simulation.rs:127, 138).  They evaluate at
 * the current state without advancing time and copy results to the host.
 *   pos_gravity_acceleration     Universe::gravity_calculate_acceleration
 *                                (universe.rs:198-303) -> out_acc[3*nb]
 *   pos_additional_effects       inertial_to_heliocentric +
 *                                Universe::calculate_additional_effects
 *                                (universe.rs:318-351, 428-538) ->
 *                                out_acc[3*nb], out_dangular_momentum_dt[3*nb]
 * Neither modifies the ensemble state.  Status bits raised by the gravity
 * checks are OR-ed into out_status[n_systems] (optional). */
int32_t pos_gravity_acceleration(pos_ens *ens, double *out_acc, uint32_t *out_status);
int32_t pos_additional_effects(pos_ens *ens, double *out_acc, double *out_dangular_momentum_dt);

/* pos_additional_effects for callers that cannot allocate: code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc.  No
 * counterpart in the reference.  It evaluates the same kernels, but writes
 * into caller-owned DEVICE buffers (caller body order, [nb][3] each, on the
 * handle's device), allocates nothing (no hipMalloc, hipFree or allocating
 * library call; its intermediate results go to a static device scratch array
 * of 2 x 3 x POS_NOALLOC_MAX_BODIES doubles, 24 MiB, that each GPU the library
 * runs on holds until the process ends) and does not synchronize: the
 * results are ready when the handle's stream (pos_get_stream) reaches this
 * point.  Ensembles of more than POS_NOALLOC_MAX_BODIES bodies are rejected
 * with POS_E_LIMIT.  Every call on one GPU shares that GPU's scratch array,
 * so calls that overlap on one GPU (two handles, or two host threads) race. */
#define POS_NOALLOC_MAX_BODIES 524288
int32_t pos_additional_effects_noalloc(pos_ens *ens, double *dev_out_acc, double *dev_out_dangular_momentum_dt);

/*
 * Historic snapshots (src/integrator/output.rs:90-172, written by
 * WHFast::iterate at whfast.rs:240-256).  After pos_history_enable, every
 * step whose snapshot condition fires appends one record per body of its
 * system to a device buffer of `capacity` records per system (a full buffer
 * drops the record and sets POS_STATUS_HISTORY_OVERFLOW).
 * pos_history_drain copies the pending records out (input order: system,
 * then record, then body) and clears the buffer; *n_out receives the count.
 * The fields are those of the reference's history file: lag_angle is the
 * body's evolving-dissipation lag angle at the snapshot (evolution.rs:548-567,
 * 0 for other models), migration_timescale the disk migration timescale of
 * its last disk evaluation (disk.rs:236, 0 without the disk) and denergy_dt
 * (tides/common.rs:263-281) is evaluated at the snapshot state.
 */
typedef struct pos_history_record {
    int32_t system;            /* input system index                      */
    int32_t body;              /* index of the body within its system     */
    double time, time_step;    /* days                                    */
    double position[3];        /* inertial, AU                            */
    double spin[3];            /* rad/day                                 */
    double velocity[3];        /* inertial, AU/day                        */
    double mass, radius, radius_of_gyration_2;
    double love_number, scaled_dissipation_factor, lag_angle, denergy_dt, migration_timescale;
} pos_history_record;
int32_t pos_history_enable(pos_ens *ens, int32_t capacity);
int32_t pos_history_drain(pos_ens *ens, pos_history_record *out, int64_t max_records, int64_t *n_out);

/*
 * Discrete adjoint (the differentiable form).  Supported systems: democratic
 * heliocentric, Jacobi or WHDS coordinates, any of tides / rotational
 * flattening / Kidder 1995 GR, with or without wind and evolution tables
 * (radius and radius of gyration following time), no disk and no evolving
 * dissipation.
 * No counterpart in the reference,
 * which has no derivatives; it differentiates the step of whfast.rs:235-305.
 *
 * pos_adjoint_record advances every system by n_steps like pos_step and
 * keeps a checkpoint of the state every segment_steps steps (0 = 1000).
 * pos_adjoint then returns the vector-Jacobian product of that run: from
 * the cotangents of the final position, velocity and spin angular momentum
 * (seed), the cotangents of the state at the start of the recording and the
 * gradients with respect to every body's CTL dissipation factor sigma
 * (dissipation_factor_scale * dissipation_factor), tidal Love number and
 * fluid Love number (grad).  It recomputes each segment from its checkpoint,
 * runs the reverse sweep over it, and leaves the ensemble's state where
 * pos_adjoint_record left it.  Arrays are in input body order, [3*nb] for
 * vectors and [nb] for parameters; NULL seed arrays count as zero and NULL
 * grad arrays are skipped.  Systems that stop early (completion or a fatal
 * status) are differentiated over the steps they executed.
 */
typedef struct pos_adjoint_seed {
    const double *position;
    const double *velocity;
    const double *angular_momentum;
} pos_adjoint_seed;
typedef struct pos_adjoint_grad {
    double *position;
    double *velocity;
    double *angular_momentum;
    double *dissipation_factor;   /* d/d sigma, sigma = scale * factor */
    double *love_number;          /* d/d tidal k2                      */
    double *fluid_love_number;    /* d/d flattening k2                 */
} pos_adjoint_grad;
int32_t pos_adjoint_record(pos_ens *ens, int64_t n_steps, int32_t segment_steps);
int32_t pos_adjoint(pos_ens *ens, const pos_adjoint_seed *seed, pos_adjoint_grad *grad);

/* Release device memory and the stream. */
void pos_destroy(pos_ens *ens);

#ifdef __cplusplus
}
#endif
#endif /* POSIDONIUS_HIP_H */
