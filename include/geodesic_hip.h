/*
 * geodesic_hip.h -- C-ABI of the MI355X (gfx950) force-evaluation engine.
 *
 * This is the drop-in boundary for GEODESIC-M's `ComputeBackend` trait
 * (geodesic-core/src/backend.rs:11-23; paths are relative to the reference
 * repository root).  A Rust FFI crate binds these symbols and implements the
 * trait over them; the reference's wgpu `GpuBackend`
 * (geodesic-gpu/src/gpu_backend.rs:11-100) is what it replaces.
 *
 * Conventions
 *   - Every function returns gm_status (0 = GM_OK).  gm_last_error() gives a
 *     message for the most recent failure on the calling thread.
 *   - Arrays are plain pointers + sizes, SoA, f64 for positions/forces/
 *     parameters and uint32 for indices, exactly the layouts of SimState
 *     (state.rs:3-23), AtomData (atoms.rs:3-12), BondedTopology
 *     (topology.rs:4-37) and ForceBuffer (buffers.rs:5-9).
 *   - gm_system_create copies every array; the caller keeps ownership.
 *   - One host thread per gm_sys at a time; distinct gm_sys objects may be
 *     used concurrently.  Each gm_sys owns one HIP stream (default priority).
 *   - Functions taking host pointers are synchronous.  The *_device variants
 *     take device pointers (HBM resident) and are ordered on the system's
 *     stream; gm_sync waits for them.
 *   - Results are bit-for-bit reproducible run to run: pair forces are
 *     computed in FP32 and accumulated in int64 fixed point (2^-32 kcal/mol/A
 *     per unit); energies are reduced in f64 in a fixed order.
 */
#ifndef GEODESIC_HIP_H
#define GEODESIC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  1..4 map 1:1 onto geodesic-core BackendError
 * (error.rs:82-91): DeviceLost, ShaderCompilation, OutOfGpuMemory, NoAdapter. */
typedef int gm_status;
enum {
    GM_OK = 0,
    GM_ERR_DEVICE_LOST = 1,   /* BackendError::DeviceLost */
    GM_ERR_KERNEL_LOAD = 2,   /* BackendError::ShaderCompilation (code object missing for this GPU) */
    GM_ERR_OOM = 3,           /* BackendError::OutOfGpuMemory */
    GM_ERR_NO_DEVICE = 4,     /* BackendError::NoAdapter */
    GM_ERR_INVALID_ARG = 5,   /* bad descriptor / index out of range / r_switch<r_cutoff<r_skin violated */
    GM_ERR_KERNEL = 6,        /* a kernel launch or HIP call failed */
    GM_ERR_NOT_BUILT = 7,     /* compute before build_neighbor_list */
    GM_ERR_CAPACITY = 8,      /* output buffer too small / fixed-point or LDS capacity exceeded */
    GM_ERR_CONVERGENCE = 9,   /* ConvergenceError::ConstraintSolver (error.rs): SHAKE/RATTLE did not converge */
    /* 10, 11: host-layer codes (geodesic_host.h) */
    GM_ERR_NUMERICAL = 12,    /* NumericalError::NanInForce / NanInPos (error.rs; geodesic/src/lib.rs:79-107):
                                 a force or position is not finite.  Outputs are still written. */
};

typedef struct gm_ctx gm_ctx; /* one per device */
typedef struct gm_sys gm_sys; /* one per simulated system */

/* Everything GpuBackend::try_new receives (gpu_backend.rs:24-53):
 * AtomData (sigma, epsilon), BondedTopology (bonds, angles, dihedrals,
 * exclusions) and the SimParams fields the force path reads (params.rs:4-23).
 * Extensions with no reference counterpart are marked [ext]. */
typedef struct {
    int64_t n_atoms;
    const double *sigma;   /* per atom, A */
    const double *epsilon; /* per atom, kcal/mol */
    double box[3];         /* SimParams::box_size */
    int32_t periodic;      /* SimParams::periodic: wrap positions at list build */
    double r_cutoff, r_switch, r_skin; /* r_skin is the list radius (topology.rs:53) */
    double scale_14;           /* [ext] LJ scale for pair14_*; 0 = reference behaviour */
    int32_t bonded_min_image;  /* [ext] min-image bonded displacements; 0 = reference raw */
    int64_t n_bonds;
    const uint32_t *bond_i, *bond_j;
    const double *bond_k, *bond_r0;
    int64_t n_angles;
    const uint32_t *angle_i, *angle_j, *angle_k; /* j is the central atom */
    const double *angle_kth, *angle_th0;
    int64_t n_dihedrals;
    const uint32_t *dihed_i, *dihed_j, *dihed_k, *dihed_l;
    const double *dihed_kphi;
    const uint32_t *dihed_n;
    const double *dihed_delta;
    int64_t n_excl;
    const uint32_t *excl_i, *excl_j;
    int64_t n_pairs14;                 /* [ext] scaled 1-4 pairs */
    const uint32_t *pair14_i, *pair14_j;
} gm_system_desc;

/* Per-term potential energies in kcal/mol (geodesic/src/lib.rs:37-45
 * EnergyReport: bond, angle, dihedral, nonbonded, total), plus the [ext]
 * scaled 1-4 term.  total = bond + angle + dihedral + lj14 + lj. */
typedef struct {
    double bond, angle, dihedral, lj, lj14, total;
} gm_energy_terms;

/* Device / context.  Replaces geodesic-gpu/src/device.rs:8-36 (try_new):
 * GM_ERR_NO_DEVICE when no gfx950 device exists at that ordinal. */
gm_status gm_init(int device_ordinal, gm_ctx **out);
gm_status gm_ctx_destroy(gm_ctx *ctx);
/* Message for the most recent failure on the calling thread ("" if none). */
const char *gm_last_error(void);
/* Number of visible HIP devices (0 without a GPU); never fails. */
int gm_device_count(void);

/* GpuBackend::try_new (gpu_backend.rs:24-53): validates and copies the
 * descriptor to the device. */
gm_status gm_system_create(gm_ctx *ctx, const gm_system_desc *desc, gm_sys **out);
gm_status gm_system_destroy(gm_sys *sys);

/* ComputeBackend::build_neighbor_list (backend.rs:15; neighbor.rs:8-55).
 * Host positions; wraps them in place into [0, box) when periodic and
 * records them as the reference positions for needs_rebuild. */
gm_status gm_build_neighbor_list(gm_sys *sys, double *x, double *y, double *z);

/* ComputeBackend::needs_rebuild (neighbor.rs:60-72). */
gm_status gm_needs_rebuild(gm_sys *sys, const double *x, const double *y, const double *z,
                           int32_t *out);

/* ComputeBackend::compute_forces + potential_energy (cpu_backend.rs:92-182,
 * gpu_backend.rs:64-82): forces into caller memory (f64 SoA, length n_atoms),
 * per-term energies into *e (may be NULL). */
gm_status gm_compute_forces(gm_sys *sys, const double *x, const double *y, const double *z,
                            double *fx, double *fy, double *fz, gm_energy_terms *e);

/* Canonical export of the current Verlet list for bit-exact checks: unordered
 * pairs (i<j) sorted i-major, exactly the set neighbor.rs:25-43 produces.
 * Writes min(n, cap) pairs; *n receives the full count. */
gm_status gm_get_pair_list(gm_sys *sys, uint32_t *pair_i, uint32_t *pair_j, int64_t cap,
                           int64_t *n);

/* Device-resident variants: x/y/z/f* are device pointers (length n_atoms).
 * Ordered on the system's stream; energies are delivered by gm_sync. */
gm_status gm_build_neighbor_list_device(gm_sys *sys, double *d_x, double *d_y, double *d_z);
gm_status gm_compute_forces_device(gm_sys *sys, const double *d_x, const double *d_y,
                                   const double *d_z, double *d_fx, double *d_fy, double *d_fz);
/* No counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * gm_compute_forces_device, bit for bit, except that the evaluation's scratch
 * (fixed-point accumulator, non-finite flags, displacement words, energy
 * partials, i-force records) lives in a static device array of 32 MiB the
 * library holds on each GPU until the process ends; it calls no allocator.
 * The system must be built (list, work items) beforehand.  GM_ERR_CAPACITY if
 * the list needs more than 32 MiB (about 200k atoms with the automatic work
 * shape).  All calls on one GPU share that array: calls that overlap on one
 * GPU (two streams or host threads) race. */
gm_status gm_compute_forces_device_noalloc(gm_sys *sys, const double *d_x, const double *d_y, const double *d_z,
                                           double *d_fx, double *d_fy, double *d_fz);
/* ComputeBackend::needs_rebuild at device positions (synchronous). */
gm_status gm_needs_rebuild_device(gm_sys *sys, const double *d_x, const double *d_y, const double *d_z,
                                  int32_t *out);
/* Waits for the system's stream; writes the last evaluation's energies (e may be NULL). */
gm_status gm_sync(gm_sys *sys, gm_energy_terms *e);

/* Resident variants: the system keeps one set of f64 positions and forces in
 * HBM (atom order).  gm_build_neighbor_list / gm_compute_forces leave their
 * (wrapped) positions there; gm_set_positions replaces them from the host.
 * gm_compute_forces_resident evaluates at the resident positions into the
 * resident forces, ordered on the system's stream; gm_get_forces copies the
 * forces out (synchronous). */
gm_status gm_set_positions(gm_sys *sys, const double *x, const double *y, const double *z);
gm_status gm_compute_forces_resident(gm_sys *sys);
gm_status gm_get_forces(gm_sys *sys, double *fx, double *fy, double *fz);

/* Device memory helpers for callers without their own HIP binding (e.g. an
 * FFI crate): allocate / free / copy on the context's device.  Copies are
 * synchronous. */
gm_status gm_device_alloc(gm_ctx *ctx, size_t bytes, void **out);
gm_status gm_device_free(gm_ctx *ctx, void *p);
gm_status gm_copy_to_device(gm_ctx *ctx, void *dst, const void *src, size_t bytes);
gm_status gm_copy_to_host(gm_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Statistics of the current list and the last evaluation. */
typedef struct {
    int64_t n_atoms;
    int64_t n_superclusters;   /* 64-slot i-groups */
    int64_t n_clusters;        /* 8-slot groups */
    int64_t n_entries;         /* (supercluster, j-cluster) list entries */
    int64_t n_tiles;           /* (i-cluster, j-cluster) 8x8 tiles with a nonzero mask */
    int64_t n_tiles_cutoff;    /* of those, tiles with a pair within r_cutoff at build time */
    int64_t n_list_pairs;      /* pairs in the Verlet list */
    int64_t n_pairs_cutoff;    /* of those, pairs with r <= r_cutoff (f64) at the build positions */
    int64_t n_work_items;      /* waves launched by the nonbonded kernel */
    float ms_nonbonded;        /* nonbonded kernel of the last profiled evaluation (hipEvent) */
    float ms_eval;             /* last profiled evaluation, whole (device time) */
    float ms_build;            /* last list build (device time) */
    int64_t n_partner_entries; /* partner-list layout: list entries, padding included (0 for tiles) */
    int32_t pair_layout;       /* GM_LAYOUT_TILES or GM_LAYOUT_PARTNERS: the layout evaluations use */
    int32_t n_lj_types;        /* distinct (sigma, epsilon) pairs, 0 when more than 256 */
} gm_stats;
gm_status gm_get_stats(gm_sys *sys, gm_stats *out);

/* Per-evaluation device timings between gm_profile_begin and gm_profile_end:
 * the first evaluation and every `every`-th one after it are timed (at most
 * 8192; later ones are counted in n_untimed), the others record nothing.
 * Timed evaluations add no event packets to the stream: the kernels'
 * dispatches stamp the HIP events themselves (the prep kernel's start, the
 * nonbonded kernel's start and end, the finish kernel's end).  Outside a
 * profiling session an evaluation records no events at all. */
typedef struct {
    int64_t n_evals;            /* evaluations timed */
    int64_t n_untimed;          /* evaluations past the event pool */
    double ms_nonbonded_total;  /* sum over timed evaluations, nonbonded kernel (its dispatch's own
                                   start and end stamps) */
    double ms_nonbonded_min, ms_nonbonded_max;
    double ms_eval_total;       /* sum over timed evaluations, prep kernel's start to the finish
                                   kernel's end (the accumulator clear before it is not included) */
    double ms_prep_total;       /* sum over timed evaluations, prep kernel's start to the nonbonded
                                   kernel's start */
} gm_profile;
gm_status gm_profile_begin(gm_sys *sys, int32_t every);
gm_status gm_profile_end(gm_sys *sys, gm_profile *out);

/* Tuning: list entries per nonbonded work item (one wavefront each, at most
 * 32); 0 picks a size that gives about 1.5 work items per resident wavefront
 * slot.  Work items are issued heaviest first.  Items of fewer than 16
 * entries leave their i forces as per-item records that the finish kernel
 * adds; larger ones use atomics (same result either way).  Takes effect
 * immediately on a built list (no rebuild). */
gm_status gm_set_work_chunk(gm_sys *sys, int32_t entries_per_item);

/* Tile pruning (default on): each evaluation bounds how far any atom has moved
 * since the list build, and skips the tiles whose closest listed pair at
 * build time must now lie beyond r_cutoff.  Those tiles would add exactly
 * zero, so forces and energies are bit-identical either way; 0 turns the
 * test off (every listed tile is evaluated).  Pruning applies only when every
 * box edge is at least 2 r_skin.  No reference counterpart (the reference
 * tests every listed pair, neighbor.rs / nonbonded.rs:38-41).  A batch reads
 * its replicas' setting when its lists are built. */
gm_status gm_set_tile_pruning(gm_sys *sys, int32_t enable);

/* Tuning: the layout of the nonbonded list that evaluations walk.
 *   GM_LAYOUT_PARTNERS: per atom, its partners within r_skin, those within
 *     r_cutoff + 0.1 A at the build first; four atoms share a wavefront (16
 *     lanes each), LJ parameters come from a per-type table in LDS, every
 *     pair is computed once from each side with the same arithmetic (so the
 *     two forces are exact negatives) and its energy halved; no atomics.
 *     With pruning on (gm_set_tile_pruning) and every atom less than about
 *     0.05 A from its build position, the second part of each list is
 *     skipped (it lies beyond r_cutoff; results bit-identical).  Entries are
 *     16-bit codes relative to the atoms' group when the system's geometry
 *     allows (neighbouring cells within 32,767 slots), else 32-bit slot
 *     numbers.  Needs at most 256 distinct (sigma, epsilon) pairs: a system
 *     with more uses tiles whatever is set.
 *   GM_LAYOUT_TILES: 8x8 tiles of clusters of 8 (one wavefront per tile row
 *     of 8 i clusters), each pair once; tiles whose closest pair has moved
 *     beyond r_cutoff are skipped, gradually as the atoms move.
 *   GM_LAYOUT_AUTO (the default): partner lists, unless the first build finds
 *     that they need 32-bit slot numbers; such a system (about 250,000 atoms
 *     and more at liquid density) evaluates tiles from then on, which are
 *     faster there once the atoms have moved.
 * Both evaluate the same listed pairs in f32 with fixed-point force sums;
 * results agree to f32 rounding, not bit for bit (the order of operations
 * inside a pair differs).  The pair list, its statistics and shell-mode
 * evaluations always read the tile list: with partner lists a build leaves
 * it out and it is made when first asked for.  Takes effect at once on a
 * built list.  A batch's replicas always use tiles (GM_ERR_INVALID_ARG for a
 * replica).  gm_stats.pair_layout says which layout evaluations use.  No
 * reference counterpart (the reference walks a flat pair list, neighbor.rs /
 * nonbonded.rs:38-41). */
enum { GM_LAYOUT_TILES = 0, GM_LAYOUT_PARTNERS = 1, GM_LAYOUT_AUTO = 2 };
gm_status gm_set_pair_layout(gm_sys *sys, int32_t layout);

/* ---- Device-resident dynamics ----------------------------------------------
 * integrator.rs:14-71 (half_kick, drift_and_constrain), constraint.rs:62-259
 * (solve, constrain_velocities) and the run loop of geodesic/src/lib.rs:255-330,
 * on device arrays, in f64 with the reference's operation order: given the
 * same forces every step is bit-for-bit the CPU path's.  Masses and
 * holonomic constraints (BondedTopology constr_i/constr_j/constr_dsq,
 * topology.rs:26-32) are attached to a system once. */
gm_status gm_dynamics_set(gm_sys *sys, const double *mass, int64_t n_constr, const uint32_t *constr_i,
                          const uint32_t *constr_j, const double *constr_dsq);
/* Tuning: the SHAKE / RATTLE solver.  When every connected component of the
 * constraint graph has at most 256 atoms and 256 constraints (X-H bond
 * constraints, rigid waters), the components are packed into groups that
 * one workgroup each iterates in LDS: a solve is two launches and one read-
 * back.  Otherwise (enable 0, or a larger component) every iteration is a
 * launch over all atoms.  iteration_cap > 0 stops the group iterations after
 * that many, leaving the rest to the per-iteration launches.  Same
 * iterations, stopping point and arithmetic (hence results) every way.
 * *n_groups (may be NULL) receives the group count, 0 when there are none.
 * No reference counterpart (constraint.rs:62-259 iterates every constraint). */
gm_status gm_set_constraint_groups(gm_sys *sys, int32_t enable, int32_t iteration_cap, int32_t *n_groups);
/* v += dt_half F / m * 20.455^2 (ordered on the system's stream). */
gm_status gm_half_kick_device(gm_sys *sys, double *d_vx, double *d_vy, double *d_vz, const double *d_fx,
                              const double *d_fy, const double *d_fz, double dt_half);
/* r += dt v, SHAKE projection, v = (r - r_old) / dt.  *iters (may be NULL)
 * receives the solver's iteration count; GM_ERR_CONVERGENCE after max_iter. */
gm_status gm_drift_and_constrain_device(gm_sys *sys, double *d_x, double *d_y, double *d_z, double *d_vx,
                                        double *d_vy, double *d_vz, double dt, int32_t max_iter, double tol,
                                        int64_t step, int32_t *iters);
/* RATTLE velocity projection (constraint.rs:159-259). */
gm_status gm_constrain_velocities_device(gm_sys *sys, const double *d_x, const double *d_y, const double *d_z,
                                         double *d_vx, double *d_vy, double *d_vz, int32_t max_iter, double tol,
                                         int64_t step, int32_t *iters);
typedef struct {
    int64_t n_steps;         /* SimParams::n_steps (of this call) */
    double dt;               /* ps */
    int32_t frame_interval;  /* steps between frames */
    int32_t max_iter;        /* constraint iterations */
    double tol;              /* constraint tolerance */
    int64_t first_step;      /* [ext] step number of this call's first step (reports, frame schedule) */
    int32_t resume;          /* [ext] 1: continue the previous call's run (no list build, initial forces,
                                velocity projection or frame at first_step) */
} gm_md_desc;
/* The reference run loop on the device (geodesic/src/lib.rs:255-330): list,
 * forces and velocity projection, then per step B, A (drift + SHAKE), list
 * rebuild when needs_rebuild fires, forces, B, RATTLE.  A frame is taken at
 * the start (unless resuming) and after every step s with (s + 1) %
 * frame_interval == 0 (s counted from first_step): the f32 positions (3 x n:
 * x block, y block, z block) go to h_frames, the potential and kinetic energy
 * (kcal/mol) to h_epot / h_ekin (each may be NULL).  *n_frames receives the
 * count, also when the run stops early.  Positions and velocities advance in
 * place.  A long run can be cut into calls that end on frame steps (resume =
 * 1, first_step advanced), with the same results as one call.
 * Failures stop the run at the step where the reference stops:
 *   - GM_ERR_NUMERICAL: a position not finite after the drift of step s
 *     (NanInPos, lib.rs:94-107, checked at :296) or a force not finite after
 *     the evaluation of step s (NanInForce, lib.rs:79-92, checked at :269 and
 *     :307); gm_md_last_failure gives step, atom and component;
 *   - GM_ERR_CONVERGENCE: SHAKE or RATTLE did not converge. */
gm_status gm_md_run(gm_sys *sys, double *d_x, double *d_y, double *d_z, double *d_vx, double *d_vy,
                    double *d_vz, const gm_md_desc *md, float *h_frames, double *h_epot, double *h_ekin,
                    int64_t *n_frames);
/* The NumericalError of the last gm_md_run on this system: kind 0 none,
 * 1 NanInPos, 2 NanInForce; component 0 X, 1 Y, 2 Z (error.rs NumericalError).
 * An atom whose force is not finite holds NaN in all three components. */
typedef struct {
    int32_t kind;
    int32_t component;
    int64_t step;
    int64_t atom;
} gm_md_failure;
gm_status gm_md_last_failure(gm_sys *sys, gm_md_failure *out);
/* Jacobi-metric factor 2 (E - V(x)) in f64 for total energy E (SimParams
 * total_energy, params.rs:21-22) and the potential of the last evaluation. */
gm_status gm_jacobi_factor(gm_sys *sys, double total_energy, double *out);

/* ---- Batched replicas [ext: no reference counterpart] ---------------------
 * Many independent systems (own atom counts, boxes, topologies, parameters)
 * evaluated by one nonbonded launch replayed from a captured hipGraph, for
 * landscape mapping.  Replica boxes must be cubic.  Positions/forces are
 * concatenated in replica order.
 * Each replica's forces and energies are bit-identical to evaluating it alone
 * in the same pair layout (no work shape changes a bit). */
typedef struct gm_batch gm_batch;
gm_status gm_batch_create(gm_ctx *ctx, const gm_system_desc *descs, int32_t n_replicas,
                          gm_batch **out);
gm_status gm_batch_destroy(gm_batch *b);
/* Borrowed handle of replica r (owned by the batch): pair lists, stats. */
gm_status gm_batch_replica(gm_batch *b, int32_t r, gm_sys **out);
/* Builds every replica's list (wrapping positions in place when periodic),
 * then the descriptor table, work queue and graph. */
gm_status gm_batch_build_neighbor_lists(gm_batch *b, double *x, double *y, double *z);
gm_status gm_batch_set_positions(gm_batch *b, const double *x, const double *y, const double *z);
/* One graph launch: every replica's forces and energies (asynchronous). */
gm_status gm_batch_compute(gm_batch *b);
/* Waits; e receives n_replicas per-replica f64 energy breakdowns (may be
 * NULL); ms the device time of the last gm_batch_compute (may be NULL). */
gm_status gm_batch_sync(gm_batch *b, gm_energy_terms *e, float *ms);
gm_status gm_batch_get_forces(gm_batch *b, double *fx, double *fy, double *fz);
/* Pair layout of every replica (GM_LAYOUT_*; applied at once when built,
 * else at the build).  GM_LAYOUT_AUTO (the default): partner lists (one
 * launch over every replica's groups) when every replica's lists hold 16-bit
 * codes, else tiles (the persistent tile kernel) for all; decided by the lists
 * alone.  GM_LAYOUT_PARTNERS: GM_ERR_INVALID_ARG at the build when a replica's
 * lists would need 32-bit codes. */
gm_status gm_batch_set_pair_layout(gm_batch *b, int32_t layout);
/* The layout the current plan evaluates: GM_LAYOUT_PARTNERS or GM_LAYOUT_TILES. */
gm_status gm_batch_get_pair_layout(gm_batch *b, int32_t *layout);
/* Tile layout work shape: list entries per tile work item (0 = automatic, at
 * most 32) and persistent wavefronts per compute unit.  Speed only: i forces
 * are rounded per fixed-point block of 8 entries, so every shape gives the
 * same bytes. */
gm_status gm_batch_set_shape(gm_batch *b, int32_t entries_per_item, int32_t waves_per_cu);
/* Tile layout: times a grid of shapes on the current lists and keeps the
 * fastest.  Partner lists have no shape: the step is timed and 0, 0 returned. */
gm_status gm_batch_autotune(gm_batch *b, int32_t *entries_per_item, int32_t *waves_per_cu,
                            float *best_ms);

/* ---- Shell-force surrogate and 4:1 multiple-time-step driver -------------
 * [ext: no reference counterpart]
 *
 * The exact model is the system described by `exact` (its r_cutoff bounds the
 * feature window).  The surrogate also creates an inner model: the same
 * atoms and topology with LJ cut at r_inner_cutoff, switched from
 * r_inner_switch, on its own list of radius r_inner_skin.  Shell force =
 * exact - inner.
 *
 * Per atom, 96 radial features: 24 Gaussians (centres evenly spaced over
 * [r_lo, r_hi], width = spacing) times 4 atom classes, faded to zero over
 * `fade` A at both window edges.  They pass through a 96 -> 128 -> 128 -> 3
 * SiLU network (BF16 weights and activations on MFMA, FP32 accumulate).  The
 * three outputs o weight three equivariant vectors, so that at an inner step
 *   F = F_inner(t) + (1 + o1) F_shell(t0) - o2 s2 K dx + o3 s3 (F_inner(t) - F_inner(t0)),
 * with t0 the last exact step, dx the displacement since t0, and K the shell
 * potential's Hessian with respect to the atom's own position at t0 (summed
 * over its shell pairs, computed with the features).  o = 0 is the plain
 * multiple-time-step method (shell force held from t0); o = (0, 1/s2, 0) is
 * the first-order Taylor correction in the atom's own displacement.
 *
 * Parameters: 29315 fp32 values, W1[128][96], b1[128], W2[128][128], b2[128],
 * W3[3][128], b3[3]; feature standardisation mean[96], inv_std[96]; scales
 * {s2, s3, sigma} (sigma normalises forces in the training loss). */
typedef struct gm_surrogate gm_surrogate;
enum { GM_SUR_NPARAM = 29315, GM_SUR_NFEAT = 96 };
typedef struct {
    double r_inner_cutoff, r_inner_switch, r_inner_skin; /* inner model */
    double r_lo, r_hi, fade;  /* feature window and edge fade, A */
    const uint8_t *atom_class; /* per atom, 0..3 */
    uint64_t seed;             /* weight initialisation */
} gm_surrogate_desc;
gm_status gm_surrogate_create(gm_ctx *ctx, const gm_system_desc *exact, const gm_surrogate_desc *d,
                              gm_surrogate **out);
gm_status gm_surrogate_destroy(gm_surrogate *s);
/* Borrowed handle of the exact (which = 0) or inner (which = 1) system. */
gm_status gm_surrogate_system(gm_surrogate *s, int32_t which, gm_sys **out);
/* Builds both lists at device positions (wrapped in place when periodic). */
gm_status gm_surrogate_build(gm_surrogate *s, double *d_x, double *d_y, double *d_z);
/* Features at device positions (lists built within their skin); h_out
 * (n x 96, may be NULL) receives a copy. */
gm_status gm_surrogate_features(gm_surrogate *s, const double *d_x, const double *d_y,
                                const double *d_z, float *h_out);
/* Shell Hessian K (n x 6: xx, yy, zz, xy, xz, yz) of the last
 * gm_surrogate_features call: for each atom, the sum over its non-excluded
 * pairs of the Hessian of the shell pair potential. */
gm_status gm_surrogate_get_hessian(gm_surrogate *s, float *h_out);
/* Any NULL argument leaves that part unchanged. */
gm_status gm_surrogate_set_params(gm_surrogate *s, const float *params, const float *fmean,
                                  const float *finv, const double *scales);
gm_status gm_surrogate_get_params(gm_surrogate *s, float *params, float *fmean, float *finv,
                                  double *scales);
/* Network outputs (n x 3) for host features (n x 96). */
gm_status gm_surrogate_infer(gm_surrogate *s, const float *h_feat, int64_t n, float *h_out);
/* Training loss and its gradient (GM_SUR_NPARAM) for host features (n x 96),
 * normalised basis vectors (n x 9: B1, B2, B3) and targets (n x 3):
 * loss = (1/n) sum |(1 + o1) B1 + o2 B2 + o3 B3 - T|^2.  No update. */
gm_status gm_surrogate_loss_grad(gm_surrogate *s, const float *h_feat, const float *h_basis,
                                 const float *h_target, int64_t n, double *loss, float *h_grad);
typedef struct {
    int64_t n_frames;       /* inner-step frames to train on (3 per exact step) */
    double mu, sigma, cap;  /* Brownian step dx = mu F + sigma xi per coordinate, clamped to +-cap (A) */
    uint64_t seed;          /* trajectory seed */
    float lr;               /* Adam learning rate */
    int64_t batch;          /* atoms per frame (contiguous, rotating) */
    int32_t update;         /* 0: evaluate only (no parameter change) */
} gm_train_desc;
typedef struct {
    double loss_first, loss_last;  /* network loss of the first / last frame */
    double loss_net, loss_stale;   /* mean over frames: network, and o = 0 */
    double loss_taylor;            /* mean over frames: o = (0, 1/s2, 0) */
    double ms_total;               /* wall time */
    int64_t frames, steps;
} gm_train_report;
/* Runs a seeded Brownian trajectory from the device positions d_x/d_y/d_z
 * (advanced in place, driven by the exact force): inner and shell forces at
 * every step, features at every 4th; each inner step is one training frame
 * (one Adam step over `batch` atoms, target = the shell force).  Lists are
 * rebuilt when either model's needs_rebuild criterion fires (checked at every
 * exact step).  The first call also fixes the feature standardisation and
 * basis scales from the first frames.  h_loss (3 per frame: network, o = 0,
 * Taylor; may be NULL) receives the loss history. */
gm_status gm_surrogate_train(gm_surrogate *s, double *d_x, double *d_y, double *d_z,
                             const gm_train_desc *td, float *h_loss, gm_train_report *rep);
/* 4:1 multiple-time-step force provider.  Step k % 4 == 0: exact forces
 * (inner + shell, was_exact = 1) and, except in GM_MTS_HELD mode, features
 * and shell Hessian (and network coefficients in GM_MTS_NET mode) are
 * refreshed.  Otherwise: inner forces plus the shell force predicted by the
 * mode: held from the last exact step, network, or Taylor correction.
 * Forces go to device arrays; ordered on the surrogate's stream (gm_sync on
 * the exact system waits). */
enum { GM_MTS_HELD = 0, GM_MTS_NET = 1, GM_MTS_TAYLOR = 2 };
gm_status gm_mts_reset(gm_surrogate *s);
gm_status gm_mts_step(gm_surrogate *s, const double *d_x, const double *d_y, const double *d_z,
                      double *d_fx, double *d_fy, double *d_fz, int32_t mode, int32_t *was_exact);
/* Seeded Gaussian displacement of every coordinate of device positions. */
gm_status gm_surrogate_random_walk(gm_surrogate *s, double *d_x, double *d_y, double *d_z,
                                   double sigma, uint64_t seed, uint64_t step);
/* One Brownian step of device positions under device forces d_f*:
 * dx = mu F + sigma xi per coordinate, clamped to +-cap (A). */
gm_status gm_surrogate_bd_step(gm_surrogate *s, double *d_x, double *d_y, double *d_z, const double *d_fx,
                               const double *d_fy, const double *d_fz, double mu, double sigma, double cap,
                               uint64_t seed, uint64_t step);
/* Steepest descent on the exact model (dx = mu F, clamped to +-cap), with
 * list rebuilds as needed; leaves both lists built.  *max_force (may be
 * NULL) receives the largest force component afterwards. */
gm_status gm_surrogate_minimize(gm_surrogate *s, double *d_x, double *d_y, double *d_z, int32_t steps,
                                double mu, double cap, double *max_force);
/* Shell force alone (exact system in shell mode) into device arrays; ordered
 * on the surrogate's stream. */
gm_status gm_surrogate_shell_forces(gm_surrogate *s, const double *d_x, const double *d_y, const double *d_z,
                                    double *d_fx, double *d_fy, double *d_fz);
/* Rebuilds both lists if either model's needs_rebuild criterion fires. */
gm_status gm_surrogate_maybe_rebuild(gm_surrogate *s, double *d_x, double *d_y, double *d_z,
                                     int32_t *rebuilt);

#ifdef __cplusplus
}
#endif
#endif /* GEODESIC_HIP_H */
