/*
 * geodesic_host.h -- C-ABI of libgeodesic_host.so, the host layer above the
 * force engine (C++ API: geodesic-m_amd/host/geodesic.hpp).
 *
 * It carries the reference's callers and data formats either side of the
 * force path (paths relative to the reference repository root):
 *   - prmtop / inpcrd reading       geodesic-io/src/prmtop.rs:182-344, inpcrd.rs:12-97
 *   - trajectory and energy log     geodesic-io/src/dcd.rs, export.rs:20-55
 *   - `geodesic energy`             geodesic/src/lib.rs:113-171 (energy_from_files)
 *   - `geodesic run`                geodesic/src/lib.rs:187-330 (run, with a parsed config)
 *   - the ComputeBackend trait      geodesic-core/src/backend.rs:11-23 (GpuBackend)
 * all over the GPU backend of geodesic_hip.h.  Status codes are gm_status
 * (geodesic_hip.h) plus the host error kinds of geodesic-core/src/error.rs
 * below; gmh_last_error() gives the message of the last failure on the
 * calling thread.  Extensions with no reference counterpart are marked [ext].
 */
#ifndef GEODESIC_HOST_H
#define GEODESIC_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "geodesic_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GMH_ERR_IO = 10,        /* IoError */
    GMH_ERR_CONFIG = 11,    /* ConfigError (parse errors, invalid values) */
    GMH_ERR_NUMERICAL = GM_ERR_NUMERICAL, /* NumericalError (NaN positions or forces, energy drift with action Stop) */
};

const char *gmh_last_error(void);

/* ---- prmtop / inpcrd ------------------------------------------------------
 * fidelity = 0: prmtop.rs as is.  fidelity = 1 [ext]: 1-4 pairs of proper
 * dihedrals whose 1-4 flag is clear, scaled by 1/SCNB (SCNB_SCALE_FACTOR,
 * default 2.0); SCEE read but unused (no electrostatics in this force field). */
typedef struct gmh_topology gmh_topology;
gm_status gmh_prmtop_read(const char *path, int32_t fidelity, gmh_topology **out);
gm_status gmh_topology_free(gmh_topology *t);
/* Field of the parsed AtomData / BondedTopology: *data points into the
 * object (valid until it is freed or changed), *count is the length.
 * Element types: f64 for GMH_F_SIGMA..GMH_F_CHARGE, bond/angle/dihedral
 * constants and GMH_F_CONSTR_DSQ; uint8 (Element: H C N O S Unknown = 0..5)
 * for GMH_F_ELEMENT; uint32 for the rest. */
enum {
    GMH_F_SIGMA = 0, GMH_F_EPSILON, GMH_F_MASS, GMH_F_CHARGE, GMH_F_ELEMENT, GMH_F_RESIDUE,
    GMH_F_BOND_I, GMH_F_BOND_J, GMH_F_BOND_K, GMH_F_BOND_R0,
    GMH_F_ANGLE_I, GMH_F_ANGLE_J, GMH_F_ANGLE_K, GMH_F_ANGLE_KTH, GMH_F_ANGLE_TH0,
    GMH_F_DIHED_I, GMH_F_DIHED_J, GMH_F_DIHED_K, GMH_F_DIHED_L, GMH_F_DIHED_KPHI, GMH_F_DIHED_N,
    GMH_F_DIHED_DELTA,
    GMH_F_CONSTR_I, GMH_F_CONSTR_J, GMH_F_CONSTR_DSQ,
    GMH_F_EXCL_I, GMH_F_EXCL_J,
    GMH_F_PAIR14_I, GMH_F_PAIR14_J, /* [ext] */
    GMH_F_COUNT
};
gm_status gmh_topology_field(const gmh_topology *t, int32_t field, const void **data, int64_t *count);
/* [ext] 1 / SCNB applied to the 1-4 pairs (0 without pairs) and SCEE as read. */
gm_status gmh_topology_scales(const gmh_topology *t, double *scale_14, double *scee);
/* constraint.rs:10-40: bonds touching a hydrogen become constraints. */
gm_status gmh_topology_promote_hydrogen_bonds(gmh_topology *t);
/* inpcrd.rs:12-84: positions and velocities (A/ps), 3 x n each (x, y, z blocks). */
gm_status gmh_inpcrd_read(const char *path, int64_t n_atoms, int32_t has_box, double *pos, double *vel);

/* ---- trajectory output ----------------------------------------------------
 * dcd.rs: frames is n_frames x (x block, y block, z block) x natom f32. */
gm_status gmh_dcd_write(const char *path, int64_t natom, int32_t frame_interval, double dt_ps, int64_t n_frames,
                        const float *frames);
/* export.rs:20-55: header and one row per entry. */
gm_status gmh_energy_log_write(const char *path, int64_t n_atoms, int64_t n_rows, const uint64_t *step,
                               const double *time_ps, const double *potential_kcal, const double *kinetic_kcal);

/* ---- geodesic energy (lib.rs:37-45, 113-171) ------------------------------ */
typedef struct {
    double bond, angle, dihedral, nonbonded, total;
    double lj14;               /* [ext] scaled 1-4 LJ (fidelity; 0 otherwise) */
    double r_switch, r_cutoff; /* the cutoffs used (10 / 12 A) */
} gmh_energy_report;
gm_status gmh_energy_from_files(const char *prmtop_path, const char *inpcrd_path, int32_t fidelity, int32_t device,
                                gmh_energy_report *out);

/* ---- geodesic run (config.rs:102-128, lib.rs:187-330) ---------------------
 * File paths as given (already resolved).  backend: "gpu" or "hybrid". */
typedef struct {
    int64_t n_steps;
    int32_t frame_interval;
    const char *backend;
    int32_t n_threads;
    const char *prmtop, *inpcrd;
    double box_size[3];
    int32_t periodic;
    double dt, total_energy;
    double r_cutoff, r_skin, r_switch;
    int32_t max_constr_iter;
    double constr_tol;
    int32_t constrain_hydrogens;
    const char *trajectory, *energy_log;
    int32_t monitoring;                 /* MonitoringConfig present */
    double energy_drift_threshold_kcal;
    int32_t energy_drift_stop;          /* DriftAction::Stop (else Warn) */
    int32_t fidelity;                   /* [ext] prmtop 1-4 flags and SCNB */
    int32_t device;
} gmh_run_config;
typedef struct {
    int64_t n_steps, n_frames;
    double final_potential, final_kinetic;
} gmh_run_summary;
gm_status gmh_run(const gmh_run_config *cfg, gmh_run_summary *out);

/* ---- the ComputeBackend trait (backend.rs:11-23) over GpuBackend -----------
 * The handle owns the backend and one SimState (state.rs:3-23). */
typedef struct gmh_backend gmh_backend;
/* GpuBackend::try_new with the SimParams fields the force path and the
 * constraint solver read. */
gm_status gmh_backend_create(const gmh_topology *t, const double box[3], int32_t periodic, double r_cutoff,
                             double r_skin, double r_switch, int32_t max_constr_iter, double constr_tol,
                             int32_t bonded_min_image, int32_t device, gmh_backend **out);
gm_status gmh_backend_free(gmh_backend *b);
/* The state's positions and velocities, 3 x n each (x, y, z blocks); step. */
gm_status gmh_backend_set_state(gmh_backend *b, const double *pos, const double *vel, uint64_t step);
gm_status gmh_backend_get_state(const gmh_backend *b, double *pos, double *vel);
gm_status gmh_backend_build_neighbor_list(gmh_backend *b);
/* compute_forces + potential_energy: forces 3 x n, per-term energies. */
gm_status gmh_backend_compute_forces(gmh_backend *b, double *forces, gm_energy_terms *e);
gm_status gmh_backend_needs_rebuild(const gmh_backend *b, int32_t *out);
/* geodesic_drift: GM_ERR_CONVERGENCE as ConvergenceError. */
gm_status gmh_backend_geodesic_drift(gmh_backend *b, double dt);
gm_status gmh_backend_n_threads(const gmh_backend *b, int64_t *out);

#ifdef __cplusplus
}
#endif

#endif /* GEODESIC_HOST_H */
