// Device-resident dynamics: argument structs and launchers shared by
// gm_dynamics.hip (kernels) and gm_dynamics.cpp (host side).
#pragma once

#include "gm_internal.hpp"

namespace gm {
namespace dyn {

struct Vec3 {
    double *x, *y, *z;
};
struct Vec3C {
    const double *x, *y, *z;
};

// Masses and constraints of one system, with each atom's incident
// constraints in ascending constraint order (CSR).
struct DynArgs {
    int64_t n;
    const double *mass;
    const uint32_t *ci, *cj;
    const double *dsq;
    const int64_t *inc_off;  // n + 1
    const int64_t *inc_con;  // 2 n_constr
};

// Constraint groups: the connected components of the constraint graph, packed
// whole into groups of at most kGroupAtoms atoms and kGroupCons constraints.
// The Jacobi iterations of different components share no atom, so one
// workgroup can iterate its group in LDS, alone, with the reference's
// arithmetic; only the stopping point is global (the first iteration whose
// largest multiplier over all constraints is below the tolerance).
constexpr int kGroupAtoms = 256;
constexpr int kGroupCons = 256;
struct GroupArgs {
    int n_groups;
    const int *atom_off;       // n_groups + 1, into atoms and inc_off
    const uint32_t *atoms;     // atom index, ascending within a group
    const int *con_off;        // n_groups + 1, into cons
    const uint32_t *cons;      // constraint index, ascending within a group
    const uint32_t *con_loc;   // per group constraint: local i | local j << 16
    const int *inc_off;        // per group atom + 1: its incident constraints in inc
    const uint16_t *inc;       // local constraint << 1 | (the atom is the constraint's j)
    const uint32_t *free_atoms;  // atoms in no constraint
    int64_t n_free;
    int *iters;                // per group: iterations done
    int *conv;                 // per group: converged at iters
    int *ctl;                  // [0] largest iters over groups, [1] some group not converged at ctl[0]
    int *ctl_next;             // the next solve's words, cleared by phase 1 (no fill launch)
};

// A guard (nullable) makes a kernel do nothing while *guard is nonzero: the
// run loop launches a step before it knows whether the previous velocity
// projection needs more iterations, and replays the step if it does.

hipError_t launch_half_kick(int64_t n, double *vx, double *vy, double *vz, const double *fx, const double *fy,
                            const double *fz, const double *mass, double dt_half, hipStream_t st,
                            const int *guard = nullptr);
hipError_t launch_drift(int64_t n, double *x, double *y, double *z, const double *vx, const double *vy,
                        const double *vz, double *rx, double *ry, double *rz, double dt, hipStream_t st,
                        const int *guard = nullptr);
hipError_t launch_resync(int64_t n, const double *x, const double *y, const double *z, const double *rx,
                         const double *ry, const double *rz, double *vx, double *vy, double *vz, double dt,
                         hipStream_t st, const int *guard = nullptr);
// One iteration `it` (maxslot: the per-iteration maxima, from iteration 0);
// it first makes the convergence test of iteration it - 1.
hipError_t launch_shake_iter(const DynArgs &D, Vec3 in, Vec3 out, Vec3C ref, unsigned long long *maxslot, int it,
                             double tol, double *lambda, int *flag, hipStream_t st);
hipError_t launch_rattle_iter(const DynArgs &D, Vec3C pos, Vec3 in, Vec3 out, unsigned long long *maxslot, int it,
                              double tol, double *lambda, int *flag, hipStream_t st);
hipError_t launch_check(const unsigned long long *maxslot, double tol, int it, int *flag, hipStream_t st);
// Group solve of SHAKE (velocity = false: buf positions, other the reference
// positions) or RATTLE (velocity = true: buf velocities, other positions).
// Phase 0: each group iterates from buf until its own largest multiplier is
// below tol, or `cap` iterations; ctl[0] gets the largest count.  Phase 1:
// every group continues to ctl[0] iterations and ctl[1] is raised if any is
// not below tol there; atoms in no constraint get the reference's + 0.0.
hipError_t launch_group_solve(const GroupArgs &G, const DynArgs &D, bool velocity, Vec3 buf, Vec3C other, int cap,
                              double tol, int phase, double *lambda, hipStream_t st, const int *guard = nullptr);
hipError_t launch_kinetic_terms(int64_t n, const double *vx, const double *vy, const double *vz,
                                const double *mass, double *out, hipStream_t st);

}  // namespace dyn
}  // namespace gm
