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

hipError_t launch_half_kick(int64_t n, double *vx, double *vy, double *vz, const double *fx, const double *fy,
                            const double *fz, const double *mass, double dt_half, hipStream_t st);
hipError_t launch_drift(int64_t n, double *x, double *y, double *z, const double *vx, const double *vy,
                        const double *vz, double *rx, double *ry, double *rz, double dt, hipStream_t st);
hipError_t launch_resync(int64_t n, const double *x, const double *y, const double *z, const double *rx,
                         const double *ry, const double *rz, double *vx, double *vy, double *vz, double dt,
                         hipStream_t st);
// One iteration `it` (maxslot: the per-iteration maxima, from iteration 0);
// it first makes the convergence test of iteration it - 1.
hipError_t launch_shake_iter(const DynArgs &D, Vec3 in, Vec3 out, Vec3C ref, unsigned long long *maxslot, int it,
                             double tol, double *lambda, int *flag, hipStream_t st);
hipError_t launch_rattle_iter(const DynArgs &D, Vec3C pos, Vec3 in, Vec3 out, unsigned long long *maxslot, int it,
                              double tol, double *lambda, int *flag, hipStream_t st);
hipError_t launch_check(const unsigned long long *maxslot, double tol, int it, int *flag, hipStream_t st);
hipError_t launch_kinetic_terms(int64_t n, const double *vx, const double *vy, const double *vz,
                                const double *mass, double *out, hipStream_t st);

}  // namespace dyn
}  // namespace gm
