// Device-resident dynamics for gfx950: velocity half-kick, drift, and the
// SHAKE / RATTLE projections, in f64 with the reference's operation order
// (this file is compiled with -ffp-contract=off), so every step is bit-for-bit
// the CPU path's given the same forces.
//
// integrator.rs:14-71 (half_kick, drift_and_constrain) and
// constraint.rs:62-259 (solve, constrain_velocities); paths relative to the
// reference repository root.  The reference's solvers are Jacobi iterations:
// each iteration computes every constraint's multiplier from the current
// positions (or velocities), accumulates the corrections per atom in
// constraint order, then applies them.  Here one thread per atom walks its
// incident constraints in constraint order, recomputing each multiplier
// exactly as the reference does, so the per-atom sums match bit for bit.
// Positions (velocities) are double-buffered across iterations.
#include "gm_dynamics.hpp"

using namespace gm;
using namespace gm::dyn;

namespace {

constexpr double kAccel = 20.455 * 20.455;  // integrator.rs:9 FORCE_TO_ACCEL_ANG_PER_PS2

// integrator.rs:14-21
__global__ void k_half_kick(int64_t n, double *vx, double *vy, double *vz, const double *fx,
                            const double *fy, const double *fz, const double *mass, double dt_half) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double inv_m = 1.0 / mass[i];
    vx[i] += dt_half * fx[i] * inv_m * kAccel;
    vy[i] += dt_half * fy[i] * inv_m * kAccel;
    vz[i] += dt_half * fz[i] * inv_m * kAccel;
}

// integrator.rs:40-48: remember the positions, then r <- r + dt v.
__global__ void k_drift(int64_t n, double *x, double *y, double *z, const double *vx, const double *vy,
                        const double *vz, double *rx, double *ry, double *rz, double dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rx[i] = x[i];
    ry[i] = y[i];
    rz[i] = z[i];
    x[i] += dt * vx[i];
    y[i] += dt * vy[i];
    z[i] += dt * vz[i];
}

// integrator.rs:64-68: v <- (r_new - r_old) / dt.
__global__ void k_resync(int64_t n, const double *x, const double *y, const double *z, const double *rx,
                         const double *ry, const double *rz, double *vx, double *vy, double *vz, double dt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    vx[i] = (x[i] - rx[i]) / dt;
    vy[i] = (y[i] - ry[i]) / dt;
    vz[i] = (z[i] - rz[i]) / dt;
}

// |v| as the bit pattern that orders like the value (a NaN above everything,
// so it reaches the convergence test as before).
__device__ inline unsigned long long abs_bits(double v) {
    return (unsigned long long)__double_as_longlong(fabs(v));
}

// The largest of the wavefront's abs_bits values into *slot: one atomic per
// wavefront (same-address atomics from every thread would serialise).  Every
// lane calls it.
__device__ inline void wave_atomic_max_bits(unsigned long long *slot, unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    if ((threadIdx.x & 63) == 0 && v != 0ull)
        __hip_atomic_fetch_max(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The convergence test of the previous iteration (k_check's, made by the
// next iteration's kernel instead of a launch of its own): true if this
// iteration is not to run, because an earlier one converged (flag[0]) or the
// previous one did (its largest |multiplier| below tol; then the first
// thread records it as k_check would).  The same for every thread.
__device__ inline bool iteration_skipped(const unsigned long long *prev, double tol, int it, int *flag) {
    if (flag[0]) return true;
    if (prev && __longlong_as_double((long long)*prev) < tol) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            flag[1] = it;
            flag[2] = it & 1;
            flag[0] = 1;
        }
        return true;
    }
    return false;
}

// One SHAKE iteration (constraint.rs:97-150) from positions `in` to `out`.
// Skipped entirely once an earlier iteration has converged.
__global__ void k_shake_iter(DynArgs D, Vec3 in, Vec3 out, Vec3C ref, unsigned long long *maxslot,
                             const unsigned long long *prev, double tol, int it, double *lambda, int *flag) {
    if (iteration_skipped(prev, tol, it, flag)) return;  // the same for every thread
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = a < D.n;
    double ddx = 0.0, ddy = 0.0, ddz = 0.0;
    unsigned long long mx = 0ull;
    const int64_t k0 = act ? D.inc_off[a] : 0, k1 = act ? D.inc_off[a + 1] : 0;
    for (int64_t k = k0; k < k1; k++) {
        const int64_t c = D.inc_con[k];
        const uint32_t i = D.ci[c], j = D.cj[c];
        const double dx = in.x[i] - in.x[j];
        const double dy = in.y[i] - in.y[j];
        const double dz = in.z[i] - in.z[j];
        const double diff = dx * dx + dy * dy + dz * dz - D.dsq[c];
        const double rx = ref.x[i] - ref.x[j];
        const double ry = ref.y[i] - ref.y[j];
        const double rz = ref.z[i] - ref.z[j];
        const double inv_mi = 1.0 / D.mass[i];
        const double inv_mj = 1.0 / D.mass[j];
        const double dot_new_old = dx * rx + dy * ry + dz * rz;
        const double denom = 2.0 * (inv_mi + inv_mj) * dot_new_old;
        const double lam = denom == 0.0 ? 0.0 : diff / denom;
        if ((uint32_t)a == i) {
            ddx -= lam * inv_mi * rx;
            ddy -= lam * inv_mi * ry;
            ddz -= lam * inv_mi * rz;
            lambda[c] = lam;
            mx = max(mx, abs_bits(lam));
        } else {
            ddx += lam * inv_mj * rx;
            ddy += lam * inv_mj * ry;
            ddz += lam * inv_mj * rz;
        }
    }
    wave_atomic_max_bits(maxslot, mx);
    if (!act) return;
    out.x[a] = in.x[a] + ddx;
    out.y[a] = in.y[a] + ddy;
    out.z[a] = in.z[a] + ddz;
}

// One RATTLE velocity iteration (constraint.rs:196-250) from velocities `in` to `out`.
__global__ void k_rattle_iter(DynArgs D, Vec3C pos, Vec3 in, Vec3 out, unsigned long long *maxslot,
                              const unsigned long long *prev, double tol, int it, double *lambda, int *flag) {
    if (iteration_skipped(prev, tol, it, flag)) return;  // the same for every thread
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = a < D.n;
    double ddx = 0.0, ddy = 0.0, ddz = 0.0;
    unsigned long long mx = 0ull;
    const int64_t k0 = act ? D.inc_off[a] : 0, k1 = act ? D.inc_off[a + 1] : 0;
    for (int64_t k = k0; k < k1; k++) {
        const int64_t c = D.inc_con[k];
        const uint32_t i = D.ci[c], j = D.cj[c];
        const double rx = pos.x[i] - pos.x[j];
        const double ry = pos.y[i] - pos.y[j];
        const double rz = pos.z[i] - pos.z[j];
        const double rsq = rx * rx + ry * ry + rz * rz;
        const double dvx = in.x[i] - in.x[j];
        const double dvy = in.y[i] - in.y[j];
        const double dvz = in.z[i] - in.z[j];
        const double rel_vel = dvx * rx + dvy * ry + dvz * rz;
        const double inv_mi = 1.0 / D.mass[i];
        const double inv_mj = 1.0 / D.mass[j];
        const double denom = (inv_mi + inv_mj) * rsq;
        const double mu = denom == 0.0 ? 0.0 : rel_vel / denom;
        if ((uint32_t)a == i) {
            ddx -= mu * inv_mi * rx;
            ddy -= mu * inv_mi * ry;
            ddz -= mu * inv_mi * rz;
            lambda[c] = mu;
            mx = max(mx, abs_bits(mu));
        } else {
            ddx += mu * inv_mj * rx;
            ddy += mu * inv_mj * ry;
            ddz += mu * inv_mj * rz;
        }
    }
    wave_atomic_max_bits(maxslot, mx);
    if (!act) return;
    out.x[a] = in.x[a] + ddx;
    out.y[a] = in.y[a] + ddy;
    out.z[a] = in.z[a] + ddz;
}

// After iteration `it`: converged when max |multiplier| < tol (constraint.rs:152).
// flag[0] = done, flag[1] = iterations run, flag[2] = buffer holding the result.
__global__ void k_check(const unsigned long long *maxslot, double tol, int it, int *flag) {
    if (threadIdx.x != 0 || flag[0]) return;
    const double m = __longlong_as_double((long long)maxslot[it]);
    flag[1] = it + 1;
    flag[2] = (it + 1) & 1;
    if (m < tol) flag[0] = 1;
}

// 0.5 * sum m v^2 / accel factor, per atom (summed on the host in atom order).
__global__ void k_kinetic_terms(int64_t n, const double *vx, const double *vy, const double *vz,
                                const double *mass, double *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = mass[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
}

inline unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }
// the constraint iterations: one wavefront per workgroup, so a system of a
// few tens of thousands of atoms spreads over every CU (each thread is a
// short chain of dependent loads)
inline unsigned grid_of64(int64_t n) { return (unsigned)((n + 63) / 64); }

}  // namespace

namespace gm {
namespace dyn {

hipError_t launch_half_kick(int64_t n, double *vx, double *vy, double *vz, const double *fx, const double *fy,
                            const double *fz, const double *mass, double dt_half, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_half_kick<<<grid_of(n), 256, 0, st>>>(n, vx, vy, vz, fx, fy, fz, mass, dt_half);
    return hipGetLastError();
}

hipError_t launch_drift(int64_t n, double *x, double *y, double *z, const double *vx, const double *vy,
                        const double *vz, double *rx, double *ry, double *rz, double dt, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_drift<<<grid_of(n), 256, 0, st>>>(n, x, y, z, vx, vy, vz, rx, ry, rz, dt);
    return hipGetLastError();
}

hipError_t launch_resync(int64_t n, const double *x, const double *y, const double *z, const double *rx,
                         const double *ry, const double *rz, double *vx, double *vy, double *vz, double dt,
                         hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_resync<<<grid_of(n), 256, 0, st>>>(n, x, y, z, rx, ry, rz, vx, vy, vz, dt);
    return hipGetLastError();
}

hipError_t launch_shake_iter(const DynArgs &D, Vec3 in, Vec3 out, Vec3C ref, unsigned long long *maxslot, int it,
                             double tol, double *lambda, int *flag, hipStream_t st) {
    if (D.n == 0) return hipSuccess;
    k_shake_iter<<<grid_of64(D.n), 64, 0, st>>>(D, in, out, ref, maxslot + it, it > 0 ? maxslot + it - 1 : nullptr,
                                                tol, it, lambda, flag);
    return hipGetLastError();
}

hipError_t launch_rattle_iter(const DynArgs &D, Vec3C pos, Vec3 in, Vec3 out, unsigned long long *maxslot, int it,
                              double tol, double *lambda, int *flag, hipStream_t st) {
    if (D.n == 0) return hipSuccess;
    k_rattle_iter<<<grid_of64(D.n), 64, 0, st>>>(D, pos, in, out, maxslot + it, it > 0 ? maxslot + it - 1 : nullptr,
                                                 tol, it, lambda, flag);
    return hipGetLastError();
}

hipError_t launch_check(const unsigned long long *maxslot, double tol, int it, int *flag, hipStream_t st) {
    k_check<<<1, 64, 0, st>>>(maxslot, tol, it, flag);
    return hipGetLastError();
}

hipError_t launch_kinetic_terms(int64_t n, const double *vx, const double *vy, const double *vz,
                                const double *mass, double *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_kinetic_terms<<<grid_of(n), 256, 0, st>>>(n, vx, vy, vz, mass, out);
    return hipGetLastError();
}

}  // namespace dyn
}  // namespace gm
