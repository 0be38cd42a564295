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
                            const double *fy, const double *fz, const double *mass, double dt_half,
                            const int *guard) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (guard && *guard)) return;
    const double inv_m = 1.0 / mass[i];
    vx[i] += dt_half * fx[i] * inv_m * kAccel;
    vy[i] += dt_half * fy[i] * inv_m * kAccel;
    vz[i] += dt_half * fz[i] * inv_m * kAccel;
}

// integrator.rs:40-48: remember the positions, then r <- r + dt v.
__global__ void k_drift(int64_t n, double *x, double *y, double *z, const double *vx, const double *vy,
                        const double *vz, double *rx, double *ry, double *rz, double dt, const int *guard) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (guard && *guard)) return;
    rx[i] = x[i];
    ry[i] = y[i];
    rz[i] = z[i];
    x[i] += dt * vx[i];
    y[i] += dt * vy[i];
    z[i] += dt * vz[i];
}

// integrator.rs:64-68: v <- (r_new - r_old) / dt.
__global__ void k_resync(int64_t n, const double *x, const double *y, const double *z, const double *rx,
                         const double *ry, const double *rz, double *vx, double *vy, double *vz, double dt,
                         const int *guard) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (guard && *guard)) return;
    vx[i] = (x[i] - rx[i]) / dt;
    vy[i] = (y[i] - ry[i]) / dt;
    vz[i] = (z[i] - rz[i]) / dt;
}

// |v| as the bit pattern that orders like the value, NaN as 0: the reference's `lambda_abs > max_lambda`
// never takes a NaN (constraint.rs:127-131), so a NaN multiplier does not
// hold the solve from converging.
__device__ inline unsigned long long abs_bits(double v) {
    return __builtin_isnan(v) ? 0ull : (unsigned long long)__double_as_longlong(fabs(v));
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

__device__ inline unsigned long long wave_max_bits(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o);
        v = w > v ? w : v;
    }
    return v;
}

// One constraint group per 256-thread workgroup (GroupArgs): thread t holds
// group constraint t in registers and group atom t's position (velocity) in a
// double-buffered LDS array.  An iteration is the global kernels' arithmetic,
// operation for operation: thread t computes constraint t's multiplier from
// the current state, then atom t adds its incident constraints' corrections
// in constraint order to its own state.  The group's largest |multiplier|
// decides, in phase 0, when it stops (its first iteration below tol, or cap
// iterations); phase 1 continues each group to the largest count of phase 0
// and reports groups still not below tol there.  Workgroups past the groups
// (phase 1 only) give the atoms in no constraint the reference's v + 0.0.
template <bool VEL>
__global__ __launch_bounds__(256) void k_group_solve(GroupArgs A, DynArgs D, Vec3 buf, Vec3C other, int cap,
                                                     double tol, int phase, double *lambda, const int *guard) {
    __shared__ double sx[2][kGroupAtoms], sy[2][kGroupAtoms], sz[2][kGroupAtoms];
    __shared__ double s_lam[kGroupCons], s_rx[kGroupCons], s_ry[kGroupCons], s_rz[kGroupCons];
    __shared__ uint16_t s_inc[2 * kGroupCons];
    __shared__ unsigned long long s_wmax[4];
    const int g = blockIdx.x, t = threadIdx.x;
    if (guard && *guard) return;  // the same for every thread
    if (g >= A.n_groups) {
        const int64_t k = (int64_t)(g - A.n_groups) * 256 + t;
        if (k < A.n_free) {
            const uint32_t a = A.free_atoms[k];
            buf.x[a] = buf.x[a] + 0.0;
            buf.y[a] = buf.y[a] + 0.0;
            buf.z[a] = buf.z[a] + 0.0;
        }
        return;
    }
    int it = 0, target = cap;
    if (phase == 1) {
        if (g == 0 && t == 0) {
            A.ctl_next[0] = 0;
            A.ctl_next[1] = 0;
        }
        target = A.ctl[0];
        it = A.iters[g];
        if (it == target) {  // this group set the count (or reached it): converged there or not
            if (t == 0 && !A.conv[g]) A.ctl[1] = 1;
            return;
        }
    }
    const int a0 = A.atom_off[g], na = A.atom_off[g + 1] - a0;
    const int c0 = A.con_off[g], nc = A.con_off[g + 1] - c0;
    const int ib = A.inc_off[a0], ni = A.inc_off[a0 + na] - ib;
    uint32_t ga = 0;
    double inv_ma = 0.0;
    int q0 = 0, q1 = 0;
    if (t < na) {
        ga = A.atoms[a0 + t];
        sx[0][t] = buf.x[ga];
        sy[0][t] = buf.y[ga];
        sz[0][t] = buf.z[ga];
        inv_ma = 1.0 / D.mass[ga];
        q0 = A.inc_off[a0 + t] - ib;
        q1 = A.inc_off[a0 + t + 1] - ib;
    }
    for (int q = t; q < ni; q += 256) s_inc[q] = A.inc[ib + q];
    uint32_t gc = 0;
    int li = 0, lj = 0;
    double cd = 0.0, rx = 0.0, ry = 0.0, rz = 0.0, inv_mi = 0.0, inv_mj = 0.0;
    if (t < nc) {
        gc = A.cons[c0 + t];
        const uint32_t lo = A.con_loc[c0 + t];
        li = (int)(lo & 0xffffu);
        lj = (int)(lo >> 16);
        const uint32_t i = D.ci[gc], j = D.cj[gc];
        inv_mi = 1.0 / D.mass[i];
        inv_mj = 1.0 / D.mass[j];
        rx = other.x[i] - other.x[j];
        ry = other.y[i] - other.y[j];
        rz = other.z[i] - other.z[j];
        if (VEL) {
            const double rsq = rx * rx + ry * ry + rz * rz;
            cd = (inv_mi + inv_mj) * rsq;  // constraint.rs:215-216
        } else {
            cd = D.dsq[gc];
        }
        s_rx[t] = rx;
        s_ry[t] = ry;
        s_rz[t] = rz;
    }
    __syncthreads();
    int cur = 0;
    double lam = 0.0;
    unsigned long long gmax = 0ull;
    while (it < target) {
        unsigned long long mx = 0ull;
        if (t < nc) {
            const double dx = sx[cur][li] - sx[cur][lj];
            const double dy = sy[cur][li] - sy[cur][lj];
            const double dz = sz[cur][li] - sz[cur][lj];
            if (VEL) {  // constraint.rs:208-217
                const double rel_vel = dx * rx + dy * ry + dz * rz;
                lam = cd == 0.0 ? 0.0 : rel_vel / cd;
            } else {  // constraint.rs:100-113
                const double diff = dx * dx + dy * dy + dz * dz - cd;
                const double dot_new_old = dx * rx + dy * ry + dz * rz;
                const double denom = 2.0 * (inv_mi + inv_mj) * dot_new_old;
                lam = denom == 0.0 ? 0.0 : diff / denom;
            }
            s_lam[t] = lam;
            mx = abs_bits(lam);
        }
        mx = wave_max_bits(mx);
        if ((t & 63) == 0) s_wmax[t >> 6] = mx;
        __syncthreads();
        if (t < na) {
            double ddx = 0.0, ddy = 0.0, ddz = 0.0;
            for (int q = q0; q < q1; q++) {
                const int e = s_inc[q], c = e >> 1;
                const double l = s_lam[c];
                if (e & 1) {
                    ddx += l * inv_ma * s_rx[c];
                    ddy += l * inv_ma * s_ry[c];
                    ddz += l * inv_ma * s_rz[c];
                } else {
                    ddx -= l * inv_ma * s_rx[c];
                    ddy -= l * inv_ma * s_ry[c];
                    ddz -= l * inv_ma * s_rz[c];
                }
            }
            sx[cur ^ 1][t] = sx[cur][t] + ddx;
            sy[cur ^ 1][t] = sy[cur][t] + ddy;
            sz[cur ^ 1][t] = sz[cur][t] + ddz;
        }
        gmax = s_wmax[0];
        for (int w = 1; w < 4; w++) gmax = s_wmax[w] > gmax ? s_wmax[w] : gmax;
        __syncthreads();
        cur ^= 1;
        it++;
        if (phase == 0 && __longlong_as_double((long long)gmax) < tol) break;  // the same in every thread
    }
    const bool conv = __longlong_as_double((long long)gmax) < tol;
    if (t < na) {
        buf.x[ga] = sx[cur][t];
        buf.y[ga] = sy[cur][t];
        buf.z[ga] = sz[cur][t];
    }
    if (t < nc) lambda[gc] = lam;
    if (t == 0) {
        if (phase == 0) {
            A.iters[g] = it;
            A.conv[g] = conv ? 1 : 0;
            __hip_atomic_fetch_max(&A.ctl[0], it, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if (!conv) {
            A.ctl[1] = 1;
        }
    }
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
                            const double *fz, const double *mass, double dt_half, hipStream_t st,
                            const int *guard) {
    if (n == 0) return hipSuccess;
    k_half_kick<<<grid_of(n), 256, 0, st>>>(n, vx, vy, vz, fx, fy, fz, mass, dt_half, guard);
    return hipGetLastError();
}

hipError_t launch_drift(int64_t n, double *x, double *y, double *z, const double *vx, const double *vy,
                        const double *vz, double *rx, double *ry, double *rz, double dt, hipStream_t st,
                        const int *guard) {
    if (n == 0) return hipSuccess;
    k_drift<<<grid_of(n), 256, 0, st>>>(n, x, y, z, vx, vy, vz, rx, ry, rz, dt, guard);
    return hipGetLastError();
}

hipError_t launch_resync(int64_t n, const double *x, const double *y, const double *z, const double *rx,
                         const double *ry, const double *rz, double *vx, double *vy, double *vz, double dt,
                         hipStream_t st, const int *guard) {
    if (n == 0) return hipSuccess;
    k_resync<<<grid_of(n), 256, 0, st>>>(n, x, y, z, rx, ry, rz, vx, vy, vz, dt, guard);
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

hipError_t launch_group_solve(const GroupArgs &G, const DynArgs &D, bool velocity, Vec3 buf, Vec3C other, int cap,
                              double tol, int phase, double *lambda, hipStream_t st, const int *guard) {
    const int64_t free_blocks = phase == 1 ? (G.n_free + 255) / 256 : 0;
    const unsigned grid = (unsigned)(G.n_groups + free_blocks);
    if (grid == 0) return hipSuccess;
    if (velocity)
        k_group_solve<true><<<grid, 256, 0, st>>>(G, D, buf, other, cap, tol, phase, lambda, guard);
    else
        k_group_solve<false><<<grid, 256, 0, st>>>(G, D, buf, other, cap, tol, phase, lambda, guard);
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
