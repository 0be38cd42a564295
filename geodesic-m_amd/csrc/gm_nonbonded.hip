// Switched Lennard-Jones over the pair-tile list, gfx950.
//
// Replaces force::nonbonded::compute_pair_forces (force/nonbonded.rs:20-82;
// paths relative to the reference repository root) as driven by
// CpuBackend::compute_forces (cpu_backend.rs:122-179) and the wgpu gather
// kernel (geodesic-gpu/src/nonbonded.wgsl:22-70).  Same physics: Lorentz-
// Berthelot combining, quintic switch over [r_switch, r_cutoff], pairs with
// r2 > r_cutoff^2 or r2 == 0 skipped, F_i -= f d/r, F_j += f d/r.
//
// Mapping: one wavefront per work item = (supercluster, run of list entries).
// Lane = (jl = lane>>3, il = lane&7): each lane owns one i atom of each of the
// supercluster's 8 clusters (il) and one j atom of the current j cluster (jl),
// so every lane evaluates one pair per (i-cluster, j-cluster) tile.
//  - i forces stay in registers across the whole run of entries and are
//    reduced over the 8 jl lanes once at the end;
//  - j forces are reduced over the 8 il lanes (DPP, within 8-lane rows) once
//    per entry, i.e. per up to 8 tiles.
// Displacements come from 32-bit fixed-point positions (fractions of the box):
// the wrapping int32 difference is the exact minimum-image displacement, and
// its conversion to f32 keeps 24 significant bits regardless of where the
// atoms sit in the box.  Pair math is FP32; force partials are converted once
// per reduction into int64 fixed point (2^-32 kcal/mol/A) and added with
// integer atomics, so the result does not depend on scheduling order.
#include "gm_internal.hpp"

using namespace gm;

namespace {

__device__ inline float dpp_sum8(float v) {
    // quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror: after the
    // three steps every lane of an 8-lane group holds the group's sum, and the
    // value is independent of which lane reads it.
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
    return v;
}

__device__ inline void fixed_add(unsigned long long *p, float v) {
    const long long q = (long long)__builtin_rintf(v * 4294967296.0f);
    __hip_atomic_fetch_add(p, (unsigned long long)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline float fixdiff(int a, int b, float scale) {
    return (float)(int)((unsigned)a - (unsigned)b) * scale;
}

__device__ inline uint32_t quantize(double x, double inv_q) {
    return (uint32_t)(int64_t)floor(x * inv_q);
}

// Current positions (atom order, f64) -> fixed point in slot order.
__global__ void k_quantize_slots(int nslots, const int *slot_atom, const double *x, const double *y,
                                 const double *z, double iqx, double iqy, double iqz, int4 *sq) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    int a = slot_atom[s];
    if (a >= 0)
        sq[s] = make_int4((int)quantize(x[a], iqx), (int)quantize(y[a], iqy), (int)quantize(z[a], iqz), 0);
}

__global__ __launch_bounds__(64) void k_nonbonded(
    const int4 *__restrict__ sq, const float2 *__restrict__ sp, const int *__restrict__ wi_sc,
    const int *__restrict__ wi_start, const int *__restrict__ wi_end, const int *__restrict__ ent_cj,
    const uint32_t *__restrict__ ent_imask, const unsigned long long *__restrict__ ent_mask,
    unsigned long long *__restrict__ F, int n_pad, double *__restrict__ e_wi, NbParams p) {
    const int w = blockIdx.x;
    const int lane = threadIdx.x;
    const int il = lane & 7, jl = lane >> 3;
    const int sc = wi_sc[w];
    const unsigned long long lanebit = 1ull << lane;

    int xi[8], yi[8], zi[8];
    float shi[8], sei[8];
    float fix[8], fiy[8], fiz[8];
#pragma unroll
    for (int ci = 0; ci < 8; ci++) {
        const int s = sc * 64 + ci * 8 + il;
        const int4 q = sq[s];
        const float2 pp = sp[s];
        xi[ci] = q.x; yi[ci] = q.y; zi[ci] = q.z;
        shi[ci] = pp.x; sei[ci] = pp.y;
        fix[ci] = 0.f; fiy[ci] = 0.f; fiz[ci] = 0.f;
    }
    double e_acc = 0.0;
    const int e0 = wi_start[w], e1 = wi_end[w];
    for (int e = e0; e < e1; e++) {
        const int cj = ent_cj[e];
        const uint32_t im = ent_imask[e];
        const int sj = cj * 8 + jl;
        const int4 qj = sq[sj];
        const float2 pj = sp[sj];
        float fjx = 0.f, fjy = 0.f, fjz = 0.f, ee = 0.f;
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            if (im & (1u << ci)) {
                const unsigned long long m = ent_mask[(size_t)e * 8 + ci];
                const float dx = fixdiff(qj.x, xi[ci], p.sx);
                const float dy = fixdiff(qj.y, yi[ci], p.sy);
                const float dz = fixdiff(qj.z, zi[ci], p.sz);
                const float r2 = dx * dx + dy * dy + dz * dz;
                if ((m & lanebit) && r2 <= p.rc2 && r2 > 0.f) {
                    const float inv_r2 = __builtin_amdgcn_rcpf(r2);
                    const float sig = shi[ci] + pj.x;
                    const float s2 = sig * sig * inv_r2;
                    const float s6 = s2 * s2 * s2;
                    const float s12 = s6 * s6;
                    const float eps = sei[ci] * pj.y;
                    float v = 4.0f * eps * (s12 - s6);
                    float fr = 24.0f * eps * (2.0f * s12 - s6) * inv_r2;  // F_radial / r
                    if (r2 > p.rsw2) {
                        const float inv_r = __builtin_amdgcn_rsqf(r2);
                        const float r = r2 * inv_r;
                        const float u = (r - p.rsw) * p.inv_w;
                        const float u2 = u * u;
                        const float sw = 1.0f + u2 * u * (-10.0f + u * (15.0f - 6.0f * u));
                        const float om = 1.0f - u;
                        const float dsdr = p.ds_coef * u2 * om * om;
                        fr = fr * sw - v * dsdr * inv_r;
                        v = v * sw;
                    }
                    const float tx = fr * dx, ty = fr * dy, tz = fr * dz;
                    fix[ci] -= tx; fiy[ci] -= ty; fiz[ci] -= tz;
                    fjx += tx; fjy += ty; fjz += tz;
                    ee += v;
                }
            }
        }
        fjx = dpp_sum8(fjx);
        fjy = dpp_sum8(fjy);
        fjz = dpp_sum8(fjz);
        if (il == 0 && im != 0u) {
            fixed_add(&F[sj], fjx);
            fixed_add(&F[(size_t)n_pad + sj], fjy);
            fixed_add(&F[2 * (size_t)n_pad + sj], fjz);
        }
        e_acc += (double)ee;
    }
#pragma unroll
    for (int ci = 0; ci < 8; ci++) {
        float vx = fix[ci], vy = fiy[ci], vz = fiz[ci];
        for (int o = 8; o < 64; o <<= 1) {
            vx += __shfl_xor(vx, o);
            vy += __shfl_xor(vy, o);
            vz += __shfl_xor(vz, o);
        }
        if (jl == 0) {
            const int s = sc * 64 + ci * 8 + il;
            fixed_add(&F[s], vx);
            fixed_add(&F[(size_t)n_pad + s], vy);
            fixed_add(&F[2 * (size_t)n_pad + s], vz);
        }
    }
    for (int o = 1; o < 64; o <<= 1) e_acc += __shfl_xor(e_acc, o);
    if (lane == 0) e_wi[w] = e_acc;
}

inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

namespace gm {

hipError_t launch_nonbonded(gm_sys *s, const double *x, const double *y, const double *z) {
    const size_t nslots = (size_t)s->nsc * 64;
    if (nslots > 0)
        k_quantize_slots<<<grid_for(nslots, 256), 256, 0, s->stream>>>(
            (int)nslots, s->slot_atom.p, x, y, z, s->bp.inv_q[0], s->bp.inv_q[1], s->bp.inv_q[2], s->sq.p);
    hipError_t e = hipEventRecord(s->cur_nb0, s->stream);
    if (e != hipSuccess) return e;
    if (s->n_wi > 0)
        k_nonbonded<<<(unsigned)s->n_wi, 64, 0, s->stream>>>(
            s->sq.p, s->sp.p, s->wi_sc.p, s->wi_start.p, s->wi_end.p, s->ent_cj.p, s->ent_imask.p,
            s->ent_mask.p, s->F.p, (int)nslots, s->e_wi.p, s->nbp);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return hipEventRecord(s->cur_nb1, s->stream);
}

}  // namespace gm
