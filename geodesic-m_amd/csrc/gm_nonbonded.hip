// Force evaluation kernel for gfx950: switched Lennard-Jones over the
// pair-tile list, with the bonded terms as extra workgroups of the same launch.
//
// Replaces force::nonbonded::compute_pair_forces (force/nonbonded.rs:20-82;
// paths relative to the reference repository root) as driven by
// CpuBackend::compute_forces (cpu_backend.rs:92-182) and the wgpu gather
// kernel (geodesic-gpu/src/nonbonded.wgsl:22-70).  Same physics: Lorentz-
// Berthelot combining, quintic switch over [r_switch, r_cutoff], pairs with
// r2 > r_cutoff^2 or r2 == 0 skipped, F_i -= f d/r, F_j += f d/r.
//
// Tile path (workgroups after the bonded ones): one wavefront per work item =
// (supercluster, run of list entries).  Lane = (jl = lane>>3, il = lane&7):
// each lane owns one i atom of each of the supercluster's 8 clusters (il) and
// one j atom of the current entry's cluster (jl), so every lane evaluates one
// pair per (i-cluster, j-cluster) tile.
//  - i forces stay in registers across the run and are reduced over the 8 jl
//    lanes once at the end; j forces are reduced over the 8 il lanes with
//    three DPP adds once per entry (up to 8 tiles).
//  - Each entry carries its j-cluster's atom indices and one mask byte per
//    lane (bit ci = "this lane's pair in tile ci is in the list"), so the next
//    entries' loads depend only on the entry index; they are issued one and
//    two entries ahead of the arithmetic.
//  - Positions are quantised on the fly to 32-bit fixed-point fractions of the
//    box; the wrapping int32 difference is the exact minimum-image
//    displacement and its f32 conversion keeps 24 significant bits anywhere in
//    the box.  Pair math is FP32; partial sums are converted once per reduction
//    to int64 fixed point (2^-32 kcal/mol/A) and added with integer atomics, so
//    the result does not depend on scheduling order.
// Bonded path (the first ceil(n_terms/64) workgroups): one thread per bonded term (f64,
// gm_terms.hpp), running beside the tiles instead of in launches of its own.
#include "gm_internal.hpp"
#include "gm_terms.hpp"

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

__device__ inline void fixed_add(unsigned long long *p, float v, float scale) {
    const long long q = (long long)__builtin_rintf(v * scale);
    __hip_atomic_fetch_add(p, (unsigned long long)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline float fixdiff(int a, int b) { return (float)(int)((unsigned)a - (unsigned)b); }

struct JAtom {
    double x, y, z;
    float2 par;
};

__device__ inline JAtom load_j(const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, const float2 *__restrict__ atom_par,
                               int ja) {
    JAtom r;
    const int a = ja < 0 ? 0 : ja;  // padding slot: any valid atom, its mask bits are 0
    r.x = x[a];
    r.y = y[a];
    r.z = z[a];
    r.par = atom_par[a];
    return r;
}

// One work item.  CUBIC: all three box edges equal, so lengths are carried in
// units of the fixed-point step (box / 2^32) and no per-pair rescaling is
// needed; the unit change is folded into the constants and into the scale at
// which forces enter the fixed-point accumulator.
// SHELL: each pair's switched LJ is weighted by S(r) - S_in(r) instead of
// S(r) (the shell part of a split of the switched LJ into inner + shell);
// below the inner switch both are 1, so short-range pairs contribute exactly 0.
template <bool CUBIC, bool SHELL>
__device__ __attribute__((always_inline)) inline void tile_work(const int w, const int lane, const TileArgs &A) {
    const double *__restrict__ x = A.x;
    const double *__restrict__ y = A.y;
    const double *__restrict__ z = A.z;
    const float2 *__restrict__ atom_par = A.atom_par;
    const int *__restrict__ slot_atom = A.slot_atom;
    const int2 *__restrict__ ent_meta = A.ent_meta;
    const int *__restrict__ ent_ja = A.ent_ja;
    const uint8_t *__restrict__ ent_lb = A.ent_lb;
    unsigned long long *__restrict__ F = A.F;
    const int n_pad = A.n_pad;
    const NbParams &p = A.p;
    const double iqx = A.iq[0], iqy = A.iq[1], iqz = A.iq[2];
    // The run's list data (mask bytes, j atoms, metadata) arrive in LDS in one
    // burst of independent loads; the loop then only gathers positions, one
    // entry ahead of the arithmetic.  Reduced j forces wait in LDS and leave
    // with atomics after the loop: gfx950 counts atomics in vmcnt, so issuing
    // them inside the loop would make every iteration wait for the previous
    // one's.
    __shared__ uint32_t s_lb[kMaxRun * 16];
    __shared__ int s_ja[kMaxRun * 8];
    __shared__ int2 s_meta[kMaxRun];
    __shared__ float jf[3][kMaxRun * 8];
    const int il = lane & 7, jl = lane >> 3;
    const int sc = A.wi_sc[w];
    const int e0 = A.wi_start[w], ne = A.wi_end[w] - e0;
    {
        const uint32_t *lb = reinterpret_cast<const uint32_t *>(ent_lb + (size_t)e0 * 64);
        uint32_t lbv[kMaxRun / 4];
        int jav[kMaxRun / 8];
#pragma unroll
        for (int r = 0; r < kMaxRun / 4; r++) {
            const int k = lane + 64 * r;
            lbv[r] = k < ne * 16 ? lb[k] : 0u;
        }
#pragma unroll
        for (int r = 0; r < kMaxRun / 8; r++) {
            const int k = lane + 64 * r;
            jav[r] = k < ne * 8 ? ent_ja[(size_t)e0 * 8 + k] : -1;
        }
        const int2 mv = lane < ne ? ent_meta[e0 + lane] : make_int2(-1, 0);
#pragma unroll
        for (int r = 0; r < kMaxRun / 4; r++) s_lb[lane + 64 * r] = lbv[r];
#pragma unroll
        for (int r = 0; r < kMaxRun / 8; r++) s_ja[lane + 64 * r] = jav[r];
        if (lane < kMaxRun) s_meta[lane] = mv;
    }
    // The supercluster's 64 i atoms live in LDS (fixed-point position + sigma/2
    // in one int4, sqrt(eps) beside it): every tile reads 8 distinct entries,
    // broadcast to 8 lanes each, which keeps 40 registers free for occupancy.
    __shared__ int4 s_iq[64];
    __shared__ float s_ie[64];
    const float inv_s = CUBIC ? p.inv_s : 1.0f;
    {
        const int a = slot_atom[sc * 64 + lane];
        const int aa = a < 0 ? 0 : a;
        const float2 pp = atom_par[aa];
        s_iq[lane] = make_int4(quantize_fixed(x[aa], iqx), quantize_fixed(y[aa], iqy),
                               quantize_fixed(z[aa], iqz), __float_as_int(pp.x * inv_s));
        s_ie[lane] = pp.y;
    }
    float fix[8], fiy[8], fiz[8];
#pragma unroll
    for (int ci = 0; ci < 8; ci++) { fix[ci] = 0.f; fiy[ci] = 0.f; fiz[ci] = 0.f; }
    __syncthreads();
    const uint8_t *s_lb8 = reinterpret_cast<const uint8_t *>(s_lb);
    const float rc2 = CUBIC ? p.rc2_u : p.rc2;
    const unsigned rc2_bits = (unsigned)__float_as_int(rc2);
    const float rsw = CUBIC ? p.rsw_u : p.rsw, inv_w = CUBIC ? p.inv_w_u : p.inv_w;
    const float ds_coef = CUBIC ? p.ds_coef_u : p.ds_coef;
    const float rin = CUBIC ? p.rin_sw_u : p.rin_sw, inv_w_in = CUBIC ? p.inv_w_in_u : p.inv_w_in;
    const float ds_coef_in = CUBIC ? p.ds_coef_in_u : p.ds_coef_in;
    double e_acc = 0.0;
    JAtom cur = load_j(x, y, z, atom_par, s_ja[jl]);
    for (int k = 0; k < ne; k++) {
        const JAtom nxt = load_j(x, y, z, atom_par, s_ja[min(k + 1, ne - 1) * 8 + jl]);
        const uint32_t im = (uint32_t)__builtin_amdgcn_readfirstlane(s_meta[k].y);
        const uint32_t bits = s_lb8[k * 64 + lane];

        const int qjx = quantize_fixed(cur.x, iqx);
        const int qjy = quantize_fixed(cur.y, iqy);
        const int qjz = quantize_fixed(cur.z, iqz);
        const float shj = cur.par.x * inv_s, sej = cur.par.y;
        float fjx = 0.f, fjy = 0.f, fjz = 0.f, ee = 0.f;
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            if (im & (1u << ci)) {
                // Branch-free tile: every lane evaluates its pair; pairs not in
                // the list, beyond r_cutoff or coincident get zero force and
                // energy through one select (which also discards the inf/NaN
                // an r2 of 0 produces).  Below r_switch, u clamps to 0, where
                // S = 1 and dS/dr = 0 exactly.
                const int4 qi = s_iq[ci * 8 + il];
                float dx = fixdiff(qjx, qi.x);
                float dy = fixdiff(qjy, qi.y);
                float dz = fixdiff(qjz, qi.z);
                if (!CUBIC) { dx *= p.sx; dy *= p.sy; dz *= p.sz; }
                const float r2 = dx * dx + dy * dy + dz * dz;
                // 0 < r2 <= rc2 as one unsigned compare of the bit patterns
                const bool ok = ((bits >> ci) & 1u) &&
                                ((unsigned)__float_as_int(r2) - 1u) < rc2_bits;
                const float inv_r2 = __builtin_amdgcn_rcpf(r2);
                const float inv_r = __builtin_amdgcn_rsqf(r2);
                const float sig = __int_as_float(qi.w) + shj;
                const float s2 = sig * sig * inv_r2;
                const float s6 = s2 * s2 * s2;
                const float s12 = s6 * s6;
                const float eps = s_ie[ci * 8 + il] * sej;
                const float v_lj = 4.0f * eps * (s12 - s6);
                const float f_lj = 24.0f * eps * (2.0f * s12 - s6) * inv_r2;  // F_radial / r
                const float r = r2 * inv_r;
                const float u = fmaxf((r - rsw) * inv_w, 0.0f);
                const float u2 = u * u;
                float sw = 1.0f + u2 * u * (-10.0f + u * (15.0f - 6.0f * u));
                const float om = 1.0f - u;
                float dsdr = ds_coef * u2 * om * om;
                if (SHELL) {
                    const float ui = fminf(fmaxf((r - rin) * inv_w_in, 0.0f), 1.0f);
                    const float ui2 = ui * ui, omi = 1.0f - ui;
                    sw -= 1.0f + ui2 * ui * (-10.0f + ui * (15.0f - 6.0f * ui));
                    dsdr -= ds_coef_in * ui2 * omi * omi;
                }
                const float fr = ok ? f_lj * sw - v_lj * dsdr * inv_r : 0.0f;
                const float v = ok ? v_lj * sw : 0.0f;
                const float tx = fr * dx, ty = fr * dy, tz = fr * dz;
                fix[ci] -= tx; fiy[ci] -= ty; fiz[ci] -= tz;
                fjx += tx; fjy += ty; fjz += tz;
                ee += v;
            }
        }
        fjx = dpp_sum8(fjx);
        fjy = dpp_sum8(fjy);
        fjz = dpp_sum8(fjz);
        if (il == 0) {
            jf[0][k * 8 + jl] = fjx;
            jf[1][k * 8 + jl] = fjy;
            jf[2][k * 8 + jl] = fjz;
        }
        e_acc += (double)ee;
        cur = nxt;
    }
    __syncthreads();
    const float fscale = CUBIC ? p.fscale_u : 4294967296.0f;
    for (int c = 0; c < 3; c++) {
        for (int idx = lane; idx < ne * 8; idx += 64) {
            const int2 m = s_meta[idx >> 3];
            const float v = jf[c][idx];
            if (m.y != 0 && v != 0.0f)
                fixed_add(&F[(size_t)c * n_pad + (size_t)m.x * 8 + (idx & 7)], v, fscale);
        }
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
            const size_t s = (size_t)sc * 64 + ci * 8 + il;
            fixed_add(&F[s], vx, fscale);
            fixed_add(&F[(size_t)n_pad + s], vy, fscale);
            fixed_add(&F[2 * (size_t)n_pad + s], vz, fscale);
        }
    }
    for (int o = 1; o < 64; o <<= 1) e_acc += __shfl_xor(e_acc, o);
    if (lane == 0) A.e_wi[w] = e_acc;
    __syncthreads();  // LDS is reused by the next work item of a persistent wave
}

template <bool CUBIC, bool SHELL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_forces(TileArgs A,
                                                                                       TermArgs ta) {
    // bonded workgroups come first so they start with the tiles instead of
    // forming a tail after them
    const int lane = threadIdx.x;
    const int64_t n_terms = ta.nb + ta.na + ta.nd + ta.n14;
    const int n_bonded_groups = (int)((n_terms + 63) / 64);
    if ((int)blockIdx.x < n_bonded_groups) {
        const int64_t k = (int64_t)blockIdx.x * 64 + lane;
        if (k < n_terms) term_dispatch(ta, k);
        return;
    }
    tile_work<CUBIC, SHELL>((int)blockIdx.x - n_bonded_groups, lane, A);
}

// Persistent batched tile kernel: a fixed grid of wavefronts walks the tile
// work items (replica, work item) in a static order, with no shared counter.
// Workgroups are dealt to the 8 XCDs in turn, so wavefront w runs on XCD
// w % 8; XCD x owns the contiguous item range [x n / 8, (x + 1) n / 8), which
// keeps neighbouring items (same replica, nearby superclusters) in one L2.
// Each item's arithmetic is the single-system kernel's, so results do not
// depend on which wavefront takes which item.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void k_batch_tiles(
    const ReplicaDesc *__restrict__ reps, const int2 *__restrict__ items, int n_items) {
    const int lane = threadIdx.x;
    const int w = blockIdx.x, nw = gridDim.x;
    const int x = w & 7, l = w >> 3;
    const int nwx = (nw - x + 7) >> 3;  // wavefronts on XCD x
    const int lo = (int)(((int64_t)n_items * x) >> 3), hi = (int)(((int64_t)n_items * (x + 1)) >> 3);
    for (int it = lo + l; it < hi; it += nwx) {
        const int2 rw = items[it];
        tile_work<true, false>(rw.y, lane, reps[rw.x].tile);  // replicas are cubic
    }
}

// Bonded terms of every replica: one workgroup per (replica, group of 64 terms).
__global__ __launch_bounds__(64) void k_batch_bonded(const ReplicaDesc *__restrict__ reps,
                                                     const int2 *__restrict__ items) {
    const int2 rg = items[blockIdx.x];
    const TermArgs &t = reps[rg.x].terms;
    const int64_t k = (int64_t)rg.y * 64 + threadIdx.x;
    if (k < t.nb + t.na + t.nd + t.n14) term_dispatch(t, k);
}

}  // namespace

namespace gm {

TermArgs term_args(gm_sys *s, const double *x, const double *y, const double *z) {
    TermArgs t{};
    t.x = x; t.y = y; t.z = z;
    for (int a = 0; a < 3; a++) t.box[a] = s->box[a];
    t.mi = s->bonded_mi;
    t.F = s->F.p;
    t.atom_slot = s->atom_slot.p;
    t.n_pad = (size_t)s->nsc * 64;
    t.bi = s->bi.p; t.bj = s->bj.p; t.bk = s->bk.p; t.br0 = s->br0.p;
    t.ai = s->ai.p; t.aj = s->aj.p; t.ak = s->ak.p; t.akth = s->akth.p; t.ath0 = s->ath0.p;
    t.di = s->di.p; t.dj = s->dj.p; t.dk = s->dk.p; t.dl = s->dl.p; t.dn = s->dn.p;
    t.dkphi = s->dkphi.p; t.ddelta = s->ddelta.p;
    t.p14i = s->p14i.p; t.p14j = s->p14j.p;
    t.sigma = s->sigma.p; t.epsilon = s->eps.p;
    t.scale14 = s->scale14;
    t.nb = s->nb; t.na = s->na; t.nd = s->nd;
    t.n14 = (s->scale14 != 0.0) ? s->n14 : 0;
    t.e_items = s->e_items.p;
    return t;
}

TileArgs tile_args(gm_sys *s, const double *x, const double *y, const double *z) {
    TileArgs A{};
    A.x = x; A.y = y; A.z = z;
    A.atom_par = s->atom_par.p;
    A.slot_atom = s->slot_atom.p;
    A.wi_sc = s->wi_sc.p; A.wi_start = s->wi_start.p; A.wi_end = s->wi_end.p;
    A.ent_meta = s->ent_meta.p; A.ent_ja = s->ent_ja.p; A.ent_lb = s->ent_lb.p;
    A.F = s->F.p;
    A.n_pad = (int)(s->nsc * 64);
    A.e_wi = s->e_wi.p;
    A.p = s->nbp;
    for (int a = 0; a < 3; a++) A.iq[a] = s->bp.inv_q[a];
    A.cubic = (s->box[0] == s->box[1] && s->box[1] == s->box[2]) ? 1 : 0;
    return A;
}

// One launch: ceil(n_terms / 64) bonded workgroups, then the tile work items.
// Shell mode: tile work items only.
hipError_t launch_forces(gm_sys *s, const double *x, const double *y, const double *z, bool shell) {
    TermArgs ta = term_args(s, x, y, z);
    TileArgs A = tile_args(s, x, y, z);
    if (shell) {
        ta.nb = ta.na = ta.nd = ta.n14 = 0;
        A.p = s->nbp_shell;
    }
    const int64_t nterms = ta.nb + ta.na + ta.nd + ta.n14;
    const int64_t grid = s->n_wi + (nterms + 63) / 64;
    if (grid > 0) {
        auto kern = shell ? (A.cubic ? k_forces<true, true> : k_forces<false, true>)
                          : (A.cubic ? k_forces<true, false> : k_forces<false, false>);
        kern<<<(unsigned)grid, 64, 0, s->stream>>>(A, ta);
    }
    return hipGetLastError();
}

hipError_t launch_batch_tiles(const ReplicaDesc *reps, const int2 *items, int n_items, int grid,
                              hipStream_t st) {
    if (n_items > 0 && grid > 0) k_batch_tiles<<<(unsigned)grid, 64, 0, st>>>(reps, items, n_items);
    return hipGetLastError();
}

hipError_t launch_batch_bonded(const ReplicaDesc *reps, const int2 *items, int n_items, hipStream_t st) {
    if (n_items > 0) k_batch_bonded<<<(unsigned)n_items, 64, 0, st>>>(reps, items);
    return hipGetLastError();
}

}  // namespace gm
