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
// Prep kernel (one thread per atom, before the force launch): each atom's
// position quantised to 32-bit fixed-point fractions of the box and written to
// its slot together with sigma/2, so the tile loop reads one contiguous int4
// per j atom; and the largest displacement of any atom since the list build,
// which bounds how far every listed pair can have come closer (tile pruning).
//
// Tile path (workgroups after the bonded ones): one wavefront per work item =
// (supercluster, run of list entries).  Lane = (jl = lane>>3, il = lane&7):
// each lane owns one i atom of each of the supercluster's 8 clusters (il) and
// one j atom of the current entry's cluster (jl), so every lane evaluates one
// pair per (i-cluster, j-cluster) tile.
//  - i forces stay in registers across the run and are reduced over the 8 jl
//    lanes once at the end; j forces are reduced over the 8 il lanes with
//    three DPP adds once per entry (up to 8 tiles).
//  - Each entry carries one mask byte per lane (bit ci = "this lane's pair in
//    tile ci is in the list") and one distance byte per tile; tiles whose
//    closest listed pair at build time is provably beyond r_cutoff now are
//    dropped from the entry's tile mask before the loop.  Such a tile would
//    only add zeros, so results are bit-identical with and without pruning.
//  - The wrapping int32 difference of fixed-point positions is the exact
//    minimum-image displacement and its f32 conversion keeps 24 significant
//    bits anywhere in the box.  Pair math is FP32; partial sums are converted
//    once per reduction to int64 fixed point (2^-32 kcal/mol/A) and added with
//    integer atomics, so the result does not depend on scheduling order.
// Bonded path (the first ceil(n_terms/64) workgroups): one thread per bonded term (f64,
// gm_terms.hpp), running beside the tiles instead of in launches of its own.
#include <hip/hip_ext.h>

#include <algorithm>

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

// v * scale rounded to an integer: the fixed-point image of one partial sum.
// The finish kernel rounds i-force records with the same expression, so the
// integers do not depend on where the rounding happens.
__device__ inline long long to_fixed_f(float v, float scale) { return (long long)__builtin_rintf(v * scale); }

// The fixed-point image of v into the accumulator word p; a value that is not
// finite or reaches 2^62 in fixed point marks the slot instead (*bad = 1).
__device__ inline void fixed_add(unsigned long long *p, float v, float scale, unsigned *bad) {
    if (!(fabsf(v * scale) < kFixedLimitF)) {
        *bad = 1u;
        return;
    }
    __hip_atomic_fetch_add(p, (unsigned long long)to_fixed_f(v, scale), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Exchange-and-add across lane bit 5: with a, b one value each per lane,
// returns a[l] + a[l + 32] on lanes l < 32 and b[l - 32] + b[l] on the others
// (v_permlane32_swap: the upper half of a trades places with the lower half of b).
__device__ inline float swap_add32(float a, float b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// The same across lane bit 4 (rows of 16 lanes): a summed over row pairs on
// even rows, b on odd rows (v_permlane16_swap: odd rows of a trade with even rows of b).
__device__ inline float swap_add16(float a, float b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// A wave-uniform value copied into a vector register (a plain move, no load).
__device__ inline float in_vgpr(float v) {
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "s"(v));
    return r;
}

// An SGPR holding the f32 bit pattern k, opaque to the optimiser (a constant
// operand of a class test would be folded back into a mask-and-compare).
__device__ inline float sgpr_bits(unsigned k) {
    float r;
    asm("s_mov_b32 %0, %1" : "=s"(r) : "i"(k));
    return r;
}

template <int CTRL>
__device__ inline float dpp_max(float v) {
    return fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false)));
}

// Largest value of a row of 16 lanes, in every lane of the row (non-negative
// values; every lane active).
__device__ inline float row_max16(float v) {
    v = dpp_max<0xB1>(v);      // quad_perm [1,0,3,2]
    v = dpp_max<0x4E>(v);      // quad_perm [2,3,0,1]
    v = dpp_max<0x141>(v);     // row_half_mirror
    return dpp_max<0x128>(v);  // row_ror:8
}

// Largest value of the wavefront, in every lane (non-negative values; every
// lane active): rows by DPP, then across rows by the gfx950 permlane swaps.
__device__ inline float wave_max(float v) {
    v = row_max16(v);
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = fmaxf(__uint_as_float(r[0]), __uint_as_float(r[1]));
    const auto q = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return fmaxf(__uint_as_float(q[0]), __uint_as_float(q[1]));
}

// One atom of the prep pass: quantise, store x, y, z and sigma/2 to the slot,
// return the displacement from the build position (A).  Every load is by
// atom index.  A position that is not finite marks the slot (its force is
// reported as NaN), stores sigma/2 as -0.0 (its sign bit tells the force
// kernel to mark every listed partner of the atom the same way; as a value it
// keeps the pairs the kernel does not evaluate at exactly zero) and returns an
// infinite displacement, which turns tile pruning off for the evaluation.
__device__ inline float prep_atom(const PrepArgs &P, int a) {
    const int s = P.atom_slot[a];
    const double x = P.x[a], y = P.y[a], z = P.z[a];
    const float sh = P.atom_par[a].x * P.sig_scale;
    const int qx = quantize_fixed(x, P.iq[0]);
    const int qy = quantize_fixed(y, P.iq[1]);
    const int qz = quantize_fixed(z, P.iq[2]);
    const bool fin = isfinite(x) && isfinite(y) && isfinite(z);
    const int w = P.wtype ? (P.wtype[a] | (fin ? 0 : (int)0x80000000)) : (fin ? __float_as_int(sh) : (int)0x80000000);
    P.qpos[s] = make_int4(qx, qy, qz, w);
    if (!fin) {
        P.bad[s] = 1u;
        return INFINITY;
    }
    const float dx = fixdiff(qx, (int)P.qref[a]) * P.sq_len[0];
    const float dy = fixdiff(qy, (int)P.qref[P.n + a]) * P.sq_len[1];
    const float dz = fixdiff(qz, (int)P.qref[2 * (size_t)P.n + a]) * P.sq_len[2];
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

__device__ inline void atomic_max_nonneg(unsigned *slot, float v) {
    // non-negative floats order like their bit patterns
    __hip_atomic_fetch_max(slot, __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One thread per atom, 256 per workgroup (spread over many CUs: the kernel is
// a load round trip and a store); each wavefront raises one of the
// kDispSlots displacement words with its largest value.
__global__ __launch_bounds__(256) void k_prep(PrepArgs P) {
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a == 0) *P.bad_first = 0x7fffffff;  // read by the finish kernel of this evaluation
    float d = 0.f;
    if (a < P.n) d = prep_atom(P, a);
    d = wave_max(d);
    if ((threadIdx.x & 63) == 0)
        atomic_max_nonneg(&P.disp[((blockIdx.x * 4 + (threadIdx.x >> 6)) % kDispSlots) * kDispStride], d);
}

// Batched prep: atom a of the concatenated replicas.  A wavefront whose atoms
// all belong to one replica (nearly all of them) reads that replica's
// descriptor once, with scalar loads, and raises its bound once; one that
// straddles a boundary reads the descriptor and raises the bound per lane.
__global__ __launch_bounds__(256) void k_batch_prep(const ReplicaDesc *__restrict__ reps,
                                                    const int *__restrict__ atom_rep,
                                                    const int64_t *__restrict__ atom_off, int64_t n_atoms,
                                                    unsigned *ctr) {
    const int64_t a = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int r = a < n_atoms ? atom_rep[a] : -1;
    if (r >= 0 && a == atom_off[r]) *reps[r].prep.bad_first = 0x7fffffff;  // the replica's first atom
    if (a < 8) ctr[a * kCtrStride] = 0u;  // the tile kernel's work counters, one per XCD
    const int r0 = __builtin_amdgcn_readfirstlane(r);
    const size_t slot = (size_t)((blockIdx.x * 4 + (threadIdx.x >> 6)) % kDispSlots) * kDispStride;
    if (__all(r == r0)) {
        if (r0 < 0) return;
        const PrepArgs &P = reps[r0].prep;
        const float d = wave_max(prep_atom(P, (int)(a - atom_off[r0])));
        if ((threadIdx.x & 63) == 0) atomic_max_nonneg(&P.disp[slot], d);
    } else if (r >= 0) {
        const PrepArgs &P = reps[r].prep;
        atomic_max_nonneg(&P.disp[slot], prep_atom(P, (int)(a - atom_off[r])));
    }
}

// One work item.  CUBIC: all three box edges equal, so lengths are carried in
// units of the fixed-point step (box / 2^32) and no per-pair rescaling is
// needed; the unit change is folded into the constants and into the scale at
// which forces enter the fixed-point accumulator.
// SHELL: each pair's switched LJ is weighted by S(r) - S_in(r) instead of
// S(r) (the shell part of a split of the switched LJ into inner + shell);
// below the inner switch both are 1, so short-range pairs contribute exactly 0.
//
// Pair arithmetic, with E4 = 4 eps = (2 sqrt eps_i)(2 sqrt eps_j), s6 = (sig/r)^6,
// u = clamp((r - r_switch) / w, 0, 1), S = 1 + u^3 (-10 + u (15 - 6 u)),
// dS/dr = (-30 / w) (u - u^2)^2:
//   energy  v = E4 (s6^2 - s6) S
//   F / r   = E4 / r^2 [6 (2 s6^2 - s6) S - (s6^2 - s6) r dS/dr]
// (the reference's 24 eps (2 s12 - s6) / r^2 S - v_lj dS/dr / r).
template <bool CUBIC, bool SHELL>
__device__ __attribute__((always_inline)) inline void tile_work(const int w, const int lane, const TileArgs &A) {
    const int4 *__restrict__ qpos = A.qpos;
    const float *__restrict__ qeps = A.qeps;
    const uint8_t *__restrict__ ent_lb = A.ent_lb;
    unsigned long long *__restrict__ F = A.F;
    const int n_pad = A.n_pad;
    const NbParams &p = A.p;
    // The run's list data (mask bytes, metadata) arrive in LDS in one burst of
    // independent loads; the loop then only gathers j positions, one entry
    // ahead of the arithmetic.  Reduced j forces wait in LDS and leave with
    // atomics after the loop: gfx950 counts atomics in vmcnt, so issuing them
    // inside the loop would make every iteration wait for the previous one's.
    __shared__ uint32_t s_lb[kMaxRun * 16];
    __shared__ int2 s_meta[kMaxRun];
    __shared__ float jf[3][kMaxRun * 8];
    const int il = lane & 7, jl = lane >> 3;
    // (supercluster | entry count << 26, first entry, first block, j-cluster
    // of the first entry): the first j positions load with the run's list data
    const int4 wd = A.wi[w];
    const int sc = (int)((unsigned)wd.x & (unsigned)kWiScMask);
    const int e0 = wd.y, ne = (int)((unsigned)wd.x >> kWiEntryShift), blk0 = wd.z;
    int4 cq = qpos[(size_t)wd.w * 8 + jl];
    float ce = qeps[(size_t)wd.w * 8 + jl];
    unsigned long long live = 0;
    {
        // keep tile ci of an entry unless byte_ci > bound (see kPruneSteps)
        float bound = 3.0e38f;
        if (A.disp) {
            float d = lane < kDispSlots ? __uint_as_float(A.disp[lane * kDispStride]) : 0.f;
            d = row_max16(d);  // lane 0 holds the largest of the kDispSlots (16) slots
            const float dmax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(d)));
            bound = fmaf(dmax, A.prune_scale, A.prune_bias);
        }
        // the run's mask bytes, 4 entries per row of 64 dwords: only the rows
        // that hold its entries are loaded (a test the whole wavefront shares);
        // the last row may run past the run into later entries, which the
        // entry buffers' spare capacity (1024 entries or more) keeps in bounds
        // and nothing reads
        const uint32_t *lb = reinterpret_cast<const uint32_t *>(ent_lb + (size_t)e0 * 64);
        uint32_t lbv[kMaxRun / 4];
#pragma unroll
        for (int r = 0; r < kMaxRun / 4; r++) lbv[r] = r * 4 < ne ? lb[lane + 64 * r] : 0u;
        int2 mv = make_int2(0, 0);
        if (lane < ne) {
            mv = A.ent_meta[e0 + lane];
            const uint2 dm = A.ent_dmin[e0 + lane];
            uint32_t keep = 0;
#pragma unroll
            for (int ci = 0; ci < 8; ci++) {
                const uint32_t b = ((ci < 4 ? dm.x : dm.y) >> (8 * (ci & 3))) & 255u;
                keep |= ((float)b > bound) ? 0u : (1u << ci);  // a NaN bound keeps every tile
            }
            mv.y &= (int)keep;
        }
        // Entries with every tile pruned contribute nothing: the run's
        // metadata is compacted to its live entries (in order), so the loop
        // below neither visits nor prefetches them.  Entry e goes to its rank
        // among them, its own index above the 8 mask bits (its pair-mask bytes
        // stay in place).
#pragma unroll
        for (int r = 0; r < kMaxRun / 4; r++)
            if (r * 4 < ne) s_lb[lane + 64 * r] = lbv[r];
        live = __ballot(mv.y != 0);
        if (mv.y != 0) s_meta[__popcll(live & ((1ull << lane) - 1ull))] = make_int2(mv.x, mv.y | (lane << 8));
    }
    const int nl = __popcll(live);  // live entries of the run
    // The supercluster's 64 i atoms live in LDS (fixed-point position + sigma/2
    // in one int4, 2 sqrt(eps) beside it): every tile reads 8 distinct entries,
    // broadcast to 8 lanes each, which keeps 40 registers free for occupancy.
    __shared__ int4 s_iq[64];
    __shared__ float s_ie[64];
    const int4 qown = qpos[(size_t)sc * 64 + lane];
    s_iq[lane] = qown;
    s_ie[lane] = qeps[(size_t)sc * 64 + lane];
    // an i atom whose position is not finite (sigma/2 stored as -0.0)?
    bool anybad = __ballot(qown.w < 0) != 0ull;
    float fix[8], fiy[8], fiz[8];
#pragma unroll
    for (int ci = 0; ci < 8; ci++) { fix[ci] = 0.f; fiy[ci] = 0.f; fiz[ci] = 0.f; }
    __syncthreads();
    const uint8_t *s_lb8 = reinterpret_cast<const uint8_t *>(s_lb);
    const float rc2 = CUBIC ? p.rc2_u : p.rc2;
    const unsigned rc2_bits = (unsigned)__float_as_int(rc2);
    // the switch offsets live in VGPRs: an FMA reads at most one scalar operand
    const float inv_w = CUBIC ? p.inv_w_u : p.inv_w, nrw = in_vgpr(p.nrw);
    const float ds_coef = CUBIC ? p.ds_coef_u : p.ds_coef;
    const float inv_w_in = CUBIC ? p.inv_w_in_u : p.inv_w_in, nrw_in = in_vgpr(p.nrw_in);
    const float ds_coef_in = CUBIC ? p.ds_coef_in_u : p.ds_coef_in;
    const float fscale = CUBIC ? p.fscale_u : 4294967296.0f;
    // Tile ci's list bit is bit ci of the lane's mask byte, tested by one
    // class compare: a class test's mask bit c selects IEEE class c, and
    // kcls[ci] is a value of class ci (signalling NaN, quiet NaN, -inf,
    // -normal, -denormal, -0, +0, +denormal).
    const float kcls[8] = {sgpr_bits(0x7f800001u), sgpr_bits(0x7fc00000u), sgpr_bits(0xff800000u),
                           sgpr_bits(0xbf800000u), sgpr_bits(0x80000001u), sgpr_bits(0x80000000u),
                           sgpr_bits(0x00000000u), sgpr_bits(0x00000001u)};
    if (nl > 0 && !(live & 1ull)) {  // the first entry was pruned: start from the first live one
        cq = qpos[(size_t)s_meta[0].x * 8 + jl];
        ce = qeps[(size_t)s_meta[0].x * 8 + jl];
    }
    // Blocks: entries [b kBlock, (b + 1) kBlock) of the run are block blk0 + b.
    // A block's i forces and energy leave once, after its last entry: records
    // (plain stores) or, without records, fixed point summed over the item's
    // blocks in registers and added with atomics at the end.
    const int nblk = (ne + kBlock - 1) / kBlock;
    long long ai[3] = {0, 0, 0};
    bool ai_ok = true;
    int k = 0;
    for (int b = 0; b < nblk; b++) {
        const int kend = __popcll(live & ((1ull << ((b + 1) * kBlock)) - 1ull));  // live entries up to block b
        const bool any = k < kend;
        double e_acc = 0.0;
        for (; k < kend; k++) {
            const int cjn = __builtin_amdgcn_readfirstlane(s_meta[min(k + 1, nl - 1)].x);
            const int4 nq = qpos[(size_t)cjn * 8 + jl];
            const float nce = qeps[(size_t)cjn * 8 + jl];
            const uint32_t mk = (uint32_t)__builtin_amdgcn_readfirstlane(s_meta[k].y);
            const uint32_t im = mk & 255u;
            float fjx = 0.f, fjy = 0.f, fjz = 0.f;
            if (im) {
                const uint32_t bits = s_lb8[(mk >> 8) * 64 + lane];
                const float shj = __int_as_float(cq.w), ej = ce;
                float ee = 0.f;
#pragma unroll
                for (int ci = 0; ci < 8; ci++) {
                    if (im & (1u << ci)) {
                        // Branch-free tile: every lane evaluates its pair; pairs not
                        // in the list, beyond r_cutoff or coincident get zero force
                        // and energy through one select each (which also discards
                        // the inf/NaN an r2 of 0 produces).  Below r_switch u
                        // clamps to 0, where S = 1 and dS/dr = 0 exactly.
                        const int4 qi = s_iq[ci * 8 + il];
                        float dx = fixdiff(cq.x, qi.x);
                        float dy = fixdiff(cq.y, qi.y);
                        float dz = fixdiff(cq.z, qi.z);
                        if (!CUBIC) { dx *= p.sx; dy *= p.sy; dz *= p.sz; }
                        const float r2p = dx * dx + dy * dy + dz * dz;
                        // 0 < r2 <= rc2 as one unsigned compare of the bit patterns
                        const bool ok = __builtin_amdgcn_classf(kcls[ci], (int)bits) &&
                                        ((unsigned)__float_as_int(r2p) - 1u) < rc2_bits;
                        // A pair that is not evaluated continues at r2 = 1e36 (far
                        // beyond any pair distance in either length unit): there
                        // (sigma/r)^6 underflows to +0 and u clamps to 1, where
                        // S = dS/dr = 0, so its force and energy are exactly +0,
                        // as a select of both would give.
                        const float r2 = ok ? r2p : 1.0e36f;
                        const float inv_r = __builtin_amdgcn_rsqf(r2);
                        // one Newton step brings rsq^2 (~2.5 ulp) to ~1 ulp, the
                        // accuracy (sigma/r)^12 needs, for two FMAs instead of a rcp
                        float inv_r2 = inv_r * inv_r;
                        inv_r2 = fmaf(inv_r2, fmaf(-r2, inv_r2, 1.0f), inv_r2);
                        const float sig = __int_as_float(qi.w) + shj;
                        const float s2 = sig * sig * inv_r2;
                        const float s6 = s2 * s2 * s2;
                        const float lj_a = fmaf(s6, s6, -s6);                 // s12 - s6
                        const float lj_b = s6 * fmaf(s6, 12.0f, -6.0f);       // 6 (2 s12 - s6)
                        const float e4 = s_ie[ci * 8 + il] * ej;
                        const float r = r2 * inv_r;
                        const float u = __builtin_amdgcn_fmed3f(fmaf(r, inv_w, nrw), 0.0f, 1.0f);
                        const float u2 = u * u;
                        const float pu = fmaf(u, fmaf(u, -6.0f, 15.0f), -10.0f);
                        const float qu = u - u2;
                        float sw, dsdr;
                        if (SHELL) {
                            const float ui = __builtin_amdgcn_fmed3f(fmaf(r, inv_w_in, nrw_in), 0.0f, 1.0f);
                            const float ui2 = ui * ui;
                            const float pi = fmaf(ui, fmaf(ui, -6.0f, 15.0f), -10.0f);
                            const float qi2 = ui - ui2;
                            sw = fmaf(u2 * u, pu, -(ui2 * ui * pi));           // S - S_in
                            dsdr = fmaf(ds_coef * qu, qu, -(ds_coef_in * qi2 * qi2));
                        } else {
                            sw = fmaf(u2 * u, pu, 1.0f);
                            dsdr = ds_coef * qu * qu;
                        }
                        const float inner = fmaf(-(lj_a * dsdr), r, lj_b * sw);
                        const float fr = e4 * inv_r2 * inner;
                        const float v = e4 * lj_a * sw;
                        fix[ci] = fmaf(-fr, dx, fix[ci]);
                        fiy[ci] = fmaf(-fr, dy, fiy[ci]);
                        fiz[ci] = fmaf(-fr, dz, fiz[ci]);
                        fjx = fmaf(fr, dx, fjx);
                        fjy = fmaf(fr, dy, fjy);
                        fjz = fmaf(fr, dz, fjz);
                        ee += v;
                    }
                }
                fjx = dpp_sum8(fjx);
                fjy = dpp_sum8(fjy);
                fjz = dpp_sum8(fjz);
                e_acc += (double)ee;
                anybad = anybad || __ballot(cq.w < 0) != 0ull;  // a j atom, likewise
            }
            if (il == 0) {
                jf[0][k * 8 + jl] = fjx;
                jf[1][k * 8 + jl] = fjy;
                jf[2][k * 8 + jl] = fjz;
            }
            cq = nq;
            ce = nce;
        }
        // i forces: lane (jl, il) holds partial sums for atom il of all 8 clusters;
        // the 8 jl lanes of each atom are reduced by a transpose: exchange halves
        // across lane bit 5, then bit 4 (gfx950 permlane swaps), then bit 3 (DPP),
        // keeping the half that belongs to the lane's own bit each time.  Lane
        // (jl, il) ends with the total of cluster jl, atom il: slot sc * 64 + lane.
        float fo[3] = {0.f, 0.f, 0.f};
        if (any) {
            const bool odd = (lane & 8) != 0;
#pragma unroll
            for (int c = 0; c < 3; c++) {  // one component at a time: few temporaries
                float *f = c == 0 ? fix : (c == 1 ? fiy : fiz);
                float h4[4];
#pragma unroll
                for (int q = 0; q < 4; q++) h4[q] = swap_add32(f[q], f[q + 4]);
                const float h20 = swap_add16(h4[0], h4[2]);
                const float h21 = swap_add16(h4[1], h4[3]);
                const float keep = odd ? h21 : h20;
                const float send = odd ? h20 : h21;
                fo[c] = keep + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(send), 0x128, 0xF, 0xF, false));
#pragma unroll
                for (int ci = 0; ci < 8; ci++) f[ci] = 0.f;
            }
        }
        if (A.fi_blk) {
            // the block's own record (plain stores; the finish kernel adds the
            // records of a supercluster's blocks): no atomics on the i side
            float *fw = A.fi_blk + (size_t)(blk0 + b) * 192 + lane;
            fw[0] = fo[0];
            fw[64] = fo[1];
            fw[128] = fo[2];
        } else if (any) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ai_ok = ai_ok && fabsf(fo[c] * fscale) < kFixedLimitF;
                ai[c] += to_fixed_f(fo[c], fscale);
            }
        }
        const double eb = any ? wave_sum_f64(e_acc) : 0.0;
        if (lane == 0) A.e_blk[blk0 + b] = eb;
    }
    __syncthreads();
    for (int idx = lane; idx < nl * 8; idx += 64) {
        const int2 m = s_meta[idx >> 3];
        const size_t js = (size_t)m.x * 8 + (idx & 7);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float v = jf[c][idx];
            if (v != 0.0f) fixed_add(&F[(size_t)c * n_pad + js], v, fscale, &A.bad[js]);
        }
    }
    if (!A.fi_blk && nl > 0) {  // large work items: into the accumulator (nothing to add for an empty run)
        const size_t s = (size_t)sc * 64 + lane;
        if (!ai_ok) {
            A.bad[s] = 1u;
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++)
                __hip_atomic_fetch_add(&F[(size_t)c * n_pad + s], (unsigned long long)ai[c], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // A listed pair with an atom whose position is not finite: the other
    // atom's force is reported as NaN too, as every listed pair of such an
    // atom gives NaN in f64 (rare; a pass of its own after everything else).
    // The energy is NaN then as well: the item's first block energy is
    // overwritten (the same wavefront stored it above).
    if (__builtin_expect(anybad, 0)) {
        const unsigned long long ib = __ballot(s_iq[lane].w < 0);
        bool hit = false;
        for (int kk = 0; kk < nl; kk++) {
            const int2 m = s_meta[kk];
            const size_t js = (size_t)m.x * 8 + jl;
            const bool jb = qpos[js].w < 0;
            const uint32_t bits = s_lb8[((uint32_t)m.y >> 8) * 64 + lane] & (uint32_t)m.y;
            for (int ci = 0; ci < 8; ci++) {
                if (!((bits >> ci) & 1u)) continue;
                const bool ibb = (ib >> (ci * 8 + il)) & 1ull;
                if (ibb) A.bad[js] = 1u;
                if (jb) A.bad[(size_t)sc * 64 + ci * 8 + il] = 1u;
                hit = hit || ibb || jb;
            }
        }
        if (__ballot(hit) != 0ull && lane == 0) A.e_blk[blk0] = __builtin_nan("");
    }
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

// ---- Partner-list path (the default) -----------------------------------------
// One wavefront per group of kPlAtoms consecutive slots, kPlLanes lanes per
// slot: lane l of a slot takes its list's entries l, l + 16, ... and keeps the
// slot's force in registers, so no force leaves a lane until the end, and
// every lane of a round evaluates a listed pair (the lists are dense; their
// padding is the slot itself, r2 = 0, skipped).  Each pair is evaluated from
// both of its slots with the same operands, so both slots see the same f32
// force (third law) and energy; the energy sums are halved.  A lane's f32 sum
// runs over fixed list positions and the 16 lanes combine in a fixed order,
// so results depend only on the lists, not on scheduling.
//
// Pruning: while no atom has moved farther than (kPlInner - 1e-4 r_cutoff) / 2
// since the build, every partner of a list's second part (farther than
// r_cutoff + kPlInner at the build) is now beyond r_cutoff, where the pair's
// terms are exactly zero: only the first part is evaluated.
// The kernel's arguments, PlArgs, are declared in gm_terms.hpp.

// Loads from device memory at byte address p, typed as global: pointers that
// come from a descriptor in memory (a batch's replica table) rather than from
// kernel arguments would otherwise be loaded as generic (flat) addresses with
// 64-bit arithmetic per load; through these the loads take a scalar base and
// a 32-bit offset, and a wave-uniform one becomes a scalar load.
typedef int gm_v4i __attribute__((ext_vector_type(4)));
typedef int gm_v2i __attribute__((ext_vector_type(2)));
__device__ inline int4 gload_i4(const char *p) {
    const gm_v4i v = *(const __attribute__((address_space(1))) gm_v4i *)(const __attribute__((address_space(1))) char *)p;
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ inline int2 gload_i2(const char *p) {
    const gm_v2i v = *(const __attribute__((address_space(1))) gm_v2i *)(const __attribute__((address_space(1))) char *)p;
    return make_int2(v.x, v.y);
}
__device__ inline int gload_i1(const char *p) {
    return *(const __attribute__((address_space(1))) int *)(const __attribute__((address_space(1))) char *)p;
}

// Item (of N) evaluated by workgroup W0 + k of a launch whose first W0
// workgroups do other work: workgroup w runs on XCD w % 8, and XCD x takes a
// contiguous range of items (neighbouring slots share one L2).
__device__ inline int xcd_group(int k, int W0, int N) {
    const int w = W0 + k, x = w & 7;
    int start = 0;
    for (int x2 = 0; x2 < x; x2++) {
        const int f2 = W0 + ((x2 - W0) & 7);
        start += f2 < W0 + N ? (W0 + N - f2 + 7) / 8 : 0;
    }
    return start + (w - (W0 + ((x - W0) & 7))) / 8;
}

// Sum over a slot's kPlLanes lanes, in every one of them (the same bits in
// each): the 8-lane sum, then the other 8 of the row of 16, then (32 lanes)
// the other row of the pair.
__device__ inline float slot_sum(float v) {
    v = dpp_sum8(v);
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xF, 0xF, false));
    if (kPlLanes == 32) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
        v = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    return v;
}

static_assert((64 / kPlAtoms) % kPlWaves == 0, "whole workgroups per supercluster");

// C16: the lists hold compact codes (slot minus the group's first slot plus
// kQposGuard, two rounds per 32-bit word); else 32-bit slot numbers.
// PT: the LJ parameters come from a table by type pair (sigma_ij^2 in the
// kernel's length unit, E4 = 4 eps_ij), T x T entries for T = 2^pt_shift <= 32
// types, filled with exactly the per-pair arithmetic it replaces (the sum of
// the two sigma/2, its square, the product of the two 2 sqrt eps): the same
// bits, three VALU instructions fewer per pair.  Else by type, as (sigma/2,
// 2 sqrt eps).
template <bool CUBIC, bool C16, bool PT>
__device__ __attribute__((always_inline)) inline void pl_work(const int g, const int lane, const PlArgs &A) {
    // Every fused multiply-add below is written out: the pair arithmetic is
    // inlined more than once (loop body and tail), and contraction left to the
    // compiler could fuse a different product in each copy, which would make a
    // pair's bits depend on the round it falls in.
#pragma clang fp contract(off)
    const NbParams &p = A.p;
    static_assert(kPlAtoms * kPlLanes == 64 && (kPlLanes == 16 || kPlLanes == 32), "one wavefront per group");
    const int al = lane / kPlLanes, ls = lane % kPlLanes;
    const int s = g * kPlAtoms + al;
    // LJ parameters by type, in LDS: sigma/2 in the kernel's length unit, 2 sqrt(eps)
    __shared__ float2 s_tp[kMaxLjTypes];
    __shared__ float2 s_pt[kPlPairTypes * kPlPairTypes];
    const int tsh = PT ? A.pt_shift : 0;
    if (PT) {
        const int T = 1 << tsh;
        for (int k = (int)threadIdx.x; k < T * T; k += (int)blockDim.x) {
            const int a = k >> tsh, b = k & (T - 1);
            float2 v = make_float2(0.f, 0.f);
            if (a < A.n_types && b < A.n_types) {
                const float2 ta = A.type_par[a], tb = A.type_par[b];
                const float sig = tb.x * A.sig_scale + ta.x * A.sig_scale;
                v = make_float2(sig * sig, ta.y * tb.y);
            }
            s_pt[k] = v;
        }
    } else {
        for (int k = (int)threadIdx.x; k < A.n_types; k += (int)blockDim.x) {
            const float2 t = A.type_par[k];
            s_tp[k] = make_float2(t.x * A.sig_scale, t.y);
        }
    }
    bool first_only = false;
    if (A.disp) {
        float d = lane < kDispSlots
                      ? __int_as_float(gload_i1(reinterpret_cast<const char *>(A.disp) + lane * kDispStride * 4))
                      : 0.f;
        d = row_max16(d);  // lane 0 holds the largest of the kDispSlots (16) words
        const float dmax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(d)));
        first_only = dmax <= A.dmax_first;  // (false for a NaN bound)
    }
    const int2 gl = gload_i2(reinterpret_cast<const char *>(A.glen) + (size_t)g * 8);
    const int rounds = (first_only ? gl.x : gl.y) / kPlLanes;
    // the group's list base (wave-uniform) and the lane's byte offset in it;
    // loads address base + 32-bit offset, so that no 64-bit address is formed per load
    const char *__restrict__ nbase =
        reinterpret_cast<const char *>(A.nbr) +
        (size_t)gload_i1(reinterpret_cast<const char *>(A.goff) + (size_t)g * 4) * (C16 ? 2 : 4);
    const unsigned lofs = C16 ? (unsigned)(al * gl.y) * 2u + (unsigned)ls * 4u : (unsigned)(al * gl.y + ls) * 4u;
    const char *__restrict__ qpb = reinterpret_cast<const char *>(A.qpos);
    auto qpos_of = [&](int c) { return gload_i4(qpb + ((unsigned)c << 4)); };
    // a partner's slot data from its list entry: the slot number, or the
    // compact code (from the guard start, offset by the group's first slot)
    const char *__restrict__ qpg = reinterpret_cast<const char *>(A.qpos - kQposGuard);
    const unsigned gslot = (unsigned)(g * kPlAtoms);
    auto partner = [&](int c) {
        return C16 ? gload_i4(qpg + ((gslot + (unsigned)c) << 4)) : qpos_of(c);
    };
    // the entries of rounds r and r + 1 (r even)
    auto codes2 = [&](int r, int &c0, int &c1) {
        if (C16) {
            const unsigned v = (unsigned)gload_i1(nbase + lofs + (unsigned)(r >> 1) * 64u);
            c0 = (int)(v & 0xffffu);
            c1 = (int)(v >> 16);
        } else {
            c0 = gload_i1(nbase + lofs + r * (kPlLanes * 4));
            c1 = gload_i1(nbase + lofs + (r + 1) * (kPlLanes * 4));
        }
    };
    // a slot word's type parameters (its low 11 bits are the table offset)
    auto par_of = [&](int w) {
        return *reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(s_tp) + (w & 0x7f8));
    };
    const int4 qi = qpos_of(s);
    __syncthreads();  // s_tp / s_pt
    const float2 ti = PT ? make_float2(0.f, 0.f) : par_of(qi.w);
    const float shi = ti.x, ei = ti.y;
    // PT: the row of slot i's type (a type offset past the table, which only a
    // slot without an atom can hold, is folded into it: its pairs are never
    // evaluated, and every entry is finite)
    const unsigned tmask = ((8u << tsh) - 1u) & ~7u;
    const unsigned ti_row = ((unsigned)qi.w & tmask) << tsh;
    const float rc2 = CUBIC ? p.rc2_u : p.rc2;
    const unsigned rc2_bits = (unsigned)__float_as_int(rc2);
    const float inv_w = CUBIC ? p.inv_w_u : p.inv_w, nrw = in_vgpr(p.nrw);
    const float ds_coef = CUBIC ? p.ds_coef_u : p.ds_coef;
    const float fscale = CUBIC ? p.fscale_u : 4294967296.0f;
    float fx = 0.f, fy = 0.f, fz = 0.f, ee = 0.f;
    int wmin = qi.w;  // smallest slot word seen: negative iff a position is not finite
    // One pair of the list: slot i against partner slot data qj.
    auto pair = [&](const int4 qj) {
        float dx = fixdiff(qj.x, qi.x);
        float dy = fixdiff(qj.y, qi.y);
        float dz = fixdiff(qj.z, qi.z);
        if (!CUBIC) { dx *= p.sx; dy *= p.sy; dz *= p.sz; }
        const float r2p = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        // 0 < r2 <= rc2 as one unsigned compare of the bit patterns; a pair not
        // evaluated continues at r2 = 1e36, where its terms are exactly +0
        const bool ok = ((unsigned)__float_as_int(r2p) - 1u) < rc2_bits;
        const float r2 = ok ? r2p : 1.0e36f;
        float sig2, e4;
        if (PT) {
            const float2 q = *reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(s_pt) +
                                                               (((unsigned)qj.w & tmask) | ti_row));
            sig2 = q.x;
            e4 = q.y;
        } else {
            const float2 tj = par_of(qj.w);
            const float sig = tj.x + shi;
            sig2 = sig * sig;
            e4 = ei * tj.y;
        }
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        float inv_r2 = inv_r * inv_r;
        inv_r2 = fmaf(inv_r2, fmaf(-r2, inv_r2, 1.0f), inv_r2);
        const float s2 = sig2 * inv_r2;
        const float s6 = s2 * s2 * s2;
        const float lj_a = fmaf(s6, s6, -s6);            // s12 - s6
        const float lj_b = s6 * fmaf(s6, 12.0f, -6.0f);  // 6 (2 s12 - s6)
        const float rr = r2 * inv_r;
        const float u = __builtin_amdgcn_fmed3f(fmaf(rr, inv_w, nrw), 0.0f, 1.0f);
        const float u2 = u * u;
        const float pu = fmaf(u, fmaf(u, -6.0f, 15.0f), -10.0f);
        const float qu = fmaf(-u, u, u);
        const float sw = fmaf(u2 * u, pu, 1.0f);
        const float dsdr = ds_coef * qu * qu;
        const float inner = fmaf(-(lj_a * dsdr), rr, lj_b * sw);
        const float fr = e4 * inv_r2 * inner;
        fx = fmaf(-fr, dx, fx);
        fy = fmaf(-fr, dy, fy);
        fz = fmaf(-fr, dz, fz);
        ee = fmaf(e4 * lj_a, sw, ee);
        wmin = min(wmin, qj.w);
    };
    // Four rounds per iteration, in two halves of two, with two register
    // sets: while one half's partner data is used, the other's is on its way,
    // requested a whole half earlier, and list codes arrive a half before
    // their partner data is requested.  (The empty asm statements are
    // compiler barriers: they keep the loads issued ahead of the arithmetic
    // instead of next to their first use.)  Codes up to three rounds past a
    // part are read ahead and not all used: they are later list entries or
    // the list buffer's slack, all valid slots.
    if (rounds > 0) {
        int a0, a1, b0, b1;
        codes2(0, a0, a1);
        codes2(2, b0, b1);
        int4 qa0 = partner(a0), qa1 = partner(a1);
        int r = 0;
        for (; r + 4 <= rounds; r += 4) {
            int n0, n1;
            codes2(r + 4, n0, n1);
            const int4 qb0 = partner(b0), qb1 = partner(b1);
            asm volatile("" ::: "memory");
            pair(qa0);
            pair(qa1);
            codes2(r + 6, b0, b1);
            qa0 = partner(n0);
            qa1 = partner(n1);
            asm volatile("" ::: "memory");
            pair(qb0);
            pair(qb1);
        }
        const int left = rounds - r;  // 0 to 3 rounds after the last whole iteration
        if (left >= 1) pair(qa0);
        if (left >= 2) pair(qa1);
        if (left >= 3) pair(partner(b0));
    }
    fx = slot_sum(fx);
    fy = slot_sum(fy);
    fz = slot_sum(fz);
    // A slot whose position, or a listed partner's, is not finite (rare): its
    // force is reported as NaN, and so is the energy when such a pair is listed.
    bool hit = false;
    if (__builtin_expect(__ballot(wmin < 0) != 0ull, 0)) {
        for (int r = 0; r < rounds; r += 2) {
            int c[2];
            codes2(r, c[0], c[1]);
            for (int h = 0; h < 2 && r + h < rounds; h++) {
                const int sj = C16 ? (int)gslot + c[h] - kQposGuard : c[h];  // a listed slot (or s)
                hit = hit || (sj != s && (qi.w < 0 || A.qpos[sj].w < 0));
            }
        }
        const bool row_hit = (__ballot(hit) >> (kPlLanes * al)) & ((1ull << kPlLanes) - 1ull);
        if (row_hit && ls == 0) A.bad[s] = 1u;
    }
    // The slot's whole LJ force is this wavefront's: a plain store into the
    // accumulator (nothing else adds to it in a partner-list evaluation; the
    // bonded terms write their own records, which the finish kernel adds).
    if (ls == 0) {
        const bool okf = fabsf(fx * fscale) < kFixedLimitF && fabsf(fy * fscale) < kFixedLimitF &&
                         fabsf(fz * fscale) < kFixedLimitF;
        if (!okf) A.bad[s] = 1u;
        A.F[s] = okf ? (unsigned long long)to_fixed_f(fx, fscale) : 0ull;
        A.F[(size_t)A.n_pad + s] = okf ? (unsigned long long)to_fixed_f(fy, fscale) : 0ull;
        A.F[2 * (size_t)A.n_pad + s] = okf ? (unsigned long long)to_fixed_f(fz, fscale) : 0ull;
    }
    const double eg = wave_sum_f64((double)ee);
    if (lane == 0) A.e_grp[g] = __ballot(hit) != 0ull ? __builtin_nan("") : 0.5 * eg;
}

// Workgroups of kPlWaves wavefronts: the bonded workgroups first (a term per
// thread; they store force records), then one group per wavefront (kPlWaves
// consecutive groups per workgroup; the group count, 64 / kPlAtoms per
// supercluster, is a multiple).
template <bool CUBIC, bool C16, bool PT>
__global__ __launch_bounds__(64 * kPlWaves) __attribute__((amdgpu_waves_per_eu(8))) void k_forces_pl(PlArgs A,
                                                                                                    TermArgs ta) {
    const int64_t n_terms = ta.nb + ta.na + ta.nd + ta.n14;
    const int n_bonded_wg = (int)((n_terms + 64 * kPlWaves - 1) / (64 * kPlWaves));
    if ((int)blockIdx.x < n_bonded_wg) {
        const int64_t k = (int64_t)blockIdx.x * (64 * kPlWaves) + threadIdx.x;
        if (k < n_terms) term_dispatch(ta, k);
        return;
    }
    const int wg = xcd_group((int)blockIdx.x - n_bonded_wg, n_bonded_wg, A.n_grp / kPlWaves);
    // (the group is the same in the whole wavefront: said so, everything
    // derived from it lives in scalar registers)
    const int g = __builtin_amdgcn_readfirstlane(wg * kPlWaves + (int)(threadIdx.x >> 6));
    pl_work<CUBIC, C16, PT>(g, (int)(threadIdx.x & 63), A);
}

// Persistent batched tile kernel: a fixed grid of wavefronts takes tile work
// items (replica, work item) from eight counters, one per XCD.  Workgroups
// are dealt to the 8 XCDs in turn, so wavefront w runs on XCD w % 8 and takes
// items of its XCD's contiguous range [x n / 8, (x + 1) n / 8) first (it
// keeps neighbouring items, same replica and nearby superclusters, in one
// L2), then helps the other XCDs' ranges in turn.  The counters (kCtrStride
// words apart) are cleared by the batched prep kernel.  The next ticket is
// fetched before the current item is evaluated, so its latency overlaps the
// item.  Each item's arithmetic is the single-system kernel's, so results do
// not depend on which wavefront takes which item.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_batch_tiles(
    const ReplicaDesc *__restrict__ reps, const int2 *__restrict__ items, int n_items, unsigned *ctr) {
    const int lane = threadIdx.x;
    const int x = blockIdx.x & 7;
    for (int d = 0; d < 8; d++) {
        const int xx = (x + d) & 7;
        const int lo = (int)(((int64_t)n_items * xx) >> 3), hi = (int)(((int64_t)n_items * (xx + 1)) >> 3);
        unsigned *c = ctr + xx * kCtrStride;
        unsigned t = 0;
        if (lane == 0) t = __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int it = lo + (int)__builtin_amdgcn_readfirstlane(t);
        while (it < hi) {
            unsigned tn = 0;
            if (lane == 0) tn = __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // (replica, item) is the same for every lane: scalar registers
            const int rep = __builtin_amdgcn_readfirstlane(items[it].x);
            const int wi = __builtin_amdgcn_readfirstlane(items[it].y);
            tile_work<true, false>(wi, lane, reps[rep].tile);  // replicas are cubic
            it = lo + (int)__builtin_amdgcn_readfirstlane(tn);
        }
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

// Partner-list groups of every replica: one workgroup per item (replica, first
// of kPlWaves consecutive groups; a replica's group count is a multiple of
// kPlWaves, so a workgroup never spans two replicas and its LDS type table is
// one replica's).  Each XCD takes a contiguous range of items, so neighbouring
// groups of one replica share its L2.  Every group runs the single-system
// code: a replica's forces and energies are those of evaluating it alone.
// Replicas are cubic and their lists hold 16-bit codes (the batch checks both).
template <bool PT>
__global__ __launch_bounds__(64 * kPlWaves) __attribute__((amdgpu_waves_per_eu(8))) void k_batch_pl(
    const ReplicaDesc *__restrict__ reps, const int2 *__restrict__ items, int n_items) {
    const int2 it = items[xcd_group((int)blockIdx.x, 0, n_items)];
    // (replica and group are the same in the whole wavefront: scalar registers)
    const int r = __builtin_amdgcn_readfirstlane(it.x);
    const int g = __builtin_amdgcn_readfirstlane(it.y + (int)(threadIdx.x >> 6));
    pl_work<true, true, PT>(g, (int)(threadIdx.x & 63), reps[r].pl);
}

}  // namespace

namespace gm {

TermArgs term_args(gm_sys *s, const double *x, const double *y, const double *z) {
    TermArgs t{};
    t.x = x; t.y = y; t.z = z;
    for (int a = 0; a < 3; a++) t.box[a] = s->box[a];
    t.mi = s->bonded_mi;
    t.F = s->ev.F;
    t.bad = s->ev.bad;
    t.atom_slot = s->atom_slot.p;
    t.n_pad = (size_t)s->nsc * 64;
    t.bi = s->bi.p; t.bj = s->bj.p; t.bk = s->bk.p; t.br0 = s->br0.p;
    t.ai = s->ai.p; t.aj = s->aj.p; t.ak = s->ak.p; t.akth = s->akth.p; t.ath0 = s->ath0.p;
    t.di = s->di.p; t.dj = s->dj.p; t.dk = s->dk.p; t.dl = s->dl.p; t.dn = s->dn.p;
    t.dkphi = s->dkphi.p; t.dphase = s->dphase.p;
    t.p14i = s->p14i.p; t.p14j = s->p14j.p;
    t.sigma = s->sigma.p; t.epsilon = s->eps.p;
    t.scale14 = s->scale14;
    t.nb = s->nb; t.na = s->na; t.nd = s->nd;
    t.n14 = (s->scale14 != 0.0) ? s->n14 : 0;
    t.e_items = s->ev.e_items;
    t.rec = s->ev.brec;
    t.rec_pos = s->bc_pos.p;
    t.n_rec = s->n_contrib;
    return t;
}

TileArgs tile_args(gm_sys *s) {
    TileArgs A{};
    A.qpos = s->qpos.p;
    A.qeps = s->qeps.p;
    A.wi = s->wi.p;
    A.ent_meta = s->ent_meta.p; A.ent_dmin = s->ent_dmin.p; A.ent_lb = s->ent_lb.p;
    A.disp = (s->prune && s->prune_geom) ? s->ev.disp : nullptr;
    // keep iff byte <= (2 disp + 1e-4 r_c) / step + 1 (see kPruneSteps)
    A.prune_scale = 2.0f * s->bp.prune_inv_step;
    A.prune_bias = (float)(1e-4 * s->rc) * s->bp.prune_inv_step + 1.0f;
    A.F = s->ev.F;
    A.bad = s->ev.bad;
    A.bad_first = bad_first_word(s);
    A.n_pad = (int)(s->nsc * 64);
    A.e_blk = s->ev.e_blk;
    A.fi_blk = s->ev.fi_blk;
    A.sc_blk_off = s->sc_blk_off.p;
    A.fi_scale = span_is_cubic(s) ? s->nbp.fscale_u : 4294967296.0f;  // the tile kernel's fscale
    A.slot_atom = s->slot_atom.p;
    A.p = s->nbp;
    A.cubic = span_is_cubic(s) ? 1 : 0;
    return A;
}

PrepArgs prep_args(gm_sys *s, const double *x, const double *y, const double *z) {
    PrepArgs P{};
    P.x = x; P.y = y; P.z = z;
    P.atom_slot = s->atom_slot.p;
    P.qref = s->qatom.p;
    P.qpos = s->qpos.p;
    P.atom_par = s->atom_par.p;
    P.sig_scale = span_is_cubic(s) ? s->nbp.inv_s : 1.0f;  // as the list build's sigma/2
    P.disp = s->ev.disp;
    P.bad = s->ev.bad;
    P.bad_first = bad_first_word(s);
    for (int a = 0; a < 3; a++) {
        P.iq[a] = s->bp.inv_q[a];
        P.sq_len[a] = s->bp.sq[a];
    }
    P.n = (int)s->n;
    return P;
}

// Slot-order fixed-point positions and the displacement bound; precedes every
// launch_forces on the same positions.
hipError_t launch_prep(gm_sys *s, const double *x, const double *y, const double *z, hipEvent_t ev_start,
                       bool pl) {
    if (s->n == 0) return ev_start ? hipEventRecord(ev_start, s->stream) : hipSuccess;
    const unsigned grid = (unsigned)((s->n + 255) / 256);
    PrepArgs P = prep_args(s, x, y, z);
    if (pl) P.wtype = s->atom_wtype.p;  // partner lists: type-table offsets in the slot words
    if (ev_start)  // stamped by the dispatch at the kernel's start
        hipExtLaunchKernelGGL(k_prep, dim3(grid), dim3(256), 0, s->stream, ev_start, nullptr, 0, P);
    else
        k_prep<<<grid, 256, 0, s->stream>>>(P);
    return hipGetLastError();
}

PlArgs pl_args(gm_sys *s) {
    PlArgs A{};
    A.qpos = s->qpos.p;
    A.type_par = s->type_par.p;
    A.n_types = s->n_types;
    A.sig_scale = span_is_cubic(s) ? s->nbp.inv_s : 1.0f;  // as the prep kernel's sigma/2
    A.goff = s->pl_goff.p;
    A.glen = s->pl_glen.p;
    A.nbr = s->pl_nbr.p;
    A.disp = (s->prune && s->prune_geom) ? s->ev.disp : nullptr;
    // 2 d + 1e-4 r_cutoff <= kPlInner: the second parts lie beyond r_cutoff
    A.dmax_first = (float)((kPlInner - 1e-4 * s->rc) * 0.5);
    A.F = s->ev.F;
    A.bad = s->ev.bad;
    A.e_grp = s->ev.e_blk;
    A.n_pad = (int)(s->nsc * 64);
    A.n_grp = (int)(s->nsc * (64 / kPlAtoms));
    A.p = s->nbp;
    // the pair table's size: the types rounded up to a power of two, at most kPlPairTypes
    int sh = 0;
    while ((1 << sh) < s->n_types) sh++;
    A.pt_shift = (1 << sh) <= kPlPairTypes ? sh : -1;
    return A;
}

// One launch: the bonded workgroups, then the partner-list groups (or the
// tile work items when the system has no partner lists).
// Shell mode: tile work items only.
hipError_t launch_forces(gm_sys *s, const double *x, const double *y, const double *z, bool shell,
                         hipEvent_t ev_start, hipEvent_t ev_stop) {
    if (!shell && pl_active(s)) {
        const TermArgs ta = term_args(s, x, y, z);
        const PlArgs A = pl_args(s);
        const int64_t grid = (int64_t)A.n_grp / kPlWaves + (ta.nb + ta.na + ta.nd + ta.n14 + 64 * kPlWaves - 1) /
                                                            (64 * kPlWaves);
        if (grid > 0) {
            const bool pt = A.pt_shift >= 0;
            auto kern = span_is_cubic(s)
                            ? (s->pl_wide ? (pt ? k_forces_pl<true, false, true> : k_forces_pl<true, false, false>)
                                          : (pt ? k_forces_pl<true, true, true> : k_forces_pl<true, true, false>))
                            : (s->pl_wide ? (pt ? k_forces_pl<false, false, true> : k_forces_pl<false, false, false>)
                                          : (pt ? k_forces_pl<false, true, true> : k_forces_pl<false, true, false>));
            if (ev_start || ev_stop)
                hipExtLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * kPlWaves), 0, s->stream, ev_start,
                                      ev_stop, 0, A, ta);
            else
                kern<<<(unsigned)grid, 64 * kPlWaves, 0, s->stream>>>(A, ta);
        } else {
            hipError_t e = hipSuccess;
            if (ev_start) e = hipEventRecord(ev_start, s->stream);
            if (e == hipSuccess && ev_stop) e = hipEventRecord(ev_stop, s->stream);
            if (e != hipSuccess) return e;
        }
        return hipGetLastError();
    }
    TermArgs ta = term_args(s, x, y, z);
    TileArgs A = tile_args(s);
    if (shell) {
        ta.nb = ta.na = ta.nd = ta.n14 = 0;
        A.p = s->nbp_shell;
    }
    const int64_t nterms = ta.nb + ta.na + ta.nd + ta.n14;
    const int64_t grid = s->n_wi + (nterms + 63) / 64;
    if (grid > 0) {
        auto kern = shell ? (A.cubic ? k_forces<true, true> : k_forces<false, true>)
                          : (A.cubic ? k_forces<true, false> : k_forces<false, false>);
        if (ev_start || ev_stop)  // the dispatch itself stamps the events: the kernel's own span
            hipExtLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64), 0, s->stream, ev_start, ev_stop, 0, A, ta);
        else
            kern<<<(unsigned)grid, 64, 0, s->stream>>>(A, ta);
    } else {
        hipError_t e = hipSuccess;
        if (ev_start) e = hipEventRecord(ev_start, s->stream);
        if (e == hipSuccess && ev_stop) e = hipEventRecord(ev_stop, s->stream);
        if (e != hipSuccess) return e;
    }
    return hipGetLastError();
}

hipError_t launch_batch_prep(const ReplicaDesc *reps, const int *atom_rep, const int64_t *atom_off,
                             int64_t n_atoms, unsigned *ctr, hipStream_t st) {
    // at least one block, which clears the tile kernel's counters
    k_batch_prep<<<(unsigned)std::max<int64_t>((n_atoms + 255) / 256, 1), 256, 0, st>>>(reps, atom_rep, atom_off,
                                                                                         n_atoms, ctr);
    return hipGetLastError();
}

hipError_t launch_batch_tiles(const ReplicaDesc *reps, const int2 *items, int n_items, int grid, unsigned *ctr,
                              hipStream_t st) {
    if (n_items > 0 && grid > 0) k_batch_tiles<<<(unsigned)grid, 64, 0, st>>>(reps, items, n_items, ctr);
    return hipGetLastError();
}

hipError_t launch_batch_pl(const ReplicaDesc *reps, const int2 *items, int n_items, bool pair_table,
                           hipStream_t st) {
    if (n_items > 0 && pair_table)
        k_batch_pl<true><<<(unsigned)n_items, 64 * kPlWaves, 0, st>>>(reps, items, n_items);
    else if (n_items > 0)
        k_batch_pl<false><<<(unsigned)n_items, 64 * kPlWaves, 0, st>>>(reps, items, n_items);
    return hipGetLastError();
}

hipError_t launch_batch_bonded(const ReplicaDesc *reps, const int2 *items, int n_items, hipStream_t st) {
    if (n_items > 0) k_batch_bonded<<<(unsigned)n_items, 64, 0, st>>>(reps, items);
    return hipGetLastError();
}

}  // namespace gm
