// End of a force evaluation, gfx950: per-term energy reduction, f64 force
// output in atom order, and clearing of the int64 accumulator for the next
// evaluation — one launch.
//
// Energies are summed by a fixed-order f64 tree (the reference sums bonded
// terms, then per-thread LJ partials, in a fixed order: cpu_backend.rs:93-166;
// paths relative to the reference repository root).  Forces leave the
// fixed-point accumulator as value * 2^-32 kcal/mol/A.
#include <hip/hip_ext.h>

#include "gm_internal.hpp"
#include "gm_terms.hpp"

using namespace gm;

namespace {

struct Segs {
    const double *ptr[8];
    long long len[8];
    int nseg;
};

// Fixed-order f64 sum of one segment by a 1024-thread block: in round r,
// thread t adds item 16384 r + 1024 k + t to running sum k (k = 0..15; an item
// past the end adds 0), so sixteen loads are in flight per round (the loads
// are unconditional, at an index clamped into the segment).  The sixteen
// sums combine as a pairwise tree, each wavefront sums its lanes with DPP and
// permlane steps, and thread 0 adds the 16 wave sums in order.  The result is
// valid in thread 0.
__device__ inline double block_sum(const double *ptr, long long len, double *sh) {
    constexpr int K = 8;
    const int t = threadIdx.x;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0.0;
    for (long long base = 0; base < len; base += K * 1024) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const long long i = base + k * 1024 + t;
            const double v = ptr[i < len ? i : len - 1];
            acc[k] += i < len ? v : 0.0;
        }
    }
#pragma unroll
    for (int w = K / 2; w >= 1; w /= 2)
#pragma unroll
        for (int k = 0; k < w; k++) acc[k] = acc[k] + acc[k + w];
    double a = wave_sum_f64(acc[0]);
    if ((t & 63) == 0) sh[t >> 6] = a;
    __syncthreads();
    double r = 0.0;
    if (t == 0)
        for (int w = 0; w < 16; w++) r += sh[w];
    return r;
}

// v * scale rounded to an integer, as the force kernel rounds the i forces
// it adds itself (so records and atomics give the same integers).
__device__ inline long long to_fixed(float v, float scale) { return (long long)__builtin_rintf(v * scale); }

// Forces of one supercluster's 64 slots, by G (1, 4, 8 or 16) wavefronts of a
// 1024-thread workgroup (wavefronts G g .. G g + G - 1 for the workgroup's
// supercluster g): wavefront q adds the i-force records of the
// supercluster's blocks q, q + G, ... in fixed point; the first then adds the
// others' sums, the accumulator (j forces, bonded terms; or the partner-list
// forces) and, when brec is given, the slot's bonded force records
// [srec[s].x, srec[s].y), clears the accumulator and writes each slot's f64
// force to its atom (atom index + atom_base).  Integer sums: the order does not change the result.  A slot
// marked in bad (a contribution that was not finite or out of the fixed-point
// range, or a position that was not finite), or with a record that is, gets
// NaN forces, and its atom is raised into *bad_first (atomic minimum).  Every
// thread of the workgroup calls it (it holds a barrier); valid says whether
// group g has a supercluster.
__device__ inline void sc_finish(bool valid, int sc, int G, unsigned long long *F, size_t n_pad,
                                 const float *fi_blk, const int *sc_blk_off, float fi_scale,
                                 const int *slot_atom, int64_t atom_base, double *fx, double *fy, double *fz,
                                 unsigned *bad, int *bad_first, const long long *brec = nullptr,
                                 int64_t n_rec = 0, const int2 *srec = nullptr) {
    __shared__ long long red[16][3][64];
    __shared__ unsigned redb[16][64];
    const int t = threadIdx.x, wv = t >> 6, l = t & 63, q = wv & (G - 1);
    const size_t s = (size_t)sc * 64 + l;
    // the summing wavefront's accumulator words and atom index load with the records
    unsigned long long f[3] = {0ull, 0ull, 0ull};
    int a = -1;
    unsigned bd = 0u;
    int2 rr = make_int2(0, 0);
    if (valid && q == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) f[c] = F[(size_t)c * n_pad + s];
        a = slot_atom[s];
        bd = bad[s];
        if (brec) rr = srec[s];
    }
    long long ax = 0, ay = 0, az = 0;
    unsigned rb = 0u;
    if (valid && fi_blk) {  // (null: the i forces are already in the accumulator)
        // the supercluster's records are consecutive: wavefront q adds every G-th
        const int k0 = sc_blk_off[sc], k1 = sc_blk_off[sc + 1];
#pragma unroll 2
        for (int k = k0 + q; k < k1; k += G) {
            const float *fw = fi_blk + (size_t)k * 192 + l;
            const float vx = fw[0], vy = fw[64], vz = fw[128];
            const float m = fmaxf(fabsf(vx), fmaxf(fabsf(vy), fabsf(vz)));
            rb |= (!(m * fi_scale < kFixedLimitF) || __builtin_isnan(vx) || __builtin_isnan(vy) ||
                   __builtin_isnan(vz)) ? 1u : 0u;
            ax += to_fixed(vx, fi_scale);
            ay += to_fixed(vy, fi_scale);
            az += to_fixed(vz, fi_scale);
        }
    }
    red[wv][0][l] = ax;
    red[wv][1][l] = ay;
    red[wv][2][l] = az;
    redb[wv][l] = rb;
    __syncthreads();
    if (!valid || q != 0) return;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        for (int o = 0; o < G; o++) f[c] += (unsigned long long)red[wv + o][c][l];
        F[(size_t)c * n_pad + s] = 0ull;
    }
    for (int o = 0; o < G; o++) bd |= redb[wv + o][l];
    if (bd) bad[s] = 0u;
    if (a < 0) return;
#pragma unroll 4
    for (int k = rr.x; k < rr.y; k++) {
        f[0] += (unsigned long long)brec[k];
        f[1] += (unsigned long long)brec[n_rec + k];
        f[2] += (unsigned long long)brec[2 * n_rec + k];
    }
    if (bd) {
        fx[atom_base + a] = __builtin_nan("");
        fy[atom_base + a] = __builtin_nan("");
        fz[atom_base + a] = __builtin_nan("");
        atomicMin(bad_first, a);
        return;
    }
    fx[atom_base + a] = (double)(long long)f[0] * kFixedInv;
    fy[atom_base + a] = (double)(long long)f[1] * kFixedInv;
    fz[atom_base + a] = (double)(long long)f[2] * kFixedInv;
}

// Blocks [0, nseg): block b sums segment b.  Blocks >= nseg: 16 / G
// superclusters each (sc_finish).  Block 0 also clears the displacement bound
// the prep kernel raised.
__global__ __launch_bounds__(1024) void k_finish(Segs sg, double *eout, const int *slot_atom,
                                                 unsigned long long *F, size_t n_pad,
                                                 const float *fi_blk, const int *sc_blk_off,
                                                 float fi_scale, int G, double *fx,
                                                 double *fy, double *fz, unsigned *disp, unsigned *bad,
                                                 int *bad_first, const long long *brec, int64_t n_rec,
                                                 const int2 *srec) {
    __shared__ double sh[16];
    const int b = blockIdx.x, t = threadIdx.x;
    if (b == 0 && t < kDispSlots) disp[t * kDispStride] = 0u;
    if (b < sg.nseg) {
        const double v = block_sum(sg.ptr[b], sg.len[b], sh);
        if (t == 0) eout[b] = v;
        return;
    }
    const int sc = (b - sg.nseg) * (16 / G) + (t >> 6) / G;
    sc_finish((size_t)sc * 64 < n_pad, sc, G, F, n_pad, fi_blk, sc_blk_off, fi_scale, slot_atom, 0,
              fx, fy, fz, bad, bad_first, brec, n_rec, srec);
}

// Batched end of evaluation: blocks [0, R) sum replica b's five segments
// (bond, angle, dihedral, 1-4, LJ) in turn, each exactly like k_finish; the remaining
// blocks finish (replica, supercluster) pairs of scl exactly like k_finish,
// writing forces at the replica's atom offset: the first n_rec pairs (of
// replicas that keep i-force records) four per block with four wavefronts
// each, the next n_atm sixteen per block with one wavefront each (which also
// adds the bonded force records of a partner-list replica).
__global__ __launch_bounds__(1024) void k_batch_finish(const ReplicaDesc *__restrict__ reps, int n_reps,
                                                       const int2 *__restrict__ scl, int64_t n_rec, int64_t n_atm,
                                                       const int64_t *__restrict__ atom_off, double *eout,
                                                       double *fx, double *fy, double *fz) {
    __shared__ double sh[1024];
    const int b = blockIdx.x, t = threadIdx.x;
    if (b < n_reps) {  // the replica's five segments in turn
        const ReplicaDesc &R = reps[b];
        if (t < kDispSlots) R.prep.disp[t * kDispStride] = 0u;  // bound of the next evaluation
        const double *eb = R.terms.e_items;
        const double *ptr[5] = {eb, eb + R.terms.nb, eb + R.terms.nb + R.terms.na,
                                eb + R.terms.nb + R.terms.na + R.terms.nd, R.tile.e_blk};
        const long long len[5] = {R.terms.nb, R.terms.na, R.terms.nd, R.terms.n14, R.n_blk};
        for (int seg = 0; seg < 5; seg++) {
            const double v = block_sum(ptr[seg], len[seg], sh);
            if (t == 0) eout[5 * b + seg] = v;
            __syncthreads();  // sh is reused by the next segment
        }
        return;
    }
    // the rest: four (replica, supercluster) pairs with records, or sixteen without, per workgroup
    const int64_t b_rec = (n_rec + 3) / 4, bb = b - n_reps;
    const int G = bb < b_rec ? 4 : 1;
    const int64_t pidx = bb < b_rec ? bb * 4 + (t >> 6) / 4 : n_rec + (bb - b_rec) * 16 + (t >> 6);
    const bool valid = bb < b_rec ? pidx < n_rec : pidx < n_rec + n_atm;
    const int2 rs = valid ? scl[pidx] : make_int2(0, 0);
    const ReplicaDesc &R = reps[rs.x];
    // (a partner-list replica's bonded forces are records: R.terms.rec; else null)
    sc_finish(valid, rs.y, G, R.terms.F, R.terms.n_pad, R.tile.fi_blk, R.tile.sc_blk_off,
              R.tile.fi_scale, R.tile.slot_atom, atom_off[rs.x], fx, fy, fz, R.tile.bad, R.tile.bad_first,
              R.terms.rec, R.terms.n_rec, R.srec);
}

}  // namespace

namespace gm {

hipError_t launch_batch_finish(const ReplicaDesc *reps, int n_reps, const int2 *scl, int64_t n_rec, int64_t n_atm,
                               const int64_t *atom_off, double *eout, double *fx, double *fy, double *fz,
                               hipStream_t st) {
    const unsigned grid = (unsigned)n_reps + (unsigned)((n_rec + 3) / 4 + (n_atm + 15) / 16);
    k_batch_finish<<<grid, 1024, 0, st>>>(reps, n_reps, scl, n_rec, n_atm, atom_off, eout, fx, fy, fz);
    return hipGetLastError();
}

hipError_t launch_finish(gm_sys *s, double *fx, double *fy, double *fz, bool lj_only, hipEvent_t ev_stop) {
    Segs sg{};
    const double *eb = s->ev.e_items;
    const int64_t n14 = (s->scale14 != 0.0) ? s->n14 : 0;
    const int64_t nb = lj_only ? 0 : s->nb, na = lj_only ? 0 : s->na, nd = lj_only ? 0 : s->nd;
    sg.ptr[0] = eb;                         sg.len[0] = nb;
    sg.ptr[1] = eb + s->nb;                 sg.len[1] = na;
    sg.ptr[2] = eb + s->nb + s->na;         sg.len[2] = nd;
    sg.ptr[3] = eb + s->nb + s->na + s->nd; sg.len[3] = lj_only ? 0 : n14;
    sg.ptr[4] = s->ev.e_blk;                sg.len[4] = lj_only ? s->n_blk : lj_records(s);
    sg.nseg = 5;
    const size_t n_pad = (size_t)s->nsc * 64;
    // wavefronts per supercluster: about one block record each, 4 to 16;
    // one when there are no records to add
    const float *fi = s->ev.fi_blk;  // (null: no records to add)
    const int64_t per_sc = (fi && s->nsc > 0) ? (s->n_blk + s->nsc - 1) / s->nsc : 1;
    const int G = !fi ? 1 : (per_sc > 8 ? 16 : (per_sc > 4 ? 8 : 4));
    const unsigned grid = 5u + (unsigned)((s->nsc + 16 / G - 1) / (16 / G));
    const float fi_scale = span_is_cubic(s) ? s->nbp.fscale_u : 4294967296.0f;
    int *bf = bad_first_word(s);
    // bonded force records (stored by a partner-list evaluation's force kernel)
    const long long *brec = lj_only ? nullptr : s->ev.brec;
    const int2 *srec = s->srec.p;
    const int64_t nrec = s->n_contrib;
    if (ev_stop)  // stamped by the dispatch at the kernel's end
        hipExtLaunchKernelGGL(k_finish, dim3(grid), dim3(1024), 0, s->stream, nullptr, ev_stop, 0, sg, s->e_out.p,
                              (const int *)s->slot_atom.p, s->ev.F, n_pad, fi, (const int *)s->sc_blk_off.p,
                              fi_scale, G, fx, fy, fz, s->ev.disp, s->ev.bad, bf, brec, nrec, srec);
    else
        k_finish<<<grid, 1024, 0, s->stream>>>(sg, s->e_out.p, s->slot_atom.p, s->ev.F, n_pad, fi, s->sc_blk_off.p,
                                                fi_scale, G, fx, fy, fz, s->ev.disp, s->ev.bad, bf, brec, nrec,
                                                srec);
    return hipGetLastError();
}

}  // namespace gm
