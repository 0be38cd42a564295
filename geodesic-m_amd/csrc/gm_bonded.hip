// End of a force evaluation, gfx950: per-term energy reduction, f64 force
// output in atom order, and clearing of the int64 accumulator for the next
// evaluation — one launch.
//
// Energies are summed by a fixed-order f64 tree (the reference sums bonded
// terms, then per-thread LJ partials, in a fixed order: cpu_backend.rs:93-166;
// paths relative to the reference repository root).  Forces leave the
// fixed-point accumulator as value * 2^-32 kcal/mol/A.
#include "gm_internal.hpp"
#include "gm_terms.hpp"

using namespace gm;

namespace {

struct Segs {
    const double *ptr[8];
    long long len[8];
    int nseg;
};

// Fixed-order f64 sum of one segment by a 1024-thread block: thread t adds
// items t, t + 1024, ... into four running sums (items 4k + 0..3 apart by
// 1024, so four loads are in flight), combines them, each wavefront sums its
// lanes with DPP and permlane steps, and thread 0 adds the 16 wave sums in
// order.  The result is valid in thread 0.
__device__ inline double block_sum(const double *ptr, long long len, double *sh) {
    const int t = threadIdx.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    long long i = t;
    for (; i + 3072 < len; i += 4096) {
        a0 += ptr[i];
        a1 += ptr[i + 1024];
        a2 += ptr[i + 2048];
        a3 += ptr[i + 3072];
    }
    for (; i < len; i += 1024) a0 += ptr[i];
    double a = wave_sum_f64((a0 + a1) + (a2 + a3));
    if ((t & 63) == 0) sh[t >> 6] = a;
    __syncthreads();
    double r = 0.0;
    if (t == 0)
        for (int w = 0; w < 16; w++) r += sh[w];
    return r;
}

// Blocks [0, nseg): block b sums segment b.  Blocks >= nseg: one thread per
// slot (reads of the accumulator are contiguous; the f64 force goes to the
// slot's atom).  Block 0 also clears the displacement bound the prep kernel raised.
__global__ __launch_bounds__(1024) void k_finish(Segs sg, double *eout, const int *slot_atom,
                                                 unsigned long long *F, size_t n_pad, double *fx,
                                                 double *fy, double *fz, unsigned *disp) {
    __shared__ double sh[16];
    const int b = blockIdx.x, t = threadIdx.x;
    if (b == 0 && t < kDispSlots) disp[t * kDispStride] = 0u;
    if (b < sg.nseg) {
        const double v = block_sum(sg.ptr[b], sg.len[b], sh);
        if (t == 0) eout[b] = v;
        return;
    }
    const size_t s = (size_t)(b - sg.nseg) * 1024 + t;
    if (s >= n_pad) return;
    const int a = slot_atom[s];
    const long long qx = (long long)F[s], qy = (long long)F[n_pad + s], qz = (long long)F[2 * n_pad + s];
    F[s] = 0ull;
    F[n_pad + s] = 0ull;
    F[2 * n_pad + s] = 0ull;
    if (a < 0) return;
    fx[a] = (double)qx * kFixedInv;
    fy[a] = (double)qy * kFixedInv;
    fz[a] = (double)qz * kFixedInv;
}

// Batched end of evaluation: blocks [0, 5 R) sum replica b/5's segment b%5
// (bond, angle, dihedral, 1-4, LJ) exactly like k_finish; the remaining
// blocks convert every atom's force of every replica and clear the
// accumulators.
__global__ __launch_bounds__(1024) void k_batch_finish(const ReplicaDesc *__restrict__ reps, int n_reps,
                                                       const int *__restrict__ atom_rep,
                                                       const int64_t *__restrict__ atom_off,
                                                       int64_t n_atoms, double *eout, double *fx,
                                                       double *fy, double *fz) {
    __shared__ double sh[1024];
    const int b = blockIdx.x, t = threadIdx.x;
    if (b < 5 * n_reps) {
        const ReplicaDesc &R = reps[b / 5];
        const int seg = b % 5;
        if (seg == 0 && t < kDispSlots) R.prep.disp[t * kDispStride] = 0u;  // bound of the next evaluation
        const double *eb = R.terms.e_items;
        const double *ptr;
        long long len;
        switch (seg) {
            case 0: ptr = eb; len = R.terms.nb; break;
            case 1: ptr = eb + R.terms.nb; len = R.terms.na; break;
            case 2: ptr = eb + R.terms.nb + R.terms.na; len = R.terms.nd; break;
            case 3: ptr = eb + R.terms.nb + R.terms.na + R.terms.nd; len = R.terms.n14; break;
            default: ptr = R.tile.e_wi; len = R.n_wi; break;
        }
        const double v = block_sum(ptr, len, sh);
        if (t == 0) eout[b] = v;
        return;
    }
    const int64_t a = (int64_t)(b - 5 * n_reps) * 1024 + t;
    if (a >= n_atoms) return;
    const int r = atom_rep[a];
    const ReplicaDesc &R = reps[r];
    unsigned long long *F = R.terms.F;
    const size_t n_pad = R.terms.n_pad;
    const size_t s = (size_t)R.terms.atom_slot[a - atom_off[r]];
    fx[a] = (double)(long long)F[s] * kFixedInv;
    fy[a] = (double)(long long)F[n_pad + s] * kFixedInv;
    fz[a] = (double)(long long)F[2 * n_pad + s] * kFixedInv;
    F[s] = 0ull;
    F[n_pad + s] = 0ull;
    F[2 * n_pad + s] = 0ull;
}

}  // namespace

namespace gm {

hipError_t launch_batch_finish(const ReplicaDesc *reps, int n_reps, const int *atom_rep,
                               const int64_t *atom_off, int64_t n_atoms, double *eout,
                               double *fx, double *fy, double *fz, hipStream_t st) {
    const unsigned grid = 5u * (unsigned)n_reps + (unsigned)((n_atoms + 1023) / 1024);
    k_batch_finish<<<grid, 1024, 0, st>>>(reps, n_reps, atom_rep, atom_off, n_atoms, eout, fx, fy, fz);
    return hipGetLastError();
}

hipError_t launch_finish(gm_sys *s, double *fx, double *fy, double *fz, bool lj_only) {
    Segs sg{};
    const double *eb = s->e_items.p;
    const int64_t n14 = (s->scale14 != 0.0) ? s->n14 : 0;
    const int64_t nb = lj_only ? 0 : s->nb, na = lj_only ? 0 : s->na, nd = lj_only ? 0 : s->nd;
    sg.ptr[0] = eb;                         sg.len[0] = nb;
    sg.ptr[1] = eb + s->nb;                 sg.len[1] = na;
    sg.ptr[2] = eb + s->nb + s->na;         sg.len[2] = nd;
    sg.ptr[3] = eb + s->nb + s->na + s->nd; sg.len[3] = lj_only ? 0 : n14;
    sg.ptr[4] = s->e_wi.p;                  sg.len[4] = s->n_wi;
    sg.nseg = 5;
    const size_t n_pad = (size_t)s->nsc * 64;
    const unsigned grid = 5u + (unsigned)((n_pad + 1023) / 1024);
    k_finish<<<grid, 1024, 0, s->stream>>>(sg, s->e_out.p, s->slot_atom.p, s->F.p, n_pad, fx, fy, fz,
                                            s->disp.p);
    return hipGetLastError();
}

}  // namespace gm
