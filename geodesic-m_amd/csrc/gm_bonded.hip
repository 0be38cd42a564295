// End of a force evaluation, gfx950: per-term energy reduction, f64 force
// output in atom order, and clearing of the int64 accumulator for the next
// evaluation — one launch.
//
// Energies are summed by a fixed-order f64 tree (the reference sums bonded
// terms, then per-thread LJ partials, in a fixed order: cpu_backend.rs:93-166;
// paths relative to the reference repository root).  Forces leave the
// fixed-point accumulator as value * 2^-32 kcal/mol/A.
#include "gm_internal.hpp"

using namespace gm;

namespace {

struct Segs {
    const double *ptr[8];
    long long len[8];
    int nseg;
};

// Blocks [0, nseg): block b sums segment b (thread t adds items t, t+1024, ...
// in order, then a fixed LDS tree).  Blocks >= nseg: one thread per atom.
__global__ __launch_bounds__(1024) void k_finish(Segs sg, double *eout, int n, const int *atom_slot,
                                                 unsigned long long *F, size_t n_pad, double *fx,
                                                 double *fy, double *fz) {
    __shared__ double sh[1024];
    const int b = blockIdx.x, t = threadIdx.x;
    if (b < sg.nseg) {
        double a = 0.0;
        for (long long i = t; i < sg.len[b]; i += 1024) a += sg.ptr[b][i];
        sh[t] = a;
        __syncthreads();
        for (int w = 512; w > 0; w >>= 1) {
            if (t < w) sh[t] = sh[t] + sh[t + w];
            __syncthreads();
        }
        if (t == 0) eout[b] = sh[0];
        return;
    }
    const int a = (b - sg.nseg) * 1024 + t;
    if (a >= n) return;
    const size_t s = (size_t)atom_slot[a];
    fx[a] = (double)(long long)F[s] * kFixedInv;
    fy[a] = (double)(long long)F[n_pad + s] * kFixedInv;
    fz[a] = (double)(long long)F[2 * n_pad + s] * kFixedInv;
    F[s] = 0ull;
    F[n_pad + s] = 0ull;
    F[2 * n_pad + s] = 0ull;
}

}  // namespace

namespace gm {

hipError_t launch_finish(gm_sys *s, double *fx, double *fy, double *fz) {
    Segs sg{};
    const double *eb = s->e_items.p;
    const int64_t n14 = (s->scale14 != 0.0) ? s->n14 : 0;
    sg.ptr[0] = eb;                         sg.len[0] = s->nb;
    sg.ptr[1] = eb + s->nb;                 sg.len[1] = s->na;
    sg.ptr[2] = eb + s->nb + s->na;         sg.len[2] = s->nd;
    sg.ptr[3] = eb + s->nb + s->na + s->nd; sg.len[3] = n14;
    sg.ptr[4] = s->e_wi.p;                  sg.len[4] = s->n_wi;
    sg.nseg = 5;
    const unsigned grid = 5u + (unsigned)((s->n + 1023) / 1024);
    k_finish<<<grid, 1024, 0, s->stream>>>(sg, s->e_out.p, (int)s->n, s->atom_slot.p, s->F.p,
                                            (size_t)s->nsc * 64, fx, fy, fz);
    return hipGetLastError();
}

}  // namespace gm
