// Bond, angle, dihedral and scaled 1-4 terms, energy reduction and output
// conversion, gfx950.
//
// The bonded terms restate force/bonded.rs (paths relative to the reference
// repository root) in f64, in the reference's operation order; this file is
// compiled with -ffp-contract=off so the arithmetic is the CPU path's.  One
// thread per term; each term's forces are converted to int64 fixed point
// (2^-32 kcal/mol/A) and added with integer atomics into the same slot-order
// accumulator the nonbonded kernel uses, so the total is order independent.
// Per-term energies go to an array and are summed by a fixed-order f64 tree.
#include "gm_internal.hpp"

using namespace gm;

namespace {

__device__ inline double min_image(double d, double L) { return d - L * round(d / L); }

__device__ inline double bdiff(double a, double b, int mi, double L) {
    double d = a - b;
    return mi ? min_image(d, L) : d;
}

__device__ inline void fixed_add_d(unsigned long long *p, double v) {
    const long long q = __double2ll_rn(v * 4294967296.0);
    __hip_atomic_fetch_add(p, (unsigned long long)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct Acc {
    unsigned long long *F;
    const int *atom_slot;
    size_t n_pad;
    __device__ void add(uint32_t atom, double fx, double fy, double fz) const {
        const size_t s = (size_t)atom_slot[atom];
        fixed_add_d(&F[s], fx);
        fixed_add_d(&F[n_pad + s], fy);
        fixed_add_d(&F[2 * n_pad + s], fz);
    }
};

struct Geo {
    const double *x, *y, *z;
    double box[3];
    int mi;
};

// force/bonded.rs:5-40
__global__ void k_bonds(int nb, const uint32_t *bi, const uint32_t *bj, const double *bk,
                        const double *br0, Geo g, Acc acc, double *e_out) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nb) return;
    const uint32_t i = bi[n], j = bj[n];
    const double k = bk[n], r0 = br0[n];
    const double dx = bdiff(g.x[j], g.x[i], g.mi, g.box[0]);
    const double dy = bdiff(g.y[j], g.y[i], g.mi, g.box[1]);
    const double dz = bdiff(g.z[j], g.z[i], g.mi, g.box[2]);
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    double e = 0.0;
    if (r != 0.0) {
        const double dv_dr = 2.0 * k * (r - r0);
        e = k * (r - r0) * (r - r0);
        const double ux = dx / r, uy = dy / r, uz = dz / r;
        acc.add(i, dv_dr * ux, dv_dr * uy, dv_dr * uz);
        acc.add(j, -(dv_dr * ux), -(dv_dr * uy), -(dv_dr * uz));
    }
    e_out[n] = e;
}

// force/bonded.rs:44-102
__global__ void k_angles(int na, const uint32_t *ai, const uint32_t *aj, const uint32_t *ak,
                         const double *akth, const double *ath0, Geo g, Acc acc, double *e_out) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= na) return;
    const uint32_t i = ai[n], j = aj[n], kk = ak[n];
    const double k_const = akth[n], theta0 = ath0[n];
    const double b1[3] = {bdiff(g.x[i], g.x[j], g.mi, g.box[0]), bdiff(g.y[i], g.y[j], g.mi, g.box[1]),
                          bdiff(g.z[i], g.z[j], g.mi, g.box[2])};
    const double b2[3] = {bdiff(g.x[kk], g.x[j], g.mi, g.box[0]), bdiff(g.y[kk], g.y[j], g.mi, g.box[1]),
                          bdiff(g.z[kk], g.z[j], g.mi, g.box[2])};
    const double r1 = sqrt(b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]);
    const double r2 = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
    double e = 0.0;
    if (r1 != 0.0 && r2 != 0.0) {
        const double dot = b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2];
        double cos_theta = dot / (r1 * r2);
        cos_theta = cos_theta < -1.0 ? -1.0 : (cos_theta > 1.0 ? 1.0 : cos_theta);
        double sin_theta = sqrt(1.0 - cos_theta * cos_theta);
        sin_theta = sin_theta < 1e-8 ? 1e-8 : sin_theta;
        const double theta = acos(cos_theta);
        const double dv_dtheta = 2.0 * k_const * (theta - theta0);
        e = k_const * (theta - theta0) * (theta - theta0);
        const double a = dv_dtheta / sin_theta;
        double f_i[3], f_k[3];
        for (int c = 0; c < 3; c++) {
            f_i[c] = a * (b2[c] / (r1 * r2) - cos_theta * b1[c] / (r1 * r1));
            f_k[c] = a * (b1[c] / (r1 * r2) - cos_theta * b2[c] / (r2 * r2));
        }
        acc.add(i, f_i[0], f_i[1], f_i[2]);
        acc.add(kk, f_k[0], f_k[1], f_k[2]);
        acc.add(j, -(f_i[0] + f_k[0]), -(f_i[1] + f_k[1]), -(f_i[2] + f_k[2]));
    }
    e_out[n] = e;
}

__device__ inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// force/bonded.rs:112-164 (Blondel-Karplus)
__global__ void k_dihedrals(int nd, const uint32_t *di, const uint32_t *dj, const uint32_t *dk,
                            const uint32_t *dl, const double *dkphi, const uint32_t *dn,
                            const double *ddelta, Geo g, Acc acc, double *e_out) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nd) return;
    const uint32_t i = di[n], j = dj[n], k = dk[n], l = dl[n];
    const double k_phi = dkphi[n], mult = (double)dn[n], delta = ddelta[n];
    const double b1[3] = {bdiff(g.x[j], g.x[i], g.mi, g.box[0]), bdiff(g.y[j], g.y[i], g.mi, g.box[1]),
                          bdiff(g.z[j], g.z[i], g.mi, g.box[2])};
    const double b2[3] = {bdiff(g.x[k], g.x[j], g.mi, g.box[0]), bdiff(g.y[k], g.y[j], g.mi, g.box[1]),
                          bdiff(g.z[k], g.z[j], g.mi, g.box[2])};
    const double b3[3] = {bdiff(g.x[l], g.x[k], g.mi, g.box[0]), bdiff(g.y[l], g.y[k], g.mi, g.box[1]),
                          bdiff(g.z[l], g.z[k], g.mi, g.box[2])};
    double m[3], nv[3], mn[3];
    cross3(b1, b2, m);
    cross3(b2, b3, nv);
    const double m_len2 = dot3(m, m);
    const double n_len2 = dot3(nv, nv);
    const double b2_len = sqrt(dot3(b2, b2));
    double e = 0.0;
    if (m_len2 != 0.0 && n_len2 != 0.0 && b2_len != 0.0) {
        const double xx = dot3(m, nv);
        cross3(m, nv, mn);
        const double yy = dot3(mn, b2) / b2_len;
        const double phi = atan2(yy, xx);
        e = k_phi * (1.0 + cos(mult * phi - delta));
        const double dv_dphi = k_phi * mult * sin(mult * phi - delta);
        const double si = -dv_dphi * b2_len / m_len2;
        const double sl = dv_dphi * b2_len / n_len2;
        const double f_i[3] = {m[0] * si, m[1] * si, m[2] * si};
        const double f_l[3] = {nv[0] * sl, nv[1] * sl, nv[2] * sl};
        const double p = dot3(b1, b2) / (b2_len * b2_len);
        const double q = dot3(b3, b2) / (b2_len * b2_len);
        const double cj = -(p + 1.0), ck = q + 1.0;
        double f_j[3], f_k[3];
        for (int c = 0; c < 3; c++) {
            f_j[c] = f_i[c] * cj + f_l[c] * q;
            f_k[c] = f_i[c] * p - f_l[c] * ck;
        }
        acc.add(i, f_i[0], f_i[1], f_i[2]);
        acc.add(j, f_j[0], f_j[1], f_j[2]);
        acc.add(k, f_k[0], f_k[1], f_k[2]);
        acc.add(l, f_l[0], f_l[1], f_l[2]);
    }
    e_out[n] = e;
}

// Scaled 1-4 LJ: scale * unswitched LJ, no cutoff, min-image displacement,
// sign convention of force/nonbonded.rs:68-78.  No reference counterpart
// (prmtop.rs:132-137 excludes 1-4 pairs entirely).
__global__ void k_lj14(int np, const uint32_t *pi, const uint32_t *pj, const double *sigma,
                       const double *epsilon, double scale, Geo g, Acc acc, double *e_out) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= np) return;
    const uint32_t i = pi[n], j = pj[n];
    const double dx = min_image(g.x[j] - g.x[i], g.box[0]);
    const double dy = min_image(g.y[j] - g.y[i], g.box[1]);
    const double dz = min_image(g.z[j] - g.z[i], g.box[2]);
    const double r2 = dx * dx + dy * dy + dz * dz;
    double e = 0.0;
    const double sig = 0.5 * (sigma[i] + sigma[j]);
    const double eps = sqrt(epsilon[i] * epsilon[j]);
    if (r2 != 0.0 && eps != 0.0) {
        const double r = sqrt(r2);
        const double s = sig / r;
        const double s2 = s * s;
        const double sr6 = s2 * s2 * s2;
        const double sr12 = sr6 * sr6;
        e = scale * (4.0 * eps * (sr12 - sr6));
        const double f_radial = scale * (24.0 * eps / r * (2.0 * sr12 - sr6));
        const double fx = f_radial * dx / r, fy = f_radial * dy / r, fz = f_radial * dz / r;
        acc.add(i, -fx, -fy, -fz);
        acc.add(j, fx, fy, fz);
    }
    e_out[n] = e;
}

// Fixed-order f64 reduction of up to 8 segments: block b sums segment b;
// thread t adds items t, t+1024, ... in order, then a fixed LDS tree.
struct Segs {
    const double *ptr[8];
    long long len[8];
};

__global__ __launch_bounds__(1024) void k_reduce(Segs sg, double *out) {
    __shared__ double sh[1024];
    const int b = blockIdx.x, t = threadIdx.x;
    double a = 0.0;
    for (long long i = t; i < sg.len[b]; i += 1024) a += sg.ptr[b][i];
    sh[t] = a;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) sh[t] = sh[t] + sh[t + w];
        __syncthreads();
    }
    if (t == 0) out[b] = sh[0];
}

__global__ void k_finalize(int n, const int *atom_slot, const unsigned long long *F, size_t n_pad,
                           double *fx, double *fy, double *fz) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const size_t s = (size_t)atom_slot[a];
    fx[a] = (double)(long long)F[s] * kFixedInv;
    fy[a] = (double)(long long)F[n_pad + s] * kFixedInv;
    fz[a] = (double)(long long)F[2 * n_pad + s] * kFixedInv;
}

inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

namespace gm {

hipError_t launch_bonded(gm_sys *s, const double *x, const double *y, const double *z) {
    hipStream_t st = s->stream;
    Geo g{x, y, z, {s->box[0], s->box[1], s->box[2]}, s->bonded_mi};
    Acc acc{s->F.p, s->atom_slot.p, (size_t)s->nsc * 64};
    double *eb = s->e_items.p, *ea = eb + s->nb, *ed = ea + s->na, *e14 = ed + s->nd;
    if (s->nb > 0)
        k_bonds<<<grid_for(s->nb, 128), 128, 0, st>>>((int)s->nb, s->bi.p, s->bj.p, s->bk.p, s->br0.p,
                                                        g, acc, eb);
    if (s->na > 0)
        k_angles<<<grid_for(s->na, 128), 128, 0, st>>>((int)s->na, s->ai.p, s->aj.p, s->ak.p,
                                                         s->akth.p, s->ath0.p, g, acc, ea);
    if (s->nd > 0)
        k_dihedrals<<<grid_for(s->nd, 128), 128, 0, st>>>((int)s->nd, s->di.p, s->dj.p, s->dk.p,
                                                            s->dl.p, s->dkphi.p, s->dn.p,
                                                            s->ddelta.p, g, acc, ed);
    if (s->n14 > 0 && s->scale14 != 0.0)
        k_lj14<<<grid_for(s->n14, 128), 128, 0, st>>>((int)s->n14, s->p14i.p, s->p14j.p, s->sigma.p,
                                                        s->eps.p, s->scale14, g, acc, e14);
    return hipGetLastError();
}

hipError_t launch_finalize(gm_sys *s, double *fx, double *fy, double *fz) {
    hipStream_t st = s->stream;
    Segs sg{};
    const double *eb = s->e_items.p;
    sg.ptr[0] = eb;                        sg.len[0] = s->nb;
    sg.ptr[1] = eb + s->nb;                sg.len[1] = s->na;
    sg.ptr[2] = eb + s->nb + s->na;        sg.len[2] = s->nd;
    sg.ptr[3] = eb + s->nb + s->na + s->nd;
    sg.len[3] = (s->scale14 != 0.0) ? s->n14 : 0;
    sg.ptr[4] = s->e_wi.p;                 sg.len[4] = s->n_wi;
    k_reduce<<<5, 1024, 0, st>>>(sg, s->e_out.p);
    if (s->n > 0)
        k_finalize<<<grid_for(s->n, 256), 256, 0, st>>>((int)s->n, s->atom_slot.p, s->F.p,
                                                          (size_t)s->nsc * 64, fx, fy, fz);
    return hipGetLastError();
}

}  // namespace gm
