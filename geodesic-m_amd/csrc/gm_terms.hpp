// Device functions for one bonded term each: bond, angle, dihedral, scaled
// 1-4 LJ.  f64 restatement of force/bonded.rs (paths relative to the reference
// repository root) in the reference's operation order; every function body
// disables FMA contraction so the arithmetic is the CPU path's whatever flags
// the including file uses.  Forces are rounded to int64 fixed point (2^-32
// kcal/mol/A) per contribution and either added to the accumulator (slot
// order) with integer atomics, or stored as the term's own records, which the
// finish kernel adds per atom (the same integers, summed without atomics).
#pragma once

#include "gm_internal.hpp"

namespace gm {

__device__ inline double term_min_image(double d, double L) {
#pragma clang fp contract(off)
    return d - L * round(d / L);
}

__device__ inline double term_diff(double a, double b, int mi, double L) {
#pragma clang fp contract(off)
    double d = a - b;
    return mi ? term_min_image(d, L) : d;
}

// v (kcal/mol/A) into the accumulator word p, in units of 2^-32.  A value
// that is not finite or whose fixed-point image reaches 2^62 is not added: it
// marks the slot (*bad = 1), and the finish kernel reports that atom's force
// as NaN.
__device__ inline void fixed_add_d(unsigned long long *p, double v, unsigned *bad) {
    const double t = v * 4294967296.0;
    if (!(fabs(t) < kFixedLimit)) {
        *bad = 1u;
        return;
    }
    const long long q = __double2ll_rn(t);
    __hip_atomic_fetch_add(p, (unsigned long long)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The same rounding into a record: 0 (and the slot marked) for a value that
// is not finite or out of range.
__device__ inline void fixed_put_d(long long *p, double v, unsigned *bad) {
    const double t = v * 4294967296.0;
    if (!(fabs(t) < kFixedLimit)) {
        *bad = 1u;
        *p = 0;
        return;
    }
    *p = __double2ll_rn(t);
}

// Everything a bonded term reads: positions (atom order, f64), topology,
// the slot map and the accumulator.
struct TermArgs {
    const double *x, *y, *z;
    double box[3];
    int mi;  // min-image bonded displacements
    unsigned long long *F;
    unsigned *bad;  // per slot: a contribution was not finite or out of the fixed-point range
    const int *atom_slot;
    size_t n_pad;
    const uint32_t *bi, *bj;
    const double *bk, *br0;
    const uint32_t *ai, *aj, *ak;
    const double *akth, *ath0;
    const uint32_t *di, *dj, *dk, *dl, *dn;
    const double *dkphi;
    const double2 *dphase;  // per dihedral: cos(delta), sin(delta)
    const uint32_t *p14i, *p14j;
    const double *sigma, *epsilon;
    double scale14;
    int64_t nb, na, nd, n14;  // n14 = 0 when scale14 == 0
    double *e_items;           // nb + na + nd + n14 energies
    long long *rec;            // non-null: force records, [x, y, z][n_rec], not F: contribution c
                               // (the term's first plus the atom's place in the term) goes to
                               // record rec_pos[c]
    const int *rec_pos;
    int64_t n_rec;
};

// Record numbers: bonds 2 each, then angles 3, dihedrals 4, 1-4 pairs 2, in
// term order; within a term, its atoms in the order the term adds them.
__device__ inline int64_t rec_angle0(const TermArgs &t) { return 2 * t.nb; }
__device__ inline int64_t rec_dihedral0(const TermArgs &t) { return 2 * t.nb + 3 * t.na; }
__device__ inline int64_t rec_lj14_0(const TermArgs &t) { return 2 * t.nb + 3 * t.na + 4 * t.nd; }

// Force (fx, fy, fz) on atom from the term's record c.
__device__ inline void term_add(const TermArgs &t, int64_t c, uint32_t atom, double fx, double fy, double fz) {
    const size_t s = (size_t)t.atom_slot[atom];
    if (t.rec) {
        const int64_t r = t.rec_pos[c];
        fixed_put_d(&t.rec[r], fx, &t.bad[s]);
        fixed_put_d(&t.rec[t.n_rec + r], fy, &t.bad[s]);
        fixed_put_d(&t.rec[2 * t.n_rec + r], fz, &t.bad[s]);
        return;
    }
    fixed_add_d(&t.F[s], fx, &t.bad[s]);
    fixed_add_d(&t.F[t.n_pad + s], fy, &t.bad[s]);
    fixed_add_d(&t.F[2 * t.n_pad + s], fz, &t.bad[s]);
}

// A term that adds no force (degenerate geometry): its records hold zeros.
__device__ inline void term_none(const TermArgs &t, int64_t c, int k) {
    if (!t.rec) return;
    for (int a = 0; a < k; a++) {
        const int64_t r = t.rec_pos[c + a];
        t.rec[r] = t.rec[t.n_rec + r] = t.rec[2 * t.n_rec + r] = 0;
    }
}

// force/bonded.rs:5-40
__device__ inline double term_bond(const TermArgs &t, int64_t n) {
#pragma clang fp contract(off)
    const uint32_t i = t.bi[n], j = t.bj[n];
    const double k = t.bk[n], r0 = t.br0[n];
    const double dx = term_diff(t.x[j], t.x[i], t.mi, t.box[0]);
    const double dy = term_diff(t.y[j], t.y[i], t.mi, t.box[1]);
    const double dz = term_diff(t.z[j], t.z[i], t.mi, t.box[2]);
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const int64_t c = 2 * n;
    if (r == 0.0) {
        term_none(t, c, 2);
        return 0.0;
    }
    const double dv_dr = 2.0 * k * (r - r0);
    const double e = k * (r - r0) * (r - r0);
    const double ux = dx / r, uy = dy / r, uz = dz / r;
    term_add(t, c, i, dv_dr * ux, dv_dr * uy, dv_dr * uz);
    term_add(t, c + 1, j, -(dv_dr * ux), -(dv_dr * uy), -(dv_dr * uz));
    return e;
}

// force/bonded.rs:44-102
__device__ inline double term_angle(const TermArgs &t, int64_t n) {
#pragma clang fp contract(off)
    const uint32_t i = t.ai[n], j = t.aj[n], kk = t.ak[n];
    const double k_const = t.akth[n], theta0 = t.ath0[n];
    const double b1[3] = {term_diff(t.x[i], t.x[j], t.mi, t.box[0]), term_diff(t.y[i], t.y[j], t.mi, t.box[1]),
                          term_diff(t.z[i], t.z[j], t.mi, t.box[2])};
    const double b2[3] = {term_diff(t.x[kk], t.x[j], t.mi, t.box[0]), term_diff(t.y[kk], t.y[j], t.mi, t.box[1]),
                          term_diff(t.z[kk], t.z[j], t.mi, t.box[2])};
    const double r1 = sqrt(b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]);
    const double r2 = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
    const int64_t c = rec_angle0(t) + 3 * n;
    if (r1 == 0.0 || r2 == 0.0) {
        term_none(t, c, 3);
        return 0.0;
    }
    const double dot = b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2];
    double cos_theta = dot / (r1 * r2);
    cos_theta = cos_theta < -1.0 ? -1.0 : (cos_theta > 1.0 ? 1.0 : cos_theta);
    double sin_theta = sqrt(1.0 - cos_theta * cos_theta);
    sin_theta = sin_theta < 1e-8 ? 1e-8 : sin_theta;
    const double theta = acos(cos_theta);
    const double dv_dtheta = 2.0 * k_const * (theta - theta0);
    const double e = k_const * (theta - theta0) * (theta - theta0);
    const double a = dv_dtheta / sin_theta;
    double f_i[3], f_k[3];
    for (int c = 0; c < 3; c++) {
        f_i[c] = a * (b2[c] / (r1 * r2) - cos_theta * b1[c] / (r1 * r1));
        f_k[c] = a * (b1[c] / (r1 * r2) - cos_theta * b2[c] / (r2 * r2));
    }
    term_add(t, c, i, f_i[0], f_i[1], f_i[2]);
    term_add(t, c + 1, kk, f_k[0], f_k[1], f_k[2]);
    term_add(t, c + 2, j, -(f_i[0] + f_k[0]), -(f_i[1] + f_k[1]), -(f_i[2] + f_k[2]));
    return e;
}

__device__ inline void term_cross(const double *a, const double *b, double *o) {
#pragma clang fp contract(off)
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double term_dot(const double *a, const double *b) {
#pragma clang fp contract(off)
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// force/bonded.rs:112-164 (Blondel-Karplus).  The reference takes
// phi = atan2(y, x) and then cos / sin of n phi - delta; here cos phi and
// sin phi are (x, y) / |(x, y)|, cos n phi + i sin n phi their n-th complex
// power (binary powering), and the phase enters by the angle-difference
// formulas with cos delta and sin delta computed once per term at system
// creation: the same values to f64 rounding, without the three
// transcendental calls that made this term the longest dependent chain of
// an evaluation.
__device__ inline double term_dihedral(const TermArgs &t, int64_t n) {
#pragma clang fp contract(off)
    const uint32_t i = t.di[n], j = t.dj[n], k = t.dk[n], l = t.dl[n];
    const uint32_t nmult = t.dn[n];
    const double k_phi = t.dkphi[n], mult = (double)nmult;
    const double2 phase = t.dphase[n];
    const double b1[3] = {term_diff(t.x[j], t.x[i], t.mi, t.box[0]), term_diff(t.y[j], t.y[i], t.mi, t.box[1]),
                          term_diff(t.z[j], t.z[i], t.mi, t.box[2])};
    const double b2[3] = {term_diff(t.x[k], t.x[j], t.mi, t.box[0]), term_diff(t.y[k], t.y[j], t.mi, t.box[1]),
                          term_diff(t.z[k], t.z[j], t.mi, t.box[2])};
    const double b3[3] = {term_diff(t.x[l], t.x[k], t.mi, t.box[0]), term_diff(t.y[l], t.y[k], t.mi, t.box[1]),
                          term_diff(t.z[l], t.z[k], t.mi, t.box[2])};
    double m[3], nv[3], mn[3];
    term_cross(b1, b2, m);
    term_cross(b2, b3, nv);
    const double m_len2 = term_dot(m, m);
    const double n_len2 = term_dot(nv, nv);
    const double b2_len = sqrt(term_dot(b2, b2));
    const int64_t c = rec_dihedral0(t) + 4 * n;
    if (m_len2 == 0.0 || n_len2 == 0.0 || b2_len == 0.0) {
        term_none(t, c, 4);
        return 0.0;
    }
    const double xx = term_dot(m, nv);
    term_cross(m, nv, mn);
    const double yy = term_dot(mn, b2) / b2_len;
    const double rinv = 1.0 / sqrt(xx * xx + yy * yy);  // (x, y) != 0: m and n are not 0
    double pc = xx * rinv, ps = yy * rinv;               // cos phi, sin phi
    double cn = 1.0, sn = 0.0;                           // cos n phi, sin n phi
    for (uint32_t b = nmult; b != 0u; b >>= 1) {
        if (b & 1u) {
            const double c = cn * pc - sn * ps;
            sn = cn * ps + sn * pc;
            cn = c;
        }
        const double c = pc * pc - ps * ps;
        ps = 2.0 * pc * ps;
        pc = c;
    }
    const double cos_v = cn * phase.x + sn * phase.y;  // cos(n phi - delta)
    const double sin_v = sn * phase.x - cn * phase.y;  // sin(n phi - delta)
    const double e = k_phi * (1.0 + cos_v);
    const double dv_dphi = k_phi * mult * sin_v;
    const double si = -dv_dphi * b2_len / m_len2;
    const double sl = dv_dphi * b2_len / n_len2;
    const double f_i[3] = {m[0] * si, m[1] * si, m[2] * si};
    const double f_l[3] = {nv[0] * sl, nv[1] * sl, nv[2] * sl};
    const double p = term_dot(b1, b2) / (b2_len * b2_len);
    const double q = term_dot(b3, b2) / (b2_len * b2_len);
    const double cj = -(p + 1.0), ck = q + 1.0;
    double f_j[3], f_k[3];
    for (int c = 0; c < 3; c++) {
        f_j[c] = f_i[c] * cj + f_l[c] * q;
        f_k[c] = f_i[c] * p - f_l[c] * ck;
    }
    term_add(t, c, i, f_i[0], f_i[1], f_i[2]);
    term_add(t, c + 1, j, f_j[0], f_j[1], f_j[2]);
    term_add(t, c + 2, k, f_k[0], f_k[1], f_k[2]);
    term_add(t, c + 3, l, f_l[0], f_l[1], f_l[2]);
    return e;
}

// Scaled 1-4 LJ: scale * unswitched LJ, no cutoff, min-image displacement,
// sign convention of force/nonbonded.rs:68-78.  No reference counterpart
// (prmtop.rs:132-137 excludes 1-4 pairs entirely).
__device__ inline double term_lj14(const TermArgs &t, int64_t n) {
#pragma clang fp contract(off)
    const uint32_t i = t.p14i[n], j = t.p14j[n];
    const double dx = term_min_image(t.x[j] - t.x[i], t.box[0]);
    const double dy = term_min_image(t.y[j] - t.y[i], t.box[1]);
    const double dz = term_min_image(t.z[j] - t.z[i], t.box[2]);
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double sig = 0.5 * (t.sigma[i] + t.sigma[j]);
    const double eps = sqrt(t.epsilon[i] * t.epsilon[j]);
    const int64_t c = rec_lj14_0(t) + 2 * n;
    if (r2 == 0.0 || eps == 0.0) {
        term_none(t, c, 2);
        return 0.0;
    }
    const double r = sqrt(r2);
    const double s = sig / r;
    const double s2 = s * s;
    const double sr6 = s2 * s2 * s2;
    const double sr12 = sr6 * sr6;
    const double e = t.scale14 * (4.0 * eps * (sr12 - sr6));
    const double f_radial = t.scale14 * (24.0 * eps / r * (2.0 * sr12 - sr6));
    const double fx = f_radial * dx / r, fy = f_radial * dy / r, fz = f_radial * dz / r;
    term_add(t, c, i, -fx, -fy, -fz);
    term_add(t, c + 1, j, fx, fy, fz);
    return e;
}

// Everything one nonbonded work item reads (one system or one replica).
struct TileArgs {
    const int4 *qpos;                  // per slot: fixed-point x, y, z and the bits of sigma/2
                                       // (kernel length unit), written by the prep kernel
    const float *qeps;                 // per slot: 2 sqrt(eps)
    const int4 *wi;  // work items: (supercluster, first entry, end entry, j-cluster of the first entry)
    const int2 *ent_meta;              // (j cluster, i-cluster mask)
    const uint2 *ent_dmin;             // per entry, one byte per tile: closest listed pair at build
    const uint8_t *ent_lb;             // per-lane mask bytes per entry
    const unsigned *disp;              // largest displacement since the build (f32 bits, A),
                                       // kDispSlots words; null: no tile pruning
    float prune_scale, prune_bias;     // keep a tile iff byte <= disp * scale + bias
    unsigned long long *F;             // int64 fixed-point accumulator, slot order (j forces, bonded)
    unsigned *bad;                     // per slot: force not finite this evaluation
    int *bad_first;                    // first atom whose force is not finite (atomic minimum)
    int n_pad;                         // slots
    double *e_blk;                     // per block LJ energy (blocks of kBlock entries)
    // i forces per block, [block][x, y, z][64 slots of its supercluster], f32
    // in the kernel's force unit: plain stores.  A supercluster's blocks are
    // [sc_blk_off[sc], sc_blk_off[sc + 1]); the finish kernel rounds each
    // record to fixed point (scaled by fi_scale) and adds them per slot.
    // Null: each block's i forces are rounded to fixed point in the force
    // kernel and go to the accumulator F with atomics instead (same integers).
    float *fi_blk;
    const int *sc_blk_off;
    float fi_scale;
    const int *slot_atom;              // slot -> atom (-1 for padding), for the finish kernels
    NbParams p;
    int cubic;
};

// Everything the per-evaluation prep kernel reads and writes (one system or replica).
struct PrepArgs {
    const double *x, *y, *z;           // positions, atom order
    const int *atom_slot;
    const uint32_t *qref;              // fixed-point positions at the list build, 3 x n, atom order
    int4 *qpos;                        // x, y, z and w = sigma/2 (NaN for a non-finite position)
    const float2 *atom_par;            // per atom (sigma/2, sqrt eps), A
    float sig_scale;                   // sigma/2 in the force kernel's length unit: A * sig_scale
    const int *wtype;                  // non-null: w = the atom's type-table offset (bit 31 for a
                                       // non-finite position) instead of sigma/2
    unsigned *disp;
    unsigned *bad;                     // per slot: set for an atom whose position is not finite
    int *bad_first;                    // reset to INT_MAX at the start of every evaluation
    double iq[3];                      // 2^32 / box
    float sq_len[3];                   // box / 2^32, A
    int n;
};

// Everything one partner-list group reads (one system or one replica).
struct PlArgs {
    const int4 *qpos;    // per slot: fixed-point x, y, z; w = the atom's type-table offset
    const float2 *type_par;  // per type: sigma/2 (A), 2 sqrt(eps)
    int n_types;
    float sig_scale;     // sigma/2 in the kernel's length unit: A * sig_scale
    const int *goff;     // per group: list offset
    const int2 *glen;    // per group: padded first-part length, whole padded length
    const int *nbr;      // partner slots
    const unsigned *disp;  // displacement words (null: no pruning)
    float dmax_first;      // evaluate the first part only while the displacement is at most this
    unsigned long long *F;  // per slot: its LJ force in fixed point (plain stores)
    unsigned *bad;
    double *e_grp;       // per group: half its LJ energy sum
    int n_pad;
    int n_grp;
    int pt_shift;        // LJ parameters by type pair, 2^pt_shift types a side (-1: by type)
    NbParams p;
};

// One replica of a batch: its arguments for either layout (the batch
// evaluates every replica in one), the bonded arguments and the item split.
struct ReplicaDesc {
    TileArgs tile;
    PlArgs pl;                         // partner-list batches
    const int2 *srec;                  // partner-list batches: per slot, its atom's bonded records
    TermArgs terms;
    PrepArgs prep;
    int n_bonded_groups;               // items [0, n) are bonded groups, then work items
    int n_wi;
    int n_blk;                         // LJ energy records: fixed-point blocks, or partner-list groups
};

// Term k of the flat range [bonds | angles | dihedrals | 1-4]; writes its energy.
__device__ __attribute__((always_inline)) inline void term_dispatch(const TermArgs &t, int64_t k) {
    double e;
    int64_t n = k;
    if (n < t.nb) {
        e = term_bond(t, n);
    } else if ((n -= t.nb) < t.na) {
        e = term_angle(t, n);
    } else if ((n -= t.na) < t.nd) {
        e = term_dihedral(t, n);
    } else {
        n -= t.nd;
        e = term_lj14(t, n);
    }
    t.e_items[k] = e;
}

}  // namespace gm
