/*
 * gm_oracle.c -- CPU f64 restatement of GEODESIC-M's force-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libgeodesic_hip.so) never links it and has no CPU fallback.
 *
 * Every function restates the reference's arithmetic in the reference's
 * operation order (file:line cited per function; paths relative to the
 * reference repository root).  Compile with -ffp-contract=off: rustc never
 * contracts a*b+c into an FMA, so neither may this file.
 *
 * Pinning (what checks each function, beyond retyping its formula):
 *   - or_pair_forces, bonded terms, cpu backend, integrator, constraints:
 *     the reference's golden 100-frame DCD of ala_dipeptide
 *     (geodesic/tests/golden/ala_dipeptide_ref.dcd), closed-form LJ values,
 *     finite-difference gradients, Newton's third law.
 *   - neighbour builds: the reference's neighbor_list.rs cases; the cell
 *     variant is checked bit-exact against the brute-force variant.
 *   - or_lj14_forces (no reference counterpart): closed form 0.5*V_LJ(r) and
 *     finite differences.
 *   - min-image bonded terms (no reference counterpart): equal to the raw
 *     terms for unwrapped molecules, translation invariance across the box.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* neighbor.rs:74-76 and force/nonbonded.rs:172-174.  C round() is half away
 * from zero, the same as Rust f64::round. */
static inline double min_image(double d, double box_len) {
    return d - box_len * round(d / box_len);
}

/* neighbor.rs:86-88 */
static inline double wrap_coord(double x, double box_len) {
    return x - box_len * floor(x / box_len);
}

/* neighbor.rs:78-84: wrap every coordinate into [0, box). */
void or_wrap_positions(int64_t n, double *x, double *y, double *z, const double *box) {
    for (int64_t i = 0; i < n; i++) {
        x[i] = wrap_coord(x[i], box[0]);
        y[i] = wrap_coord(y[i], box[1]);
        z[i] = wrap_coord(z[i], box[2]);
    }
}

/* ---- exclusion set (neighbor.rs:18-23: HashSet of (min,max)) ------------- */

typedef struct {
    int64_t *off; /* n+1 */
    uint32_t *nbr; /* partners > i, sorted ascending, unique */
} excl_csr;

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

static int excl_build(int64_t n, int64_t ne, const uint32_t *ei, const uint32_t *ej, excl_csr *out) {
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(ne > 0 ? ne : 1));
    out->off = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    out->nbr = NULL;
    if (!keys || !out->off) { free(keys); return -1; }
    int64_t m = 0;
    for (int64_t k = 0; k < ne; k++) {
        uint32_t a = ei[k], b = ej[k];
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        if (lo == hi || hi >= (uint64_t)n) continue; /* a self-pair never matches i<j */
        keys[m++] = ((uint64_t)lo << 32) | hi;
    }
    qsort(keys, (size_t)m, sizeof(uint64_t), cmp_u64);
    int64_t u = 0;
    for (int64_t k = 0; k < m; k++)
        if (u == 0 || keys[k] != keys[u - 1]) keys[u++] = keys[k];
    out->nbr = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(u > 0 ? u : 1));
    if (!out->nbr) { free(keys); return -1; }
    for (int64_t k = 0; k < u; k++) out->off[(keys[k] >> 32) + 1]++;
    for (int64_t i = 0; i < n; i++) out->off[i + 1] += out->off[i];
    for (int64_t k = 0; k < u; k++) out->nbr[k] = (uint32_t)(keys[k] & 0xffffffffu);
    free(keys);
    return 0;
}

static void excl_free(excl_csr *e) { free(e->off); free(e->nbr); }

static inline int excl_has(const excl_csr *e, uint32_t i, uint32_t j) {
    int64_t lo = e->off[i], hi = e->off[i + 1];
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (e->nbr[mid] < j) lo = mid + 1; else hi = mid;
    }
    return lo < e->off[i + 1] && e->nbr[lo] == j;
}

/* neighbor.rs:25-43: serial O(N^2) half list, i-major, exclusion filter,
 * min-image distance, include iff r2 <= r_skin^2.  Positions are used as
 * given (the caller wraps first when periodic, neighbor.rs:13-15).  Returns
 * the pair count; writes at most cap pairs.  Returns -1 on allocation failure. */
int64_t or_neighbor_build_brute(int64_t n, const double *x, const double *y, const double *z,
                                const double *box, double r_skin, int64_t ne, const uint32_t *ei,
                                const uint32_t *ej, uint32_t *pi, uint32_t *pj, int64_t cap) {
    excl_csr ex;
    if (excl_build(n, ne, ei, ej, &ex)) { excl_free(&ex); return -1; }
    const double r_skin_sq = r_skin * r_skin;
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t p = ex.off[i], pe = ex.off[i + 1];
        for (int64_t j = i + 1; j < n; j++) {
            while (p < pe && ex.nbr[p] < (uint32_t)j) p++;
            if (p < pe && ex.nbr[p] == (uint32_t)j) continue;
            double dx = min_image(x[j] - x[i], box[0]);
            double dy = min_image(y[j] - y[i], box[1]);
            double dz = min_image(z[j] - z[i], box[2]);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= r_skin_sq) {
                if (cnt < cap) { pi[cnt] = (uint32_t)i; pj[cnt] = (uint32_t)j; }
                cnt++;
            }
        }
    }
    excl_free(&ex);
    return cnt;
}

/* Same pair set and order as or_neighbor_build_brute (identical f64 decision
 * per pair), found through a cell grid so 250k-atom systems finish in
 * seconds.  Cells are at least r_skin wide; with fewer than 3 cells along an
 * axis every cell along that axis is visited exactly once, so small boxes are
 * neither missed nor double counted. */
static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

int64_t or_neighbor_build_cells(int64_t n, const double *x, const double *y, const double *z,
                                const double *box, double r_skin, int64_t ne, const uint32_t *ei,
                                const uint32_t *ej, uint32_t *pi, uint32_t *pj, int64_t cap) {
    excl_csr ex;
    if (excl_build(n, ne, ei, ej, &ex)) { excl_free(&ex); return -1; }
    int nc[3];
    double cs[3];
    for (int a = 0; a < 3; a++) {
        double f = floor(box[a] / (r_skin * (1.0 + 1e-9)));
        nc[a] = f < 1.0 ? 1 : (f > 1024.0 ? 1024 : (int)f);
        cs[a] = box[a] / nc[a];
    }
    int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    int64_t *cstart = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
    int64_t *cell_of = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    uint32_t *sorted = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n > 0 ? n : 1));
    if (!cstart || !cell_of || !sorted || !cand) {
        free(cstart); free(cell_of); free(sorted); free(cand); excl_free(&ex); return -1;
    }
    const double *pos[3] = {x, y, z};
    for (int64_t i = 0; i < n; i++) {
        int c[3];
        for (int a = 0; a < 3; a++) {
            double w = wrap_coord(pos[a][i], box[a]);
            int64_t k = (int64_t)floor(w / cs[a]);
            if (k >= nc[a]) k = (k == nc[a]) ? 0 : nc[a] - 1;
            if (k < 0) k = 0;
            c[a] = (int)k;
        }
        cell_of[i] = ((int64_t)c[0] * nc[1] + c[1]) * nc[2] + c[2];
        cstart[cell_of[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncell > 0 ? ncell : 1));
    memcpy(fill, cstart, sizeof(int64_t) * (size_t)ncell);
    for (int64_t i = 0; i < n; i++) sorted[fill[cell_of[i]]++] = (uint32_t)i;
    free(fill);

    int offs[3][3], noff[3];
    for (int a = 0; a < 3; a++) {
        if (nc[a] >= 3) { noff[a] = 3; offs[a][0] = -1; offs[a][1] = 0; offs[a][2] = 1; }
        else { noff[a] = nc[a]; for (int k = 0; k < nc[a]; k++) offs[a][k] = k; }
    }
    const double r_skin_sq = r_skin * r_skin;
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t ci = cell_of[i];
        int c0 = (int)(ci / ((int64_t)nc[1] * nc[2])), c1 = (int)((ci / nc[2]) % nc[1]), c2 = (int)(ci % nc[2]);
        int64_t m = 0;
        for (int a = 0; a < noff[0]; a++)
            for (int b = 0; b < noff[1]; b++)
                for (int d = 0; d < noff[2]; d++) {
                    int e0 = nc[0] >= 3 ? (c0 + offs[0][a] + nc[0]) % nc[0] : offs[0][a];
                    int e1 = nc[1] >= 3 ? (c1 + offs[1][b] + nc[1]) % nc[1] : offs[1][b];
                    int e2 = nc[2] >= 3 ? (c2 + offs[2][d] + nc[2]) % nc[2] : offs[2][d];
                    int64_t cj = ((int64_t)e0 * nc[1] + e1) * nc[2] + e2;
                    for (int64_t s = cstart[cj]; s < cstart[cj + 1]; s++) {
                        uint32_t j = sorted[s];
                        if ((int64_t)j <= i) continue;
                        if (excl_has(&ex, (uint32_t)i, j)) continue;
                        double dx = min_image(x[j] - x[i], box[0]);
                        double dy = min_image(y[j] - y[i], box[1]);
                        double dz = min_image(z[j] - z[i], box[2]);
                        double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 <= r_skin_sq) cand[m++] = j;
                    }
                }
        qsort(cand, (size_t)m, sizeof(uint32_t), cmp_u32);
        for (int64_t k = 0; k < m; k++) {
            if (cnt < cap) { pi[cnt] = (uint32_t)i; pj[cnt] = cand[k]; }
            cnt++;
        }
    }
    free(cstart); free(cell_of); free(sorted); free(cand); excl_free(&ex);
    return cnt;
}

/* neighbor.rs:60-72 */
int or_needs_rebuild(int64_t n, const double *x, const double *y, const double *z,
                     const double *ref_x, const double *ref_y, const double *ref_z,
                     double r_skin, double r_cutoff) {
    double half_skin = (r_skin - r_cutoff) / 2.0;
    double threshold_sq = half_skin * half_skin;
    for (int64_t i = 0; i < n; i++) {
        double dx = x[i] - ref_x[i];
        double dy = y[i] - ref_y[i];
        double dz = z[i] - ref_z[i];
        if (dx * dx + dy * dy + dz * dz > threshold_sq) return 1;
    }
    return 0;
}

/* force/nonbonded.rs:20-82: switched LJ with Lorentz-Berthelot combining over
 * the given pairs; accumulates into f{x,y,z}; returns the energy. */
double or_pair_forces(const double *x, const double *y, const double *z, const double *sigma,
                      const double *epsilon, int64_t np, const uint32_t *pi, const uint32_t *pj,
                      double r_cutoff, double r_switch, const double *box,
                      double *fx, double *fy, double *fz) {
    double potential = 0.0;
    for (int64_t p = 0; p < np; p++) {
        uint32_t i = pi[p], j = pj[p];
        double dx = min_image(x[j] - x[i], box[0]);
        double dy = min_image(y[j] - y[i], box[1]);
        double dz = min_image(z[j] - z[i], box[2]);
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > r_cutoff * r_cutoff || r2 == 0.0) continue;
        double r = sqrt(r2);
        double sig = 0.5 * (sigma[i] + sigma[j]);
        double eps = sqrt(epsilon[i] * epsilon[j]);
        if (eps == 0.0) continue;
        /* powi(6): repeated squaring, ((s^2)^2)*s^2 as LLVM's powi lowering */
        double s = sig / r;
        double s2 = s * s;
        double sr6 = s2 * s2 * s2;
        double sr12 = sr6 * sr6;
        double v_lj = 4.0 * eps * (sr12 - sr6);
        double f_lj = 24.0 * eps / r * (2.0 * sr12 - sr6);
        double v, f_radial;
        if (r <= r_switch) {
            v = v_lj;
            f_radial = f_lj;
        } else {
            double u = (r - r_switch) / (r_cutoff - r_switch);
            double u2 = u * u;
            double sw = 1.0 - 10.0 * u2 * u + 15.0 * u2 * u2 - 6.0 * u2 * u2 * u;
            double ds_dr = -30.0 * u2 * (1.0 - u) * (1.0 - u) / (r_cutoff - r_switch);
            v = v_lj * sw;
            f_radial = f_lj * sw - v_lj * ds_dr;
        }
        potential += v;
        double ffx = f_radial * dx / r;
        double ffy = f_radial * dy / r;
        double ffz = f_radial * dz / r;
        fx[i] -= ffx; fy[i] -= ffy; fz[i] -= ffz;
        fx[j] += ffx; fy[j] += ffy; fz[j] += ffz;
    }
    return potential;
}

/* Displacement used by the bonded terms: raw difference as the reference
 * (bonded.rs:19-21, 59-68, 129-131), or min-imaged when min_img != 0 (an
 * extension for periodic chains that the reference does not have). */
static inline double bdiff(double a, double b, int min_img, double box_len) {
    double d = a - b;
    return min_img ? min_image(d, box_len) : d;
}

/* force/bonded.rs:5-40 */
double or_bond_forces(const double *x, const double *y, const double *z, int64_t nb,
                      const uint32_t *bi, const uint32_t *bj, const double *bk, const double *br0,
                      int min_img, const double *box, double *fx, double *fy, double *fz) {
    double potential = 0.0;
    for (int64_t n = 0; n < nb; n++) {
        uint32_t i = bi[n], j = bj[n];
        double k = bk[n], r0 = br0[n];
        double dx = bdiff(x[j], x[i], min_img, box[0]);
        double dy = bdiff(y[j], y[i], min_img, box[1]);
        double dz = bdiff(z[j], z[i], min_img, box[2]);
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (r == 0.0) continue;
        double dv_dr = 2.0 * k * (r - r0);
        potential += k * (r - r0) * (r - r0);
        double ux = dx / r, uy = dy / r, uz = dz / r;
        fx[i] += dv_dr * ux; fy[i] += dv_dr * uy; fz[i] += dv_dr * uz;
        fx[j] -= dv_dr * ux; fy[j] -= dv_dr * uy; fz[j] -= dv_dr * uz;
    }
    return potential;
}

/* force/bonded.rs:44-102 */
double or_angle_forces(const double *x, const double *y, const double *z, int64_t na,
                       const uint32_t *ai, const uint32_t *aj, const uint32_t *ak,
                       const double *akth, const double *ath0, int min_img, const double *box,
                       double *fx, double *fy, double *fz) {
    double potential = 0.0;
    for (int64_t n = 0; n < na; n++) {
        uint32_t i = ai[n], j = aj[n], kk = ak[n];
        double k_const = akth[n], theta0 = ath0[n];
        double b1[3] = {bdiff(x[i], x[j], min_img, box[0]), bdiff(y[i], y[j], min_img, box[1]),
                        bdiff(z[i], z[j], min_img, box[2])};
        double b2[3] = {bdiff(x[kk], x[j], min_img, box[0]), bdiff(y[kk], y[j], min_img, box[1]),
                        bdiff(z[kk], z[j], min_img, box[2])};
        double r1 = sqrt(b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]);
        double r2 = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
        if (r1 == 0.0 || r2 == 0.0) continue;
        double dot = b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2];
        double cos_theta = dot / (r1 * r2);
        if (cos_theta < -1.0) cos_theta = -1.0;
        if (cos_theta > 1.0) cos_theta = 1.0;
        double sin_theta = sqrt(1.0 - cos_theta * cos_theta);
        if (sin_theta < 1e-8) sin_theta = 1e-8;
        double theta = acos(cos_theta);
        double dv_dtheta = 2.0 * k_const * (theta - theta0);
        potential += k_const * (theta - theta0) * (theta - theta0);
        double a = dv_dtheta / sin_theta;
        double f_i[3], f_k[3];
        for (int c = 0; c < 3; c++) {
            f_i[c] = a * (b2[c] / (r1 * r2) - cos_theta * b1[c] / (r1 * r1));
            f_k[c] = a * (b1[c] / (r1 * r2) - cos_theta * b2[c] / (r2 * r2));
        }
        fx[i] += f_i[0]; fy[i] += f_i[1]; fz[i] += f_i[2];
        fx[kk] += f_k[0]; fy[kk] += f_k[1]; fz[kk] += f_k[2];
        fx[j] -= f_i[0] + f_k[0]; fy[j] -= f_i[1] + f_k[1]; fz[j] -= f_i[2] + f_k[2];
    }
    return potential;
}

static inline void cross3(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* force/bonded.rs:112-164 (Blondel-Karplus forces) */
double or_dihedral_forces(const double *x, const double *y, const double *z, int64_t nd,
                          const uint32_t *di, const uint32_t *dj, const uint32_t *dk,
                          const uint32_t *dl, const double *dkphi, const uint32_t *dn,
                          const double *ddelta, int min_img, const double *box,
                          double *fx, double *fy, double *fz) {
    double potential = 0.0;
    for (int64_t n = 0; n < nd; n++) {
        uint32_t i = di[n], j = dj[n], k = dk[n], l = dl[n];
        double k_phi = dkphi[n], mult = (double)dn[n], delta = ddelta[n];
        double b1[3] = {bdiff(x[j], x[i], min_img, box[0]), bdiff(y[j], y[i], min_img, box[1]),
                        bdiff(z[j], z[i], min_img, box[2])};
        double b2[3] = {bdiff(x[k], x[j], min_img, box[0]), bdiff(y[k], y[j], min_img, box[1]),
                        bdiff(z[k], z[j], min_img, box[2])};
        double b3[3] = {bdiff(x[l], x[k], min_img, box[0]), bdiff(y[l], y[k], min_img, box[1]),
                        bdiff(z[l], z[k], min_img, box[2])};
        double m[3], nv[3], mn[3];
        cross3(b1, b2, m);
        cross3(b2, b3, nv);
        double m_len2 = dot3(m, m);
        double n_len2 = dot3(nv, nv);
        double b2_len = sqrt(dot3(b2, b2));
        if (m_len2 == 0.0 || n_len2 == 0.0 || b2_len == 0.0) continue;
        double xx = dot3(m, nv);
        cross3(m, nv, mn);
        double yy = dot3(mn, b2) / b2_len;
        double phi = atan2(yy, xx);
        potential += k_phi * (1.0 + cos(mult * phi - delta));
        double dv_dphi = k_phi * mult * sin(mult * phi - delta);
        double si = -dv_dphi * b2_len / m_len2;
        double sl = dv_dphi * b2_len / n_len2;
        double f_i[3] = {m[0] * si, m[1] * si, m[2] * si};
        double f_l[3] = {nv[0] * sl, nv[1] * sl, nv[2] * sl};
        double p = dot3(b1, b2) / (b2_len * b2_len);
        double q = dot3(b3, b2) / (b2_len * b2_len);
        double cj = -(p + 1.0), ck = q + 1.0;
        double f_j[3], f_k[3];
        for (int c = 0; c < 3; c++) {
            f_j[c] = f_i[c] * cj + f_l[c] * q;
            f_k[c] = f_i[c] * p - f_l[c] * ck;
        }
        fx[i] += f_i[0]; fy[i] += f_i[1]; fz[i] += f_i[2];
        fx[j] += f_j[0]; fy[j] += f_j[1]; fz[j] += f_j[2];
        fx[k] += f_k[0]; fy[k] += f_k[1]; fz[k] += f_k[2];
        fx[l] += f_l[0]; fy[l] += f_l[1]; fz[l] += f_l[2];
    }
    return potential;
}

/* Scaled 1-4 Lennard-Jones over an explicit pair list.  No reference
 * counterpart (prmtop.rs:132-137 drops the 1-4 flags and fully excludes 1-4
 * pairs).  scale * unswitched LJ, no cutoff, min-image displacement and the
 * same sign convention as force/nonbonded.rs:68-78. */
double or_lj14_forces(const double *x, const double *y, const double *z, const double *sigma,
                      const double *epsilon, int64_t np, const uint32_t *pi, const uint32_t *pj,
                      double scale, const double *box, double *fx, double *fy, double *fz) {
    double potential = 0.0;
    for (int64_t p = 0; p < np; p++) {
        uint32_t i = pi[p], j = pj[p];
        double dx = min_image(x[j] - x[i], box[0]);
        double dy = min_image(y[j] - y[i], box[1]);
        double dz = min_image(z[j] - z[i], box[2]);
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 == 0.0) continue;
        double r = sqrt(r2);
        double sig = 0.5 * (sigma[i] + sigma[j]);
        double eps = sqrt(epsilon[i] * epsilon[j]);
        if (eps == 0.0) continue;
        double s = sig / r;
        double s2 = s * s;
        double sr6 = s2 * s2 * s2;
        double sr12 = sr6 * sr6;
        double v = scale * (4.0 * eps * (sr12 - sr6));
        double f_radial = scale * (24.0 * eps / r * (2.0 * sr12 - sr6));
        potential += v;
        double ffx = f_radial * dx / r, ffy = f_radial * dy / r, ffz = f_radial * dz / r;
        fx[i] -= ffx; fy[i] -= ffy; fz[i] -= ffz;
        fx[j] += ffx; fy[j] += ffy; fz[j] += ffz;
    }
    return potential;
}

/* ---- CpuBackend::compute_forces (cpu_backend.rs:92-182) ------------------ */

typedef struct {
    int64_t n_atoms;
    const double *x, *y, *z, *sigma, *epsilon;
    double box[3];
    int periodic;
    double r_cutoff, r_switch;
    int64_t nb; const uint32_t *bi, *bj; const double *bk, *br0;
    int64_t na; const uint32_t *ai, *aj, *ak; const double *akth, *ath0;
    int64_t nd; const uint32_t *di, *dj, *dk, *dl; const double *dkphi; const uint32_t *dn;
    const double *ddelta;
    int64_t n14; const uint32_t *p14i, *p14j; double scale14;
    int bonded_min_image;
    int64_t np; const uint32_t *pi, *pj;
    int n_threads;
} or_system;

typedef struct {
    const or_system *s;
    int64_t start, end;
    double *fx, *fy, *fz;
    double energy;
} strip_job;

static void *strip_worker(void *arg) {
    strip_job *jb = (strip_job *)arg;
    const or_system *s = jb->s;
    memset(jb->fx, 0, sizeof(double) * (size_t)s->n_atoms);
    memset(jb->fy, 0, sizeof(double) * (size_t)s->n_atoms);
    memset(jb->fz, 0, sizeof(double) * (size_t)s->n_atoms);
    jb->energy = 0.0;
    if (jb->start < jb->end)
        jb->energy = or_pair_forces(s->x, s->y, s->z, s->sigma, s->epsilon, jb->end - jb->start,
                                    s->pi + jb->start, s->pj + jb->start, s->r_cutoff, s->r_switch,
                                    s->box, jb->fx, jb->fy, jb->fz);
    return NULL;
}

/* terms[6] = {bond, angle, dihedral, lj, lj14, total}.  With scale14 == 0 and
 * bonded_min_image == 0 this is the reference CpuBackend exactly: serial
 * bonded terms, T static strips of ceil(np/T) pairs with private buffers,
 * energy and forces reduced in thread-index order.  Returns 0, or -1 on
 * allocation / thread failure. */
int or_cpu_backend_compute(const or_system *s, double *fx, double *fy, double *fz, double *terms) {
    int64_t n = s->n_atoms;
    int T = s->n_threads < 1 ? 1 : s->n_threads;
    double *bfx = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *bfy = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *bfz = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    double *tb = (double *)malloc(sizeof(double) * 3 * (size_t)T * (size_t)(n > 0 ? n : 1));
    strip_job *jobs = (strip_job *)calloc((size_t)T, sizeof(strip_job));
    pthread_t *th = (pthread_t *)calloc((size_t)T, sizeof(pthread_t));
    if (!bfx || !bfy || !bfz || !tb || !jobs || !th) {
        free(bfx); free(bfy); free(bfz); free(tb); free(jobs); free(th); return -1;
    }
    double potential = 0.0;
    int mi = s->bonded_min_image;
    double e_bond = or_bond_forces(s->x, s->y, s->z, s->nb, s->bi, s->bj, s->bk, s->br0, mi, s->box, bfx, bfy, bfz);
    potential += e_bond;
    double e_angle = or_angle_forces(s->x, s->y, s->z, s->na, s->ai, s->aj, s->ak, s->akth, s->ath0, mi, s->box, bfx, bfy, bfz);
    potential += e_angle;
    double e_dihed = or_dihedral_forces(s->x, s->y, s->z, s->nd, s->di, s->dj, s->dk, s->dl, s->dkphi, s->dn, s->ddelta, mi, s->box, bfx, bfy, bfz);
    potential += e_dihed;
    double e14 = 0.0;
    if (s->n14 > 0 && s->scale14 != 0.0) {
        e14 = or_lj14_forces(s->x, s->y, s->z, s->sigma, s->epsilon, s->n14, s->p14i, s->p14j, s->scale14, s->box, bfx, bfy, bfz);
        potential += e14;
    }
    int64_t chunk = (s->np + T - 1) / T;
    if (chunk < 1) chunk = 1;
    int rc = 0;
    for (int k = 0; k < T; k++) {
        jobs[k].s = s;
        jobs[k].start = (int64_t)k * chunk < s->np ? (int64_t)k * chunk : s->np;
        jobs[k].end = (int64_t)(k + 1) * chunk < s->np ? (int64_t)(k + 1) * chunk : s->np;
        jobs[k].fx = tb + (size_t)(3 * k) * (size_t)n;
        jobs[k].fy = tb + (size_t)(3 * k + 1) * (size_t)n;
        jobs[k].fz = tb + (size_t)(3 * k + 2) * (size_t)n;
        if (T == 1) strip_worker(&jobs[k]);
        else if (pthread_create(&th[k], NULL, strip_worker, &jobs[k])) { rc = -1; T = k; break; }
    }
    if (s->n_threads > 1)
        for (int k = 0; k < T; k++) pthread_join(th[k], NULL);
    double e_lj = 0.0;
    for (int k = 0; k < T; k++) { e_lj += jobs[k].energy; potential += jobs[k].energy; }
    for (int64_t i = 0; i < n; i++) { fx[i] = bfx[i]; fy[i] = bfy[i]; fz[i] = bfz[i]; }
    for (int k = 0; k < T; k++)
        for (int64_t i = 0; i < n; i++) {
            fx[i] += jobs[k].fx[i]; fy[i] += jobs[k].fy[i]; fz[i] += jobs[k].fz[i];
        }
    terms[0] = e_bond; terms[1] = e_angle; terms[2] = e_dihed;
    terms[3] = e_lj; terms[4] = e14; terms[5] = potential;
    free(bfx); free(bfy); free(bfz); free(tb); free(jobs); free(th);
    return rc;
}

/* ---- integrator and constraints (for the golden-trajectory pin only) ----- */

#define FORCE_TO_ACCEL_ANG_PER_PS2 (20.455 * 20.455) /* integrator.rs:9 */

/* integrator.rs:14-21 */
void or_half_kick(int64_t n, double *vx, double *vy, double *vz, const double *fx,
                  const double *fy, const double *fz, const double *mass, double dt_half) {
    for (int64_t i = 0; i < n; i++) {
        double inv_m = 1.0 / mass[i];
        vx[i] += dt_half * fx[i] * inv_m * FORCE_TO_ACCEL_ANG_PER_PS2;
        vy[i] += dt_half * fy[i] * inv_m * FORCE_TO_ACCEL_ANG_PER_PS2;
        vz[i] += dt_half * fz[i] * inv_m * FORCE_TO_ACCEL_ANG_PER_PS2;
    }
}

/* constraint.rs:133-223 (Jacobi SHAKE).  Returns iterations (>0) or -1 when
 * max_iter is exhausted. */
int or_shake_solve(int64_t n, int64_t nc, const uint32_t *ci, const uint32_t *cj,
                   const double *dsq, const double *mass, const double *rx0, const double *ry0,
                   const double *rz0, double *x, double *y, double *z, int max_iter, double tol) {
    if (nc == 0) return 0;
    double *ddx = (double *)malloc(sizeof(double) * (size_t)n);
    double *ddy = (double *)malloc(sizeof(double) * (size_t)n);
    double *ddz = (double *)malloc(sizeof(double) * (size_t)n);
    int result = -1;
    for (int it = 0; it < max_iter; it++) {
        memset(ddx, 0, sizeof(double) * (size_t)n);
        memset(ddy, 0, sizeof(double) * (size_t)n);
        memset(ddz, 0, sizeof(double) * (size_t)n);
        double max_lambda = 0.0;
        for (int64_t c = 0; c < nc; c++) {
            uint32_t i = ci[c], j = cj[c];
            double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
            double diff = dx * dx + dy * dy + dz * dz - dsq[c];
            double rx = rx0[i] - rx0[j], ry = ry0[i] - ry0[j], rz = rz0[i] - rz0[j];
            double inv_mi = 1.0 / mass[i], inv_mj = 1.0 / mass[j];
            double dot_new_old = dx * rx + dy * ry + dz * rz;
            double denom = 2.0 * (inv_mi + inv_mj) * dot_new_old;
            double lambda = denom == 0.0 ? 0.0 : diff / denom;
            ddx[i] -= lambda * inv_mi * rx; ddy[i] -= lambda * inv_mi * ry; ddz[i] -= lambda * inv_mi * rz;
            ddx[j] += lambda * inv_mj * rx; ddy[j] += lambda * inv_mj * ry; ddz[j] += lambda * inv_mj * rz;
            double la = fabs(lambda);
            if (la > max_lambda) max_lambda = la;
        }
        for (int64_t a = 0; a < n; a++) { x[a] += ddx[a]; y[a] += ddy[a]; z[a] += ddz[a]; }
        if (max_lambda < tol) { result = it + 1; break; }
    }
    free(ddx); free(ddy); free(ddz);
    return result;
}

/* integrator.rs:32-71 */
int or_drift_and_constrain(int64_t n, double *x, double *y, double *z, double *vx, double *vy,
                           double *vz, int64_t nc, const uint32_t *ci, const uint32_t *cj,
                           const double *dsq, const double *mass, double dt, int max_iter, double tol) {
    double *rx = (double *)malloc(sizeof(double) * (size_t)n);
    double *ry = (double *)malloc(sizeof(double) * (size_t)n);
    double *rz = (double *)malloc(sizeof(double) * (size_t)n);
    memcpy(rx, x, sizeof(double) * (size_t)n);
    memcpy(ry, y, sizeof(double) * (size_t)n);
    memcpy(rz, z, sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; i++) { x[i] += dt * vx[i]; y[i] += dt * vy[i]; z[i] += dt * vz[i]; }
    int it = or_shake_solve(n, nc, ci, cj, dsq, mass, rx, ry, rz, x, y, z, max_iter, tol);
    if (it >= 0)
        for (int64_t i = 0; i < n; i++) {
            vx[i] = (x[i] - rx[i]) / dt; vy[i] = (y[i] - ry[i]) / dt; vz[i] = (z[i] - rz[i]) / dt;
        }
    free(rx); free(ry); free(rz);
    return it;
}

/* constraint.rs:241-330 (RATTLE velocity projection, Jacobi). */
int or_constrain_velocities(int64_t n, const double *x, const double *y, const double *z,
                            double *vx, double *vy, double *vz, int64_t nc, const uint32_t *ci,
                            const uint32_t *cj, const double *mass, int max_iter, double tol) {
    if (nc == 0) return 0;
    double *ddx = (double *)malloc(sizeof(double) * (size_t)n);
    double *ddy = (double *)malloc(sizeof(double) * (size_t)n);
    double *ddz = (double *)malloc(sizeof(double) * (size_t)n);
    int result = -1;
    for (int it = 0; it < max_iter; it++) {
        memset(ddx, 0, sizeof(double) * (size_t)n);
        memset(ddy, 0, sizeof(double) * (size_t)n);
        memset(ddz, 0, sizeof(double) * (size_t)n);
        double max_mu = 0.0;
        for (int64_t c = 0; c < nc; c++) {
            uint32_t i = ci[c], j = cj[c];
            double rx = x[i] - x[j], ry = y[i] - y[j], rz = z[i] - z[j];
            double rsq = rx * rx + ry * ry + rz * rz;
            double dvx = vx[i] - vx[j], dvy = vy[i] - vy[j], dvz = vz[i] - vz[j];
            double rel_vel = dvx * rx + dvy * ry + dvz * rz;
            double inv_mi = 1.0 / mass[i], inv_mj = 1.0 / mass[j];
            double denom = (inv_mi + inv_mj) * rsq;
            double mu = denom == 0.0 ? 0.0 : rel_vel / denom;
            ddx[i] -= mu * inv_mi * rx; ddy[i] -= mu * inv_mi * ry; ddz[i] -= mu * inv_mi * rz;
            ddx[j] += mu * inv_mj * rx; ddy[j] += mu * inv_mj * ry; ddz[j] += mu * inv_mj * rz;
            double ma = fabs(mu);
            if (ma > max_mu) max_mu = ma;
        }
        for (int64_t a = 0; a < n; a++) { vx[a] += ddx[a]; vy[a] += ddy[a]; vz[a] += ddz[a]; }
        if (max_mu < tol) { result = it + 1; break; }
    }
    free(ddx); free(ddy); free(ddz);
    return result;
}
