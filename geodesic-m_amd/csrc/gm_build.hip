// Verlet pair-tile list build for gfx950.
//
// Replaces neighbor::build (neighbor.rs:8-55; paths relative to the reference
// repository root) and the CSR expansion of neighbor_csr.rs:2-23.  The pair
// set is identical to the reference's: every (i<j) not excluded with
// min-image r2 <= r_skin^2, decided in f64 with the reference's expression.
// This file is compiled with -ffp-contract=off so that decision matches the
// f64 CPU path bit for bit.
//
// Layout produced:
//   slots      atoms reordered spatially; supercluster = 64 slots, cluster = 8
//   entries    per supercluster, the j-clusters whose bounding box comes within
//              r_skin of it, each with an 8-bit i-cluster mask and eight 64-bit
//              pair masks (bit jl*8+il: i slot ci*8+il vs j slot cj*8+jl)
// Every unordered pair appears in exactly one tile: i-cluster < j-cluster,
// or the same cluster with i slot < j slot.
#include "gm_internal.hpp"
#include "gm_geom.hpp"

using namespace gm;

namespace {

__device__ inline double wrap_coord(double x, double L) { return x - L * floor(x / L); }
__device__ inline double min_image(double d, double L) { return d - L * round(d / L); }

__device__ inline uint32_t quantize(double x, double inv_q) { return (uint32_t)quantize_fixed(x, inv_q); }

__global__ void k_fill_i32(int *p, size_t n, int v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Minimum over a full wavefront (DPP within rows, permlane swaps across them);
// every lane receives it.  All 64 lanes must be active.
__device__ inline int wave_min_i32(int v) {
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
    v = min(v, __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false));  // row_ror:8
    const auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    v = min((int)r[0], (int)r[1]);
    const auto q = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return min((int)q[0], (int)q[1]);
}

// counter[key] += 1 for every active lane, with one atomic per distinct key
// in the wavefront (atoms in input order are often spatially coherent, and 64
// same-address atomics in one instruction serialise).  Returns the counter's
// old value plus the lane's rank among the active lanes with its key: the
// lane's slot in its bucket.  Every lane of the wavefront must call it.
__device__ inline int aggregated_fetch_add(int *counter, int key, bool active) {
    const int lane = threadIdx.x & 63;
    bool todo = active;
    int old = 0;
    while (true) {
        const unsigned long long pend = __ballot(todo);
        if (pend == 0) break;
        const int leader = __ffsll((long long)pend) - 1;
        const int k0 = __shfl(key, leader);
        const bool mine = todo && key == k0;
        const unsigned long long grp = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counter[k0], __popcll(grp));
        base = __shfl(base, leader);
        if (mine) {
            old = base + __popcll(grp & ((1ull << lane) - 1ull));
            todo = false;
        }
    }
    return old;
}

// neighbor.rs:13-15 + 78-88: wrap (periodic builds only), snapshot the
// reference positions (neighbor.rs:48-50), quantise, bin into xy columns.
__global__ void k_wrap_quantize(int n, double *x, double *y, double *z, double *rx, double *ry,
                                double *rz, uint32_t *q, int *col_count, BuildParams bp, int wrap) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i0 < n;
    const int i = active ? i0 : 0;
    int col = 0;
    if (active) {
        double xi = x[i], yi = y[i], zi = z[i];
        if (wrap) {
            xi = wrap_coord(xi, bp.box[0]);
            yi = wrap_coord(yi, bp.box[1]);
            zi = wrap_coord(zi, bp.box[2]);
            x[i] = xi; y[i] = yi; z[i] = zi;
        }
        rx[i] = xi; ry[i] = yi; rz[i] = zi;
        uint32_t qx = quantize(xi, bp.inv_q[0]);
        uint32_t qy = quantize(yi, bp.inv_q[1]);
        uint32_t qz = quantize(zi, bp.inv_q[2]);
        q[i] = qx; q[n + i] = qy; q[2 * n + i] = qz;
        int cx = (int)(((uint64_t)qx * (uint64_t)bp.ncx) >> 32);
        int cy = (int)(((uint64_t)qy * (uint64_t)bp.ncy) >> 32);
        col = cx * bp.ncy + cy;
    }
    (void)aggregated_fetch_add(col_count, col, active);
}

// Exclusive scans over columns: atom offsets and supercluster offsets; total
// supercluster count, column overflow flag and largest column into meta.
__global__ __launch_bounds__(1024) void k_scan_columns(int ncol, const int *cnt, int *start,
                                                      int *colsc, int *meta, int cap) {
    __shared__ int s1[1024], s2[1024], sm[1024];
    int t = threadIdx.x;
    int per = (ncol + 1023) / 1024;
    int b = t * per, e = min(b + per, ncol);
    int a1 = 0, a2 = 0, mx = 0;
    for (int i = b; i < e; i++) {
        int c = cnt[i];
        a1 += c; a2 += (c + 63) / 64; mx = max(mx, c);
    }
    s1[t] = a1; s2[t] = a2; sm[t] = mx;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v1 = t >= off ? s1[t - off] : 0;
        int v2 = t >= off ? s2[t - off] : 0;
        int v3 = t >= off ? sm[t - off] : 0;
        __syncthreads();
        s1[t] += v1; s2[t] += v2; sm[t] = max(sm[t], v3);
        __syncthreads();
    }
    int ex1 = s1[t] - a1, ex2 = s2[t] - a2;
    for (int i = b; i < e; i++) {
        int c = cnt[i];
        start[i] = ex1; colsc[i] = ex2;
        ex1 += c; ex2 += (c + 63) / 64;
    }
    if (t == 1023) {
        meta[0] = s2[1023];
        meta[1] = sm[1023] > cap ? 1 : 0;
        meta[2] = sm[1023];
    }
}

__global__ void k_scatter(int n, const uint32_t *q, const int *col_start, int *col_fill, int *bucket,
                          BuildParams bp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < n;
    int c = 0;
    if (active) {
        const int cx = (int)(((uint64_t)q[i] * (uint64_t)bp.ncx) >> 32);
        const int cy = (int)(((uint64_t)q[n + i] * (uint64_t)bp.ncy) >> 32);
        c = cx * bp.ncy + cy;
    }
    const int slot = aggregated_fetch_add(col_fill, c, active);
    if (active) bucket[col_start[c] + slot] = i;
}

// The same two passes with a per-workgroup column histogram in LDS (up to
// kLdsColumns columns): one global atomic per (workgroup, column) instead of
// one per (wavefront, column), all issued at once rather than one returning
// atomic after another.
__global__ __launch_bounds__(1024) void k_wrap_quantize_lds(int n, double *x, double *y, double *z, double *rx,
                                                           double *ry, double *rz, uint32_t *q, int *col_count,
                                                           BuildParams bp, int wrap) {
    extern __shared__ int hist[];
    const int t = threadIdx.x, i = blockIdx.x * 1024 + t;
    for (int c = t; c < bp.ncol; c += 1024) hist[c] = 0;
    __syncthreads();
    if (i < n) {
        double xi = x[i], yi = y[i], zi = z[i];
        if (wrap) {
            xi = wrap_coord(xi, bp.box[0]);
            yi = wrap_coord(yi, bp.box[1]);
            zi = wrap_coord(zi, bp.box[2]);
            x[i] = xi; y[i] = yi; z[i] = zi;
        }
        rx[i] = xi; ry[i] = yi; rz[i] = zi;
        const uint32_t qx = quantize(xi, bp.inv_q[0]), qy = quantize(yi, bp.inv_q[1]), qz = quantize(zi, bp.inv_q[2]);
        q[i] = qx; q[n + i] = qy; q[2 * n + i] = qz;
        const int cx = (int)(((uint64_t)qx * (uint64_t)bp.ncx) >> 32);
        const int cy = (int)(((uint64_t)qy * (uint64_t)bp.ncy) >> 32);
        atomicAdd(&hist[cx * bp.ncy + cy], 1);
    }
    __syncthreads();
    for (int c = t; c < bp.ncol; c += 1024)
        if (hist[c]) atomicAdd(&col_count[c], hist[c]);
}

__global__ __launch_bounds__(1024) void k_scatter_lds(int n, const uint32_t *q, const int *col_start, int *col_fill,
                                                     int *bucket, BuildParams bp) {
    extern __shared__ int hist[];  // counts, then each column's base for this workgroup
    const int t = threadIdx.x, i = blockIdx.x * 1024 + t;
    for (int c = t; c < bp.ncol; c += 1024) hist[c] = 0;
    __syncthreads();
    int c = 0, rank = 0;
    if (i < n) {
        const int cx = (int)(((uint64_t)q[i] * (uint64_t)bp.ncx) >> 32);
        const int cy = (int)(((uint64_t)q[n + i] * (uint64_t)bp.ncy) >> 32);
        c = cx * bp.ncy + cy;
        rank = atomicAdd(&hist[c], 1);
    }
    __syncthreads();
    for (int k = t; k < bp.ncol; k += 1024) {
        const int h = hist[k];
        if (h) hist[k] = atomicAdd(&col_fill[k], h);
    }
    __syncthreads();
    if (i < n) bucket[col_start[c] + hist[c] + rank] = i;
}

// One workgroup per column: bitonic sort of (z, atom) in LDS, which fixes the
// order independently of the scatter's atomics; cut into superclusters of 64.
__global__ __launch_bounds__(1024) void k_column_sort(int n, const uint32_t *q, const int *bucket,
                                                     const int *col_start, const int *col_count,
                                                     const int *col_sc, int *sorted_atoms,
                                                     int *sc_base, int *sc_m, int *sc_col) {
    extern __shared__ unsigned long long keys[];
    const int c = blockIdx.x;
    const int m = col_count[c];
    if (m > kColumnCap || m == 0) return;
    const int base = col_start[c];
    int p2 = 1;
    while (p2 < m) p2 <<= 1;
    for (int t = threadIdx.x; t < p2; t += blockDim.x) {
        if (t < m) {
            int a = bucket[base + t];
            keys[t] = ((unsigned long long)q[2 * n + a] << 32) | (unsigned)a;
        } else {
            keys[t] = ~0ull;
        }
    }
    __syncthreads();
    for (int k = 2; k <= p2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < p2; t += blockDim.x) {
                int ixj = t ^ j;
                if (ixj > t) {
                    unsigned long long a = keys[t], b = keys[ixj];
                    bool up = (t & k) == 0;
                    if ((a > b) == up) { keys[t] = b; keys[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < m; t += blockDim.x) sorted_atoms[base + t] = (int)(keys[t] & 0xffffffffu);
    const int nsc_c = (m + 63) / 64, s0 = col_sc[c];
    for (int k = threadIdx.x; k < nsc_c; k += blockDim.x) {
        sc_base[s0 + k] = base + 64 * k;
        sc_m[s0 + k] = min(64, m - 64 * k);
        sc_col[s0 + k] = c;
    }
}

// Split a full supercluster (64 atoms, z-sorted) into 2x2x2 clusters of 8:
// halves by x, quarters by y, eighths by z (ties broken by atom index).
// Partial superclusters keep z order and are cut into consecutive eights.
__global__ __launch_bounds__(64) void k_sc_split(const int *meta, int n, const uint32_t *q,
                                                 const int *sorted_atoms, const int *sc_base,
                                                 const int *sc_m, int *slot_atom) {
    __shared__ unsigned long long key[64];
    const int sc = blockIdx.x;
    if (meta[1] || sc >= meta[0]) return;  // a column overflowed: sc_base is not set
    const int t = threadIdx.x;
    const int m = sc_m[sc];
    const int a = t < m ? sorted_atoms[sc_base[sc] + t] : -1;
    int pos = t;
    if (m == 64) {
        for (int level = 0; level < 3; level++) {
            const int gsize = 64 >> level;
            const unsigned long long mine = ((unsigned long long)q[level * n + a] << 32) | (unsigned)a;
            key[pos] = mine;
            __syncthreads();
            const int g = pos / gsize;
            int r = 0;
            for (int u = 0; u < gsize; u++) r += key[g * gsize + u] < mine ? 1 : 0;
            __syncthreads();
            pos = g * gsize + r;
        }
    }
    slot_atom[(size_t)sc * 64 + pos] = a;
}

// Slot tables.  qpos/qeps (the force kernel's per-slot inputs) are filled for
// every slot, padding included, so the per-evaluation prep kernel only needs
// to overwrite the atoms' slots.
// sq.w = 1 marks an atom whose position is not finite: such an atom is in no
// listed pair (the reference's r2 <= r_skin^2 is false for a NaN distance).
__global__ void k_slot_setup(int nslots, int n, const int *slot_atom, const uint32_t *q,
                             const float2 *atom_par, const uint2 *excl_bits, int *atom_slot, int4 *sq,
                             float2 *sp, int4 *qpos, float *qeps, uint2 *sxb, float inv_s,
                             const double *x, const double *y, const double *z) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    int a = slot_atom[s];
    if (a >= 0) {
        atom_slot[a] = s;
        const bool fin = isfinite(x[a]) && isfinite(y[a]) && isfinite(z[a]);
        const int4 v = make_int4((int)q[a], (int)q[n + a], (int)q[2 * n + a], fin ? 0 : 1);
        const float2 p = atom_par[a];
        sq[s] = v;
        sp[s] = p;
        qpos[s] = make_int4(v.x, v.y, v.z, __float_as_int(p.x * inv_s));
        qeps[s] = 2.0f * p.y;
        sxb[s] = excl_bits[a];
    } else {
        sq[s] = make_int4(0, 0, 0, 0);
        sp[s] = make_float2(0.f, 0.f);
        qpos[s] = make_int4(0, 0, 0, 0);
        qeps[s] = 0.f;
        sxb[s] = make_uint2(0u, 0u);
    }
}

// Bounding boxes (fixed-point centre, half extents in A) of every cluster and
// supercluster.  A cluster is valid iff its first slot holds an atom.  Offsets
// are taken relative to the group's first atom (minimum image), so min <= 0 <=
// max; spans can exceed half the box in small boxes, hence computed in float.
__global__ __launch_bounds__(64) void k_bbox(const int *meta, const int *slot_atom, const int4 *sq,
                                             BuildParams bp, int4 *cl_ref, float4 *cl_half,
                                             int4 *sc_ref, float4 *sc_half) {
    const int sc = blockIdx.x;
    if (sc >= meta[0]) return;
    const int t = threadIdx.x;
    const size_t s = (size_t)sc * 64 + t;
    const bool valid = slot_atom[s] >= 0;
    const int4 v = sq[s];
    const size_t sr = (size_t)sc * 64 + (t & ~7);
    const int4 r = sq[sr];
    const bool cvalid = slot_atom[sr] >= 0;
    int dx = valid ? wrapdiff(v.x, r.x) : 0, dy = valid ? wrapdiff(v.y, r.y) : 0,
        dz = valid ? wrapdiff(v.z, r.z) : 0;
    int mnx = dx, mxx = dx, mny = dy, mxy = dy, mnz = dz, mxz = dz;
    for (int o = 1; o < 8; o <<= 1) {
        mnx = min(mnx, __shfl_xor(mnx, o)); mxx = max(mxx, __shfl_xor(mxx, o));
        mny = min(mny, __shfl_xor(mny, o)); mxy = max(mxy, __shfl_xor(mxy, o));
        mnz = min(mnz, __shfl_xor(mnz, o)); mxz = max(mxz, __shfl_xor(mxz, o));
    }
    if ((t & 7) == 0) {
        const int c = sc * 8 + t / 8;
        cl_ref[c] = make_int4((int)((unsigned)r.x + (unsigned)((mnx + mxx) / 2)),
                              (int)((unsigned)r.y + (unsigned)((mny + mxy) / 2)),
                              (int)((unsigned)r.z + (unsigned)((mnz + mxz) / 2)), cvalid ? 1 : 0);
        cl_half[c] = make_float4((((float)mxx - (float)mnx) * 0.5f + 1.0f) * bp.sq[0],
                                 (((float)mxy - (float)mny) * 0.5f + 1.0f) * bp.sq[1],
                                 (((float)mxz - (float)mnz) * 0.5f + 1.0f) * bp.sq[2], 0.f);
    }
    const int4 r0 = sq[(size_t)sc * 64];
    dx = valid ? wrapdiff(v.x, r0.x) : 0;
    dy = valid ? wrapdiff(v.y, r0.y) : 0;
    dz = valid ? wrapdiff(v.z, r0.z) : 0;
    mnx = mxx = dx; mny = mxy = dy; mnz = mxz = dz;
    for (int o = 1; o < 64; o <<= 1) {
        mnx = min(mnx, __shfl_xor(mnx, o)); mxx = max(mxx, __shfl_xor(mxx, o));
        mny = min(mny, __shfl_xor(mny, o)); mxy = max(mxy, __shfl_xor(mxy, o));
        mnz = min(mnz, __shfl_xor(mnz, o)); mxz = max(mxz, __shfl_xor(mxz, o));
    }
    if (t == 0) {
        sc_ref[sc] = make_int4((int)((unsigned)r0.x + (unsigned)((mnx + mxx) / 2)),
                               (int)((unsigned)r0.y + (unsigned)((mny + mxy) / 2)),
                               (int)((unsigned)r0.z + (unsigned)((mnz + mxz) / 2)), 1);
        sc_half[sc] = make_float4((((float)mxx - (float)mnx) * 0.5f + 1.0f) * bp.sq[0],
                                  (((float)mxy - (float)mny) * 0.5f + 1.0f) * bp.sq[1],
                                  (((float)mxz - (float)mnz) * 0.5f + 1.0f) * bp.sq[2], 0.f);
    }
}

// One wave per supercluster: scan the j-clusters cj >= its first cluster that
// lie in xy columns within reach of its own column (columns are periodic in
// the fixed-point frame, so every pair within r_skin is in reach).  An entry
// is (sc, cj) with a nonzero 8-bit i-cluster mask (ci <= cj, both valid,
// boxes within r_skin).  Columns are visited in ascending index, and clusters
// are numbered column by column, so entries come in ascending cj.  Pass 0
// counts, pass 1 writes entries at sc_start[sc].
template <int FILL>
__global__ __launch_bounds__(64) void k_pair_search(const int *meta, BuildParams bp, const int *sc_col,
                                                   const int *col_sc, const int *col_count,
                                                   const int4 *cl_ref, const float4 *cl_half,
                                                   const int4 *sc_ref, const float4 *sc_half,
                                                   int *sc_count, const int *sc_start, int *ent_sc,
                                                   int *ent_cj, uint32_t *ent_imask, int fill) {
    __shared__ int4 cir[8];
    __shared__ float4 cih[8];
    const int sc = blockIdx.x;
    if (meta[1] || sc >= meta[0] || (fill && meta[kMetaEntOver])) return;
    const int t = threadIdx.x;
    if (t < 8) { cir[t] = cl_ref[sc * 8 + t]; cih[t] = cl_half[sc * 8 + t]; }
    __syncthreads();
    const int4 sr = sc_ref[sc];
    const float4 sh = sc_half[sc];
    const int first = sc * 8;
    const int base = FILL ? sc_start[sc] : 0;
    const int c0 = sc_col[sc];
    const int cx0 = c0 / bp.ncy, cy0 = c0 - (c0 / bp.ncy) * bp.ncy;
    int2 px[2], py[2];
    const int npx = ring_pieces(cx0, bp.kx, bp.ncx, px);
    const int npy = ring_pieces(cy0, bp.ky, bp.ncy, py);
    const float rs2 = bp.r_search * bp.r_search;
    int count = 0;
    const unsigned long long lt_mask = (t == 0) ? 0ull : (~0ull >> (64 - t));
    for (int a = 0; a < npx; a++) {
        for (int cx = px[a].x; cx <= px[a].y; cx++) {
            const float gx = column_gap(cx - cx0, bp.ncx, bp.wcol[0]);
            for (int b = 0; b < npy; b++) {
                for (int cy = py[b].x; cy <= py[b].y; cy++) {
                    const float gy = column_gap(cy - cy0, bp.ncy, bp.wcol[1]);
                    const int c = cx * bp.ncy + cy;
                    if (c < c0 || gx * gx + gy * gy > rs2) continue;
                    const int m = col_count[c];
                    const int lo = (c == c0) ? first : col_sc[c] * 8;
                    const int hi = col_sc[c] * 8 + ((m + 63) / 64) * 8;
                    for (int j0 = lo; j0 < hi; j0 += 64) {
                        const int cj = j0 + t;
                        uint32_t imask = 0;
                        if (cj < hi) {
                            const int4 jr = cl_ref[cj];
                            const float4 jh = cl_half[cj];
                            if (jr.w && boxes_close(sr, sh, jr, jh, bp)) {
                                for (int ci = 0; ci < 8; ci++) {
                                    if (first + ci <= cj && cir[ci].w &&
                                        boxes_close(cir[ci], cih[ci], jr, jh, bp))
                                        imask |= 1u << ci;
                                }
                            }
                        }
                        const unsigned long long bl = __ballot(imask != 0);
                        if (FILL && imask) {
                            const int off = base + count + __popcll(bl & lt_mask);
                            ent_sc[off] = sc;
                            ent_cj[off] = cj;
                            ent_imask[off] = imask;
                        }
                        count += __popcll(bl);
                    }
                }
            }
        }
    }
    if (!FILL && t == 0) sc_count[sc] = count;
}

__device__ inline bool excluded(const int64_t *off, const uint32_t *nbr, uint32_t lo, uint32_t hi) {
    int64_t a = off[lo], b = off[lo + 1];
    while (a < b) {
        int64_t mid = (a + b) >> 1;
        if (nbr[mid] < hi) a = mid + 1; else b = mid;
    }
    return a < off[lo + 1] && nbr[a] == hi;
}

// One wavefront per entry (a strided walk over an XCD's range of entries,
// four wavefronts per workgroup): exact pair masks.  Lane = (jl = lane>>3, il = lane&7).
// The inclusion test is neighbor.rs:31-38: min-image r2 <= r_skin^2 in f64 on
// the wrapped positions, with i = min(atom), j = max(atom), then the exclusion
// set (neighbor.rs:33-35).  r2 is first estimated in f32 from the exact
// fixed-point min-image displacement; a pair whose estimate lies within the
// band of a threshold is decided by the f64 expression itself, and the band
// is many times the estimate's worst error, so every decision is the f64 one.
// The exclusion lookup runs only for pairs inside r_skin, and only when the
// a bit of the atom's mask of partners lo+1..lo+64, or for a partner farther
// up, a binary search when it lies in the atom's exclusion-partner range.
// Per tile, the byte floor((d_lo - r_cutoff) / step) (0 if negative, 255 for
// a tile without listed pairs) goes to ent_dmin, d_lo a lower bound on the
// distance of the tile's closest listed pair (the f32 estimate less a
// relative 2e-6 and 1e-5 A, far above its error).
__global__ __launch_bounds__(256) void k_masks(const int *meta, const int *ent_sc, const int *ent_cj,
                                               uint32_t *ent_imask, unsigned long long *ent_mask,
                                               int2 *ent_meta, uint2 *ent_dmin, uint8_t *ent_lb,
                                               const int *slot_atom, const int4 *sq, const double *x,
                                               const double *y, const double *z,
                                               const int64_t *excl_off, const uint32_t *excl_nbr,
                                               const uint2 *excl_rng, const uint2 *sxb,
                                               BuildParams bp,
                                               unsigned long long *parts) {
    const int t = threadIdx.x & 63;
    const int il = t & 7, jl = t >> 3;
    // the entries the pair search wrote (none when they overflowed the buffers)
    const int nent = (meta[1] || meta[kMetaEntOver]) ? 0 : meta[kMetaNent];
    // Workgroups are dealt to the 8 XCDs in turn (workgroup g runs on XCD
    // g % 8); group xg of the P = min(8, grid) residues walks the contiguous
    // entry range [xg n / P, (xg + 1) n / P), so an XCD's L2 holds the slot data
    // of about an eighth of the system.
    const int P = (int)min(gridDim.x, 8u);
    const int xg = (int)(blockIdx.x % (unsigned)P);
    const int nwx = (int)((gridDim.x - (unsigned)xg + (unsigned)P - 1u) / (unsigned)P) * 4;  // wavefronts of group xg
    const int l = (int)(blockIdx.x / (unsigned)P) * 4 + (int)(threadIdx.x >> 6);
    const int e_lo = (int)((int64_t)nent * xg / P), e_hi = (int)((int64_t)nent * (xg + 1) / P);
    unsigned long long tot_pairs = 0, tot_tiles = 0, tot_ctiles = 0, tot_cpairs = 0;
    for (int e = e_lo + l; e < e_hi; e += nwx) {
        const int sc = ent_sc[e], cj = ent_cj[e];
        const uint32_t im = ent_imask[e];
        const size_t js = (size_t)cj * 8 + jl;
        const int4 qb = sq[js];
        const int b = qb.w ? -1 : slot_atom[js];  // (a non-finite position is never listed)
        const uint2 xbj = sxb[js];
        // per-cluster values in arrays of scalars (no vector structs), which
        // stay in registers once the loops are unrolled; the exclusion masks
        // of the slots' atoms arrive with their coordinates
        int a[8], qax[8], qay[8], qaz[8];
        uint32_t exl[8], exh[8];
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            a[ci] = -1;
            qax[ci] = qay[ci] = qaz[ci] = 0;
            exl[ci] = exh[ci] = 0u;
            if (im & (1u << ci)) {
                const size_t is = (size_t)sc * 64 + ci * 8 + il;
                const int4 q = sq[is];
                a[ci] = q.w ? -1 : slot_atom[is];
                qax[ci] = q.x;
                qay[ci] = q.y;
                qaz[ci] = q.z;
                const uint2 xb = sxb[is];
                exl[ci] = xb.x;
                exh[ci] = xb.y;
            }
        }
        // exclusions, reduced at once to one bit per tile (the masks' registers
        // are then free): a partner within 64 indices is a bit of the lower
        // atom's mask; a farther one (far) is looked up later if listed
        uint32_t xhit = 0, far = 0;
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            if (a[ci] >= 0 && b >= 0) {
                const uint32_t lo = (uint32_t)min(a[ci], b), hi = (uint32_t)max(a[ci], b);
                const uint32_t k = hi - lo - 1u;
                const uint32_t w = k < 32u ? (a[ci] < b ? exl[ci] : xbj.x) : (a[ci] < b ? exh[ci] : xbj.y);
                if (k >= 64u) far |= 1u << ci;
                else if (((w >> (k & 31u)) & 1u) != 0u) xhit |= 1u << ci;
            }
        }
        uint32_t inb = 0, cutb = 0, band = 0;
        float r2v[8];
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            const int cg = sc * 8 + ci;
            r2v[ci] = 0.f;
            if (a[ci] >= 0 && b >= 0 && (cg < cj || il < jl)) {
                const float dx = (float)wrapdiff(qb.x, qax[ci]) * bp.sq[0];
                const float dy = (float)wrapdiff(qb.y, qay[ci]) * bp.sq[1];
                const float dz = (float)wrapdiff(qb.z, qaz[ci]) * bp.sq[2];
                const float r2f = dx * dx + dy * dy + dz * dz;
                r2v[ci] = r2f;
                const bool in = r2f <= bp.r_skin2f, cut = r2f <= bp.r_cut2f;
                if (fabsf(r2f - bp.r_skin2f) <= bp.band_skin || fabsf(r2f - bp.r_cut2f) <= bp.band_cut)
                    band |= 1u << ci;
                inb |= in ? (1u << ci) : 0u;
                cutb |= (in && cut) ? (1u << ci) : 0u;
            }
        }
        // pairs near a threshold (rare): the f64 expression decides, one at a
        // time, outside the unrolled loop
        while (band) {
            const int ci = __ffs(band) - 1;
            band &= band - 1u;
            int aci = a[0];
#pragma unroll
            for (int c = 1; c < 8; c++) aci = ci == c ? a[c] : aci;
            const uint32_t lo = (uint32_t)min(aci, b), hi = (uint32_t)max(aci, b);
            const double ddx = min_image(x[hi] - x[lo], bp.box[0]);
            const double ddy = min_image(y[hi] - y[lo], bp.box[1]);
            const double ddz = min_image(z[hi] - z[lo], bp.box[2]);
            const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const bool in = r2 <= bp.r_skin2, cut = r2 <= bp.r_cut2;
            inb = (inb & ~(1u << ci)) | (in ? (1u << ci) : 0u);
            cutb = (cutb & ~(1u << ci)) | ((in && cut) ? (1u << ci) : 0u);
        }
        inb &= ~xhit;
        cutb &= ~xhit;
        far &= inb;
        while (far) {
            // the register arrays are read through selects, never by a variable index
            const int ci = __ffs(far) - 1;
            far &= far - 1u;
            int aci = a[0];
#pragma unroll
            for (int c = 1; c < 8; c++) aci = ci == c ? a[c] : aci;
            const uint32_t lo = (uint32_t)min(aci, b), hi = (uint32_t)max(aci, b);
            const uint2 rg = excl_rng[lo];
            if (hi <= rg.y && excluded(excl_off, excl_nbr, lo, hi)) {
                inb &= ~(1u << ci);
                cutb &= ~(1u << ci);
            }
        }
        uint32_t newmask = 0;
        unsigned long long npairs = 0, ncut_tiles = 0, ncut_pairs = 0;
        uint32_t dlo = 0, dhi = 0;
        // lane ci collects tile ci's pair mask: one 64-byte store per entry (a
        // store per tile would wait for the previous one to free its registers)
        unsigned long long mlane = 0;
        // lane ci also collects the smallest r2 estimate of tile ci's listed
        // pairs (f32 bits of non-negative values order as integers; +inf for
        // a tile without one); the distance byte is a non-decreasing function
        // of r2, so converting the minimum gives the minimum byte
        int r2min = 0x7f800000;
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            const unsigned long long m = __ballot((inb >> ci) & 1u);
            const unsigned long long mc = __ballot((cutb >> ci) & 1u);
            mlane = t == ci ? m : mlane;
            if (m) { newmask |= 1u << ci; npairs += __popcll(m); }
            if (mc) ncut_tiles++;
            ncut_pairs += __popcll(mc);
            const int rb = wave_min_i32(((inb >> ci) & 1u) ? __float_as_int(r2v[ci]) : 0x7f800000);
            r2min = t == ci ? rb : r2min;
        }
        int db = 255;
        if (r2min != 0x7f800000) {
            const float dl = sqrtf(__int_as_float(r2min)) * (1.0f - 2e-6f) - 1e-5f;
            const float st = floorf((dl - bp.prune_rc) * bp.prune_inv_step);
            db = st <= 0.f ? 0 : (st >= 255.f ? 255 : (int)st);
        }
#pragma unroll
        for (int ci = 0; ci < 4; ci++) {
            dlo |= (uint32_t)__builtin_amdgcn_readlane(db, ci) << (8 * ci);
            dhi |= (uint32_t)__builtin_amdgcn_readlane(db, ci + 4) << (8 * ci);
        }
        ent_lb[(size_t)e * 64 + t] = (uint8_t)inb;
        if (t < 8) ent_mask[(size_t)e * 8 + t] = mlane;
        if (t == 0) {
            ent_imask[e] = newmask;
            ent_meta[e] = make_int2(cj, (int)newmask);
            ent_dmin[e] = make_uint2(dlo, dhi);
        }
        tot_pairs += npairs;
        tot_tiles += (unsigned long long)__popc(newmask);
        tot_ctiles += ncut_tiles;
        tot_cpairs += ncut_pairs;
    }
    // per-wavefront partial counts, summed by k_sum_parts: same-address
    // atomics from every wavefront would serialise (about 10 ns each)
    if (t == 0) {
        unsigned long long *pw = parts + ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 4;
        pw[0] = tot_pairs;
        pw[1] = tot_tiles;
        pw[2] = tot_ctiles;
        pw[3] = tot_cpairs;
    }
}

// out[c] = sum over the n partial records of parts[4 k + c], in a fixed order.
__global__ __launch_bounds__(1024) void k_sum_parts(const unsigned long long *parts, int n,
                                                   unsigned long long *out) {
    __shared__ unsigned long long sm[4][1024];
    const int t = threadIdx.x;
    unsigned long long a[4] = {0, 0, 0, 0};
    for (int k = t; k < n; k += 1024)
        for (int c = 0; c < 4; c++) a[c] += parts[(size_t)k * 4 + c];
    for (int c = 0; c < 4; c++) sm[c][t] = a[c];
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (t < off)
            for (int c = 0; c < 4; c++) sm[c][t] += sm[c][t + off];
        __syncthreads();
    }
    if (t < 4) out[t] = sm[t][0];
}

// Clears the rebuild test's words: flag[0] (rebuild) = 0, flag[1] (first
// atom with a non-finite position) = INT_MAX, and copies the first atom with
// a non-finite force of the last evaluation into flag[2].
__global__ void k_rebuild_words(int *flag, const int *bad_first) {
    flag[0] = 0;
    flag[1] = 0x7fffffff;
    flag[2] = *bad_first;
}

// Exclusive scan of n ints by one 1024-thread block (in place of the host
// loop): out[k] = sum of in[0..k), out[n] = total, returned to every thread.
// f(k) gives the k-th input (so a count can be transformed on the fly).
template <typename F>
__device__ inline int block_exclusive_scan(int n, F f, int *out, int *sh) {
    const int t = threadIdx.x;
    int carry = 0;
    for (int base = 0; base < n; base += 1024) {
        const int k = base + t;
        const int v = k < n ? f(k) : 0;
        // inclusive scan within the block: wavefront scan by DPP-free shuffles, then the 16 wave totals
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if ((t & 63) >= o) x += y;
        }
        if ((t & 63) == 63) sh[t >> 6] = x;
        __syncthreads();
        if (t < 16) {
            int w = sh[t];
            for (int o = 1; o < 16; o <<= 1) {
                const int y = __shfl_up(w, o, 16);
                if ((t & 15) >= o) w += y;
            }
            sh[16 + t] = w;  // inclusive over waves
        }
        __syncthreads();
        const int wave_base = (t >> 6) ? sh[16 + (t >> 6) - 1] : 0;
        if (k < n) out[k] = carry + wave_base + x - v;
        carry += sh[16 + 15];
        __syncthreads();  // sh is rewritten by the next tile
    }
    if (t == 0) out[n] = carry;
    return carry;
}

// One block: entry offsets per supercluster (sc_start, exclusive scan of the
// pair-search counts), the total entry count and whether it exceeds the entry
// buffers' capacity, the work shape (entries per work item: chunk_req, or for
// 0 about 1.5 items per resident wavefront slot, in whole blocks of kBlock,
// 8..kMaxRun), and per supercluster its first work item and first block.
// Also clears the bins of the work-item sort.
__global__ __launch_bounds__(1024) void k_sc_scan(int *meta, const int *sc_count, int *sc_start, int *item_off,
                                                  int *blk_off, int ent_cap, int chunk_req, int slots,
                                                  int *bins) {
    __shared__ int sh[32];
    __shared__ int s_chunk;
    const int t = threadIdx.x;
    for (int k = t; k < 2 * kCostBins; k += 1024) bins[k] = 0;
    const int nsc = meta[1] ? 0 : meta[0];
    const int nent = block_exclusive_scan(nsc, [&](int k) { return sc_count[k]; }, sc_start, sh);
    if (t == 0) {
        meta[kMetaNent] = nent;
        meta[kMetaEntOver] = nent > ent_cap ? 1 : 0;
        int c = chunk_req;
        if (c <= 0) c = (int)((2LL * nent + 3LL * slots - 1) / (3LL * slots));
        c = min(max(kBlock, (c + kBlock - 1) / kBlock * kBlock), kMaxRun);
        s_chunk = c;
        meta[kMetaChunk] = c;
    }
    __syncthreads();
    const int c = s_chunk;
    const int nwi = block_exclusive_scan(nsc, [&](int k) { return (sc_count[k] + c - 1) / c; }, item_off, sh);
    const int nblk = block_exclusive_scan(nsc, [&](int k) { return (sc_count[k] + kBlock - 1) / kBlock; }, blk_off, sh);
    if (t == 0) {
        meta[kMetaNwi] = nwi;
        meta[kMetaNblk] = nblk;
    }
}

// One thread per work item: a run of `chunk` consecutive entries of one
// supercluster (whole blocks, counted from its first entry), with its first
// block and the j-cluster of its first entry, and its estimated cost in
// half-tile units (tiles plus 0.5 per entry plus 4 per item).
__global__ void k_wi_make(const int *meta, const int *sc_start, const int *item_off, const int *blk_off,
                          const uint32_t *ent_imask, const int *ent_cj, int4 *wi_tmp, int *cost,
                          int *bins) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (meta[kMetaEntOver] || it >= meta[kMetaNwi]) return;
    const int nsc = meta[1] ? 0 : meta[0];
    const int chunk = meta[kMetaChunk];
    int lo = 0, hi = nsc - 1;  // the last supercluster whose first item is <= it
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (item_off[mid] <= it) lo = mid; else hi = mid - 1;
    }
    const int sc = lo, k = it - item_off[sc];
    const int e0 = sc_start[sc] + k * chunk, e1 = min(e0 + chunk, sc_start[sc + 1]);
    int c = 8;
    for (int e = e0; e < e1; e++) c += 2 * __popc(ent_imask[e] & 0xffu) + 1;
    wi_tmp[it] = make_int4(sc | ((e1 - e0) << kWiEntryShift), e0, blk_off[sc] + k * (chunk / kBlock), ent_cj[e0]);
    cost[it] = c;
    atomicAdd(&bins[c], 1);
}

// One block: bin offsets in decreasing cost order (heaviest items first).
__global__ __launch_bounds__(1024) void k_cost_scan(int *bins) {
    __shared__ int sh[32];
    // bins[0..kCostBins) counts -> bins[kCostBins + b] = items costlier than b
    block_exclusive_scan(kCostBins, [&](int k) { return bins[kCostBins - 1 - k]; }, bins + kCostBins, sh);
    __syncthreads();
    // reversed offsets: start of cost b = bins[kCostBins + (kCostBins - 1 - b)]; keep starts in bins[0..)
    const int t = threadIdx.x;
    int v[(kCostBins + 1023) / 1024];
    for (int r = 0; r * 1024 + t < kCostBins; r++) v[r] = bins[kCostBins + (kCostBins - 1 - (r * 1024 + t))];
    __syncthreads();
    for (int r = 0; r * 1024 + t < kCostBins; r++) bins[r * 1024 + t] = v[r];
}

// Items to their place in the heaviest-first order (the order within one cost
// is whatever the atomics give: results do not depend on the item order).
__global__ void k_wi_scatter(const int *meta, const int4 *wi_tmp, const int *cost, int *bins, int4 *wi) {
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (meta[kMetaEntOver] || it >= meta[kMetaNwi]) return;
    const int pos = atomicAdd(&bins[cost[it]], 1);
    wi[pos] = wi_tmp[it];
}

// neighbor.rs:60-72, same f64 expression; any thread past the threshold sets
// flag[0].  A position that is not finite (never past the threshold, as in
// the reference) is raised into flag[1] (atomic minimum of the atom index):
// the run loop's check_positions_finite (geodesic/src/lib.rs:94-107).
__global__ void k_needs_rebuild(int n, const double *x, const double *y, const double *z,
                                const double *rx, const double *ry, const double *rz,
                                double threshold_sq, int *flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i], yi = y[i], zi = z[i];
    double dx = xi - rx[i];
    double dy = yi - ry[i];
    double dz = zi - rz[i];
    if (dx * dx + dy * dy + dz * dz > threshold_sq) flag[0] = 1;
    if (!(isfinite(xi) && isfinite(yi) && isfinite(zi))) atomicMin(&flag[1], i);
}

inline unsigned grid_for(size_t n, unsigned block) { return (unsigned)((n + block - 1) / block); }

// Bounding box of the positions (one block): out[0..2] min, out[3..5] max,
// out[6] = 1 if a coordinate is not finite.
__global__ __launch_bounds__(1024) void k_extent(int n, const double *x, const double *y, const double *z,
                                                double *out) {
    __shared__ double sm[6][1024];
    __shared__ int bad;
    const int t = threadIdx.x;
    if (t == 0) bad = 0;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool nf = false;
    for (int i = t; i < n; i += 1024) {
        const double v[3] = {x[i], y[i], z[i]};
        for (int a = 0; a < 3; a++) {
            nf = nf || !isfinite(v[a]);
            lo[a] = fmin(lo[a], v[a]);
            hi[a] = fmax(hi[a], v[a]);
        }
    }
    for (int a = 0; a < 3; a++) {
        sm[a][t] = lo[a];
        sm[3 + a][t] = hi[a];
    }
    __syncthreads();
    if (nf) bad = 1;
    for (int off = 512; off > 0; off >>= 1) {
        if (t < off)
            for (int a = 0; a < 3; a++) {
                sm[a][t] = fmin(sm[a][t], sm[a][t + off]);
                sm[3 + a][t] = fmax(sm[3 + a][t], sm[3 + a][t + off]);
            }
        __syncthreads();
    }
    if (t < 6) out[t] = sm[t][0];
    if (t == 6) out[6] = bad ? 1.0 : 0.0;
    if (t == 7) out[7] = 0.0;
}

// sigma/2 in the force kernel's length unit: fixed-point steps for cubic boxes, else A
inline float kernel_sigma_scale(const gm_sys *s) {
    return span_is_cubic(s) ? s->nbp.inv_s : 1.0f;
}

}  // namespace

namespace gm {

hipError_t launch_fill_i32(int *p, size_t n, int v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill_i32<<<grid_for(n, 256), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill_u64<<<grid_for(n, 256), 256, 0, st>>>(p, n);
    return hipGetLastError();
}

// Wrap/quantise, bin, sort, split and box.  meta[0] (supercluster count),
// meta[1] (column overflow) are valid once the stream reaches this point.
hipError_t launch_extent(gm_sys *s, double *out) {
    k_extent<<<1, 1024, 0, s->stream>>>((int)s->n, s->x.p, s->y.p, s->z.p, out);
    return hipGetLastError();
}

hipError_t launch_build(gm_sys *s, bool wrap) {
    hipStream_t st = s->stream;
    const int n = (int)s->n;
    const BuildParams &bp = s->bp;
    hipError_t e;
    if ((e = launch_fill_i32(s->col_count.p, bp.ncol, 0, st)) != hipSuccess) return e;
    if ((e = launch_fill_i32(s->col_fill.p, bp.ncol, 0, st)) != hipSuccess) return e;
    if ((e = launch_fill_i32(s->meta.p, 8, 0, st)) != hipSuccess) return e;
    if (n > 0) {
        if (bp.ncol <= kLdsColumns)
            k_wrap_quantize_lds<<<grid_for(n, 1024), 1024, sizeof(int) * (size_t)bp.ncol, st>>>(
                n, s->x.p, s->y.p, s->z.p, s->refx.p, s->refy.p, s->refz.p, s->qatom.p, s->col_count.p, bp,
                wrap ? 1 : 0);
        else
            k_wrap_quantize<<<grid_for(n, 256), 256, 0, st>>>(n, s->x.p, s->y.p, s->z.p, s->refx.p,
                                                                s->refy.p, s->refz.p, s->qatom.p,
                                                                s->col_count.p, bp, wrap ? 1 : 0);
    }
    k_scan_columns<<<1, 1024, 0, st>>>(bp.ncol, s->col_count.p, s->col_start.p, s->col_sc.p,
                                       s->meta.p, kColumnCap);
    if (n > 0) {
        if (bp.ncol <= kLdsColumns)
            k_scatter_lds<<<grid_for(n, 1024), 1024, sizeof(int) * (size_t)bp.ncol, st>>>(
                n, s->qatom.p, s->col_start.p, s->col_fill.p, s->bucket.p, bp);
        else
            k_scatter<<<grid_for(n, 256), 256, 0, st>>>(n, s->qatom.p, s->col_start.p, s->col_fill.p,
                                                          s->bucket.p, bp);
    }
    k_column_sort<<<bp.ncol, 1024, kColumnCap * sizeof(unsigned long long), st>>>(
        n, s->qatom.p, s->bucket.p, s->col_start.p, s->col_count.p, s->col_sc.p,
        s->sorted_atoms.p, s->sc_base.p, s->sc_m.p, s->sc_col.p);
    if ((e = launch_fill_i32(s->slot_atom.p, (size_t)s->max_sc * 64, -1, st)) != hipSuccess) return e;
    k_sc_split<<<(unsigned)s->max_sc, 64, 0, st>>>(s->meta.p, n, s->qatom.p, s->sorted_atoms.p,
                                                   s->sc_base.p, s->sc_m.p, s->slot_atom.p);
    k_slot_setup<<<grid_for((size_t)s->max_sc * 64, 256), 256, 0, st>>>(
        (int)(s->max_sc * 64), n, s->slot_atom.p, s->qatom.p, s->atom_par.p, s->excl_bits.p,
        s->atom_slot.p, s->sq.p, s->sp.p, s->qpos.p, s->qeps.p, s->sxb.p, kernel_sigma_scale(s), s->x.p, s->y.p,
        s->z.p);
    k_bbox<<<(unsigned)s->max_sc, 64, 0, st>>>(s->meta.p, s->slot_atom.p, s->sq.p, bp, s->cl_ref.p,
                                               s->cl_half.p, s->sc_ref.p, s->sc_half.p);
    return hipGetLastError();
}

// Grids of the list kernels cover the capacities (the supercluster bound
// max_sc, the entry buffers' ent_cap); each kernel reads the actual counts
// from meta and leaves the rest alone, so the build needs no host round trip
// between its kernels.
hipError_t launch_pair_count(gm_sys *s) {
    if (s->max_sc == 0) return hipSuccess;
    k_pair_search<0><<<(unsigned)s->max_sc, 64, 0, s->stream>>>(
        s->meta.p, s->bp, s->sc_col.p, s->col_sc.p, s->col_count.p, s->cl_ref.p, s->cl_half.p,
        s->sc_ref.p, s->sc_half.p, s->sc_count.p, nullptr, nullptr, nullptr, nullptr, 0);
    return hipGetLastError();
}

hipError_t launch_sc_scan(gm_sys *s, int64_t ent_cap, int chunk_req) {
    const int slots = 20 * (s->ctx->n_cu > 0 ? s->ctx->n_cu : 256);
    k_sc_scan<<<1, 1024, 0, s->stream>>>(s->meta.p, s->sc_count.p, s->sc_start.p, s->item_off.p, s->sc_blk_off.p,
                                         (int)std::min<int64_t>(ent_cap, 0x7fffffff), chunk_req, slots,
                                         s->cost_bins.p);
    return hipGetLastError();
}

hipError_t launch_pair_fill_and_masks(gm_sys *s) {
    hipStream_t st = s->stream;
    if (s->max_sc > 0)
        k_pair_search<1><<<(unsigned)s->max_sc, 64, 0, st>>>(
            s->meta.p, s->bp, s->sc_col.p, s->col_sc.p, s->col_count.p, s->cl_ref.p, s->cl_half.p,
            s->sc_ref.p, s->sc_half.p, nullptr, s->sc_start.p, s->ent_sc.p, s->ent_cj.p,
            s->ent_imask.p, 1);
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((s->ent_cap + 3) / 4, kMaskBlocks));
    hipError_t e = s->mask_parts.ensure((size_t)blocks * 4 * 4);  // 4 counts per wavefront
    if (e != hipSuccess) return e;
    s->mask_waves = blocks * 4;
    k_masks<<<(unsigned)blocks, 256, 0, st>>>(s->meta.p, s->ent_sc.p, s->ent_cj.p, s->ent_imask.p,
                                              s->ent_mask.p, s->ent_meta.p, s->ent_dmin.p, s->ent_lb.p,
                                              s->slot_atom.p, s->sq.p, s->x.p, s->y.p, s->z.p,
                                              s->excl_off.p, s->excl_nbr.p, s->excl_rng.p, s->sxb.p,
                                              s->bp, s->mask_parts.p);
    return hipGetLastError();
}

hipError_t launch_work_items(gm_sys *s) {
    hipStream_t st = s->stream;
    const unsigned g = (unsigned)std::max<int64_t>(1, (s->item_cap + 255) / 256);
    k_wi_make<<<g, 256, 0, st>>>(s->meta.p, s->sc_start.p, s->item_off.p, s->sc_blk_off.p, s->ent_imask.p,
                                 s->ent_cj.p, s->wi_tmp.p, s->wi_cost.p, s->cost_bins.p);
    k_cost_scan<<<1, 1024, 0, st>>>(s->cost_bins.p);
    k_wi_scatter<<<g, 256, 0, st>>>(s->meta.p, s->wi_tmp.p, s->wi_cost.p, s->cost_bins.p, s->wi.p);
    return hipGetLastError();
}

hipError_t launch_list_stats(gm_sys *s) {
    unsigned long long *counters = (unsigned long long *)(s->meta.p + 8);
    k_sum_parts<<<1, 1024, 0, s->stream>>>(s->mask_parts.p, s->mask_waves, counters);
    return hipGetLastError();
}

hipError_t launch_needs_rebuild(gm_sys *s, const double *x, const double *y, const double *z, int *flag) {
    const double half = (s->rskin - s->rc) / 2.0;
    const double thr = half * half;
    if (!flag) flag = s->meta.p + 5;
    k_rebuild_words<<<1, 1, 0, s->stream>>>(flag, bad_first_word(s));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (s->n > 0)
        k_needs_rebuild<<<grid_for(s->n, 256), 256, 0, s->stream>>>(
            (int)s->n, x, y, z, s->refx.p, s->refy.p, s->refz.p, thr, flag);
    return hipGetLastError();
}

}  // namespace gm
