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
// The positions are read from the caller's arrays (px, py, pz), wrapped in
// place there, and copied into the system's own (x, y, z).
__global__ void k_wrap_quantize(int n, double *px, double *py, double *pz, double *x, double *y, double *z,
                                double *rx, double *ry, double *rz, uint32_t *q, int *col_count, BuildParams bp,
                                int wrap) {
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i0 < n;
    const int i = active ? i0 : 0;
    int col = 0;
    if (active) {
        double xi = px[i], yi = py[i], zi = pz[i];
        if (wrap) {
            xi = wrap_coord(xi, bp.box[0]);
            yi = wrap_coord(yi, bp.box[1]);
            zi = wrap_coord(zi, bp.box[2]);
            px[i] = xi; py[i] = yi; pz[i] = zi;
        }
        x[i] = xi; y[i] = yi; z[i] = zi;
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

// Exclusive scans over columns (one block of 1024): atom offsets and
// supercluster offsets; total supercluster count, column overflow flag and
// largest column into meta.  Also clears the scatter's column fill counters.
__device__ inline void scan_columns_block(int ncol, const int *cnt, int *start, int *colsc, int *meta, int cap,
                                          int *col_fill) {
    __shared__ int s1[1024], s2[1024], sm[1024];
    int t = threadIdx.x;
    for (int i = t; i < ncol; i += 1024) col_fill[i] = 0;
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
        colsc[ncol] = s2[1023];  // (column c's superclusters: colsc[c] .. colsc[c + 1])
        meta[0] = s2[1023];
        meta[1] = sm[1023] > cap ? 1 : 0;
        meta[2] = sm[1023];
    }
}

__global__ __launch_bounds__(1024) void k_scan_columns(int ncol, const int *cnt, int *start, int *colsc, int *meta,
                                                      int cap, int *col_fill) {
    scan_columns_block(ncol, cnt, start, colsc, meta, cap, col_fill);
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
__global__ __launch_bounds__(1024) void k_wrap_quantize_lds(int n, double *px, double *py, double *pz, double *x,
                                                           double *y, double *z, double *rx, double *ry, double *rz,
                                                           uint32_t *q, int *col_count, BuildParams bp, int wrap) {
    extern __shared__ int hist[];
    const int t = threadIdx.x, i = blockIdx.x * 1024 + t;
    for (int c = t; c < bp.ncol; c += 1024) hist[c] = 0;
    __syncthreads();
    if (i < n) {
        double xi = px[i], yi = py[i], zi = pz[i];
        if (wrap) {
            xi = wrap_coord(xi, bp.box[0]);
            yi = wrap_coord(yi, bp.box[1]);
            zi = wrap_coord(zi, bp.box[2]);
            px[i] = xi; py[i] = yi; pz[i] = zi;
        }
        x[i] = xi; y[i] = yi; z[i] = zi;
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

// One wavefront per supercluster slot range (all max_sc of them):
//  - split: a full supercluster (64 atoms, z-sorted) into 2x2x2 clusters of
//    8: halves by x, quarters by y, eighths by z (ties broken by atom index);
//    a partial one keeps z order and is cut into consecutive eights;
//  - slot tables: qpos/qeps (the force kernel's per-slot inputs) are filled
//    for every slot, padding included, so the per-evaluation prep kernel only
//    needs to overwrite the atoms' slots; sq.w = 1 marks an atom whose
//    position is not finite: such an atom is in no listed pair (the
//    reference's r2 <= r_skin^2 is false for a NaN distance);
//  - bounding boxes (fixed-point centre, half extents in A) of the 8 clusters
//    and the supercluster.  A cluster is valid iff its first slot holds an
//    atom.  Offsets are taken relative to the group's first atom (minimum
//    image), so min <= 0 <= max; spans can exceed half the box in small
//    boxes, hence computed in float;
//  - the fixed-point force accumulator's words of these slots are cleared
//    (the slot count changed: the next evaluation starts from zero).
// Superclusters past the count (or all, after a column overflow) get padding.
__global__ __launch_bounds__(64) void k_sc_setup(const int *meta, int n, int max_sc, const uint32_t *q,
                                                 const int *sorted_atoms, const int *sc_base, const int *sc_m,
                                                 int *slot_atom, const float2 *atom_par, const uint4 *excl_bits,
                                                 int *atom_slot, int4 *sq, float2 *sp, int4 *qpos, float *qeps,
                                                 uint4 *sxb, float inv_s, const double *x, const double *y,
                                                 const double *z, BuildParams bp, int4 *cl_ref, float4 *cl_half,
                                                 int4 *sc_ref, float4 *sc_half, unsigned long long *F) {
    __shared__ unsigned long long key[64];
    __shared__ int4 lsq[64];
    const int sc = blockIdx.x;
    const int t = threadIdx.x;
    const size_t npad = (size_t)max_sc * 64;
    F[(size_t)sc * 64 + t] = 0ull;
    F[npad + (size_t)sc * 64 + t] = 0ull;
    F[2 * npad + (size_t)sc * 64 + t] = 0ull;
    const bool live = !meta[1] && sc < meta[0];  // (a column overflowed: sc_base is not set)
    int a = -1, pos = t;
    if (live) {
        const int m = sc_m[sc];
        a = t < m ? sorted_atoms[sc_base[sc] + t] : -1;
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
    }
    const size_t s = (size_t)sc * 64 + pos;  // this thread's atom's slot
    slot_atom[s] = a;
    int4 v = make_int4(0, 0, 0, 0);
    if (a >= 0) {
        atom_slot[a] = (int)s;
        const bool fin = isfinite(x[a]) && isfinite(y[a]) && isfinite(z[a]);
        v = make_int4((int)q[a], (int)q[n + a], (int)q[2 * n + a], fin ? 0 : 1);
        const float2 p = atom_par[a];
        sq[s] = v;
        sp[s] = p;
        qpos[s] = make_int4(v.x, v.y, v.z, __float_as_int(p.x * inv_s));
        qeps[s] = 2.0f * p.y;
        sxb[s] = excl_bits[a];
    } else {
        sq[s] = v;
        sp[s] = make_float2(0.f, 0.f);
        qpos[s] = v;
        qeps[s] = 0.f;
        sxb[s] = make_uint4(0u, 0u, 0u, 0u);
    }
    if (!live) return;  // the same for the whole block
    // bounding boxes over the slots, from LDS (slot t's values; w: 1 = atom)
    lsq[pos] = make_int4(v.x, v.y, v.z, a >= 0 ? 1 : 0);
    __syncthreads();
    const int4 vs = lsq[t];
    const bool valid = vs.w != 0;
    const int4 r = lsq[t & ~7];
    const bool cvalid = r.w != 0;
    int dx = valid ? wrapdiff(vs.x, r.x) : 0, dy = valid ? wrapdiff(vs.y, r.y) : 0,
        dz = valid ? wrapdiff(vs.z, r.z) : 0;
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
    const int4 r0 = lsq[0];
    dx = valid ? wrapdiff(vs.x, r0.x) : 0;
    dy = valid ? wrapdiff(vs.y, r0.y) : 0;
    dz = valid ? wrapdiff(vs.z, r0.z) : 0;
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

// Exclusive scan of n ints by one block of NT threads: out[k] = sum of
// in[0..k), out[n] = total, returned to every thread.  Thread t supplies
// input base + t of each tile of NT as f(base + t) (so a count can be
// transformed on the fly).  sh: 2 * (NT / 64) ints of LDS.
template <int NT, typename F>
__device__ inline int block_exclusive_scan(int n, F f, int *out, int *sh) {
    constexpr int NW = NT / 64;
    const int t = threadIdx.x;
    int carry = 0;
    for (int base = 0; base < n; base += NT) {
        const int k = base + t;
        const int v = k < n ? f(k) : 0;
        // inclusive scan within each wavefront, then over the wave totals
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if ((t & 63) >= o) x += y;
        }
        if ((t & 63) == 63) sh[t >> 6] = x;
        __syncthreads();
        if (t < NW) {
            int w = sh[t];
            for (int o = 1; o < NW; o <<= 1) {
                const int y = __shfl_up(w, o, NW);
                if (t >= o) w += y;
            }
            sh[NW + t] = w;  // inclusive over waves
        }
        __syncthreads();
        const int wave_base = (t >> 6) ? sh[NW + (t >> 6) - 1] : 0;
        if (k < n) out[k] = carry + wave_base + x - v;
        carry += sh[2 * NW - 1];
        __syncthreads();  // sh is rewritten by the next tile
    }
    if (t == 0) out[n] = carry;
    return carry;
}

// Arguments of the per-supercluster scans.
struct ScScanArgs {
    int *sc_start, *item_off, *blk_off, *bins;
    int ent_cap, chunk_req, slots;
};

// One block of NT: entry offsets per supercluster (sc_start, exclusive scan of
// the pair-search counts), the total entry count and whether it exceeds the
// entry buffers' capacity, the work shape (entries per work item: chunk_req,
// or for 0 about 1.5 items per resident wavefront slot, in whole blocks of
// kBlock, 8..kMaxRun), and per supercluster its first work item and first
// block.  Also clears the bins of the work-item sort.
template <int NT>
__device__ inline void sc_scan_block(int *meta, const int *sc_count, const ScScanArgs &A) {
    __shared__ int sh[2 * (NT / 64)];
    __shared__ int s_chunk;
    const int t = threadIdx.x;
    for (int k = t; k < 2 * kCostBins; k += NT) A.bins[k] = 0;
    const int nsc = meta[1] ? 0 : meta[0];
    const int nent = block_exclusive_scan<NT>(nsc, [&](int k) { return sc_count[k]; }, A.sc_start, sh);
    if (t == 0) {
        meta[kMetaNent] = nent;
        meta[kMetaEntOver] = nent > A.ent_cap ? 1 : 0;
        int c = A.chunk_req;
        if (c <= 0) c = (int)((2LL * nent + 3LL * A.slots - 1) / (3LL * A.slots));
        c = min(max(kBlock, (c + kBlock - 1) / kBlock * kBlock), kMaxRun);
        s_chunk = c;
        meta[kMetaChunk] = c;
    }
    __syncthreads();
    const int c = s_chunk;
    const int nwi = block_exclusive_scan<NT>(nsc, [&](int k) { return (sc_count[k] + c - 1) / c; }, A.item_off, sh);
    const int nblk =
        block_exclusive_scan<NT>(nsc, [&](int k) { return (sc_count[k] + kBlock - 1) / kBlock; }, A.blk_off, sh);
    if (t == 0) {
        meta[kMetaNwi] = nwi;
        meta[kMetaNblk] = nblk;
    }
}

__global__ __launch_bounds__(1024) void k_sc_scan(int *meta, const int *sc_count, ScScanArgs A) {
    sc_scan_block<1024>(meta, sc_count, A);
}

// The ix-th column of at most two ascending pieces of a ring.
__device__ inline int piece_at(int2 p0, int2 p1, int ix) {
    const int n0 = p0.y - p0.x + 1;
    return ix < n0 ? p0.x + ix : p1.x + (ix - n0);
}

// One workgroup of 256 per supercluster: test the j-clusters cj >= its first
// cluster that lie in xy columns within reach of its own column (columns are
// periodic in the fixed-point frame, so every pair within r_skin is in
// reach).  An entry is (sc, cj) with a nonzero 8-bit i-cluster mask (ci <= cj,
// both valid, boxes within r_skin).  Columns are taken in ascending index,
// 256 at a time; their candidate clusters are laid end to end (a scan of the
// per-column counts) and tested 256 per round, one per thread, so a round
// waits for one load of cluster boxes.  Clusters are numbered column by
// column, so entries come in ascending cj.  Pass 0 counts, pass 1 writes
// entries at sc_start[sc], each round's in thread order.
template <int FILL, bool FULL = false>
__device__ inline int pair_search_sc(const int *meta, const BuildParams &bp, const int *sc_col,
                                     const int *col_sc, const int4 *cl_ref,
                                     const float4 *cl_half, const int4 *sc_ref, const float4 *sc_half,
                                     const int *sc_start, int *ent_sc, int *ent_cj, uint32_t *ent_imask) {
    __shared__ int4 cir[8];
    __shared__ float4 cih[8];
    __shared__ int s_lo[256];       // first candidate cluster of each column of the chunk
    __shared__ int s_off[257];      // exclusive scan of the columns' candidate counts
    __shared__ int s_sh[8];
    __shared__ int s_wc[4];         // entries per wavefront in the round
    const int sc = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t < 8) { cir[t] = cl_ref[sc * 8 + t]; cih[t] = cl_half[sc * 8 + t]; }
    const int4 sr = sc_ref[sc];
    const float4 sh = sc_half[sc];
    const int first = sc * 8;
    const int base = FILL ? sc_start[sc] : 0;
    const int c0 = sc_col[sc];
    const int cx0 = c0 / bp.ncy, cy0 = c0 - (c0 / bp.ncy) * bp.ncy;
    int2 px0, px1, py0, py1;
    ring_pieces(cx0, bp.kx, bp.ncx, px0, px1);
    ring_pieces(cy0, bp.ky, bp.ncy, py0, py1);
    const int NX = (px0.y - px0.x + 1) + (px1.y - px1.x + 1);  // an absent piece is [0, -1]
    const int NY = (py0.y - py0.x + 1) + (py1.y - py1.x + 1);
    const int ncols = NX * NY;  // ascending (ix, iy) visits ascending column index
    const float rs2 = bp.r_search * bp.r_search;
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int count = 0;
    __syncthreads();  // cir, cih
    for (int k0 = 0; k0 < ncols; k0 += 256) {
        int lo = 0, n = 0;
        const int k = k0 + t;
        if (k < ncols) {
            const int ix = k / NY, iy = k - (k / NY) * NY;
            const int cx = piece_at(px0, px1, ix), cy = piece_at(py0, py1, iy);
            const float gx = column_gap(cx - cx0, bp.ncx, bp.wcol[0]);
            const float gy = column_gap(cy - cy0, bp.ncy, bp.wcol[1]);
            const int c = cx * bp.ncy + cy;
            if ((FULL || c >= c0) && gx * gx + gy * gy <= rs2) {
                lo = (c == c0 && !FULL) ? first : col_sc[c] * 8;
                n = max(col_sc[c + 1] * 8 - lo, 0);
            }
        }
        s_lo[t] = lo;
        const int ncand = block_exclusive_scan<256>(256, [&](int) { return n; }, s_off, s_sh);
        const int nc = min(256, ncols - k0);
        __syncthreads();  // s_lo, s_off
        for (int q0 = 0; q0 < ncand; q0 += 256) {
            const int q = q0 + t;
            uint32_t imask = 0;
            int cj = 0;
            if (q < ncand) {
                int a = 0, b = nc - 1;  // the last column whose candidates start at or before q
                while (a < b) {
                    const int mid = (a + b + 1) >> 1;
                    if (s_off[mid] <= q) a = mid; else b = mid - 1;
                }
                cj = s_lo[a] + (q - s_off[a]);
                const int4 jr = cl_ref[cj];
                const float4 jh = cl_half[cj];
                if (jr.w && boxes_close(sr, sh, jr, jh, bp)) {
                    for (int ci = 0; ci < 8; ci++) {
                        if ((FULL || first + ci <= cj) && cir[ci].w && boxes_close(cir[ci], cih[ci], jr, jh, bp))
                            imask |= 1u << ci;
                    }
                }
            }
            const unsigned long long bl = __ballot(imask != 0);
            if (lane == 0) s_wc[w] = __popcll(bl);
            __syncthreads();
            int woff = 0, rtot = 0;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int cv = s_wc[v];
                woff += v < w ? cv : 0;
                rtot += cv;
            }
            if (FILL && imask) {
                const int off = base + count + woff + __popcll(bl & lt_mask);
                ent_sc[off] = sc;
                ent_cj[off] = cj;
                if (!FULL) ent_imask[off] = imask;
            }
            count += rtot;
            __syncthreads();  // s_wc is rewritten by the next round
        }
    }
    return count;
}

// Pass 0 counts, pass 1 writes the entries.
__global__ __launch_bounds__(256) void k_pair_count(const int *meta, BuildParams bp, const int *sc_col,
                                                    const int *col_sc, const int4 *cl_ref,
                                                    const float4 *cl_half, const int4 *sc_ref,
                                                    const float4 *sc_half, int *sc_count) {
    const int sc = blockIdx.x;
    if (meta[1] || sc >= meta[0]) return;  // the same for the whole block
    const int count = pair_search_sc<0>(meta, bp, sc_col, col_sc, cl_ref, cl_half, sc_ref, sc_half, nullptr,
                                        nullptr, nullptr, nullptr);
    if (threadIdx.x == 0) sc_count[sc] = count;
}

__global__ __launch_bounds__(256) void k_pair_fill(const int *meta, BuildParams bp, const int *sc_col,
                                                   const int *col_sc, const int4 *cl_ref,
                                                   const float4 *cl_half, const int4 *sc_ref,
                                                   const float4 *sc_half, const int *sc_start, int *ent_sc,
                                                   int *ent_cj, uint32_t *ent_imask) {
    const int sc = blockIdx.x;
    if (meta[1] || sc >= meta[0] || meta[kMetaEntOver]) return;  // the same for the whole block
    (void)pair_search_sc<1>(meta, bp, sc_col, col_sc, cl_ref, cl_half, sc_ref, sc_half, sc_start,
                            ent_sc, ent_cj, ent_imask);
}

__device__ inline bool excluded(const int64_t *off, const uint32_t *nbr, uint32_t lo, uint32_t hi) {
    int64_t a = off[lo], b = off[lo + 1];
    while (a < b) {
        int64_t mid = (a + b) >> 1;
        if (nbr[mid] < hi) a = mid + 1; else b = mid;
    }
    return a < off[lo + 1] && nbr[a] == hi;
}

// Per tile ci (one value per tile in every lane), the minimum over the
// wavefront's 64 lanes, returned in the 8 lanes of the tile (lane >> 3 == ci).
// A transposed reduction: each exchange halves the tiles a lane carries, so
// the 8 minima cost about as much as 2 separate ones.
__device__ inline int wave_min8_i32(const int (&v)[8]) {
    const int lane = (int)__lane_id();
    // lanes L and L + 32: the lower half keeps tiles 0-3, the upper tiles 4-7
    int w[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const auto r = __builtin_amdgcn_permlane32_swap((unsigned)v[k], (unsigned)v[k + 4], false, false);
        w[k] = min((int)r[0], (int)r[1]);
    }
    // lanes L and L + 16 within each half: bit 4 clear keeps w[0..1], set w[2..3]
    int u[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const auto r = __builtin_amdgcn_permlane16_swap((unsigned)w[k], (unsigned)w[k + 2], false, false);
        u[k] = min((int)r[0], (int)r[1]);
    }
    // lanes L and L ^ 8: bit 3 clear keeps u[0], set u[1]
    const bool b3 = (lane & 8) != 0;
    const int keep = b3 ? u[1] : u[0], send = b3 ? u[0] : u[1];
    int x = min(keep, __builtin_amdgcn_update_dpp(0, send, 0x128, 0xF, 0xF, false));  // row_ror:8
    // the 8 lanes of the tile
    x = min(x, __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    x = min(x, __builtin_amdgcn_update_dpp(0, x, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    x = min(x, __builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, false));  // row_half_mirror
    return x;
}

// Bitwise OR over the wavefront's 64 lanes, in every lane.
__device__ inline int wave_or_i32(int v) {
    v |= __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    v |= __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    v |= __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
    v |= __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);
    const auto r = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    v = (int)(r[0] | r[1]);
    const auto q = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return (int)(q[0] | q[1]);
}

// One wavefront per entry (a strided walk over an XCD's range of entries,
// four wavefronts per workgroup): exact pair masks.  Lane = (jl = lane>>3, il = lane&7).
// The inclusion test is neighbor.rs:31-38: min-image r2 <= r_skin^2 in f64 on
// the wrapped positions, with i = min(atom), j = max(atom), then the exclusion
// set (neighbor.rs:33-35).  r2 is first estimated in f32 from the exact
// fixed-point min-image displacement; a pair whose estimate lies within the
// band of the threshold is decided by the f64 expression itself, and the band
// is many times the estimate's worst error, so every decision is the f64 one.
// Exclusions: a partner within 64 indices above the lower atom is a bit of
// that atom's mask; a partner farther up is looked up (binary search in the
// lower atom's sorted partners) only for pairs inside r_skin whose upper atom
// is at or below the lower atom's largest partner.
// Per tile, the byte floor((d_lo - r_cutoff) / step) (0 if negative, 255 for
// a tile without listed pairs) goes to ent_dmin, d_lo a lower bound on the
// distance of the tile's closest listed pair (the f32 estimate less a
// relative 2e-6 and 1e-5 A, far above its error).  The list statistics are
// counted separately, on demand (k_list_counts).
__global__ __launch_bounds__(256) void k_masks(const int *__restrict__ meta, const int *__restrict__ ent_sc,
                                               const int *__restrict__ ent_cj, uint32_t *__restrict__ ent_imask,
                                               int2 *__restrict__ ent_meta, uint2 *__restrict__ ent_dmin,
                                               uint8_t *__restrict__ ent_lb, const int *__restrict__ slot_atom,
                                               const int4 *__restrict__ sq, const double *__restrict__ x,
                                               const double *__restrict__ y, const double *__restrict__ z,
                                               const int64_t *__restrict__ excl_off,
                                               const uint32_t *__restrict__ excl_nbr,
                                               const uint4 *__restrict__ sxb, BuildParams bp,
                                               int *__restrict__ col_count) {
    // the column counters (last read by the column sort) start the next build
    // at zero
    for (int c = blockIdx.x * 256 + threadIdx.x; c < bp.ncol; c += gridDim.x * 256) col_count[c] = 0;
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
    const int l = __builtin_amdgcn_readfirstlane((int)(blockIdx.x / (unsigned)P) * 4 + (int)(threadIdx.x >> 6));
    const int e_lo = (int)((int64_t)nent * xg / P), e_hi = (int)((int64_t)nent * (xg + 1) / P);
    // The entry header (supercluster, j cluster, i-cluster mask) is the same in
    // every lane: it lives in scalar registers and is loaded one entry ahead.
    // The slot data of all 8 i clusters is loaded at once, whether the mask
    // lists a cluster or not (every slot of a supercluster exists), so that an
    // entry waits for one round of loads rather than one per cluster.
    int e = e_lo + l;
    int n_sc = 0, n_cj = 0;
    uint32_t n_im = 0;
    if (e < e_hi) { n_sc = ent_sc[e]; n_cj = ent_cj[e]; n_im = ent_imask[e]; }
    for (; e < e_hi; e += nwx) {
        const int sc = __builtin_amdgcn_readfirstlane(n_sc), cj = __builtin_amdgcn_readfirstlane(n_cj);
        const uint32_t im = (uint32_t)__builtin_amdgcn_readfirstlane((int)n_im);
        if (e + nwx < e_hi) { n_sc = ent_sc[e + nwx]; n_cj = ent_cj[e + nwx]; n_im = ent_imask[e + nwx]; }
        const size_t js = (size_t)cj * 8 + jl;
        const int4 qb = sq[js];
        const int bslot = slot_atom[js];
        const uint4 xbj = sxb[js];
        // per-cluster values in arrays of scalars (no vector structs), which
        // stay in registers once the loops are unrolled; the exclusion masks
        // of the slots' atoms arrive with their coordinates
        int a[8], qax[8], qay[8], qaz[8];
        uint32_t exl[8], exh[8], exf[8];
        const int4 *sqi = sq + (size_t)sc * 64 + il;
        const int *sai = slot_atom + (size_t)sc * 64 + il;
        const uint4 *sxi = sxb + (size_t)sc * 64 + il;
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            const int4 q = sqi[ci * 8];
            const int at = sai[ci * 8];
            const uint4 xb = sxi[ci * 8];
            a[ci] = (((im >> ci) & 1u) && q.w == 0) ? at : -1;
            qax[ci] = q.x;
            qay[ci] = q.y;
            qaz[ci] = q.z;
            exl[ci] = xb.x;
            exh[ci] = xb.y;
            exf[ci] = xb.z;
        }
        const int b = qb.w ? -1 : bslot;  // (a non-finite position is never listed)
        // exclusions, reduced at once to one bit per tile (the masks' registers
        // are then free): pairs within 64 indices (near) test the lower atom's
        // mask, which a tile does only if it has such a pair; a farther pair
        // up to the lower atom's largest partner (far) is looked up later if
        // it is listed
        uint32_t xhit = 0, far = 0;
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            const bool ilo = a[ci] < b;
            const int lo = ilo ? a[ci] : b, hi = ilo ? b : a[ci];
            const uint32_t k = (uint32_t)(hi - lo - 1);
            const bool valid = lo >= 0;  // both slots hold atoms
            const bool near = valid && k < 64u;
            if (valid && k >= 64u && (uint32_t)hi <= (ilo ? exf[ci] : xbj.z)) far |= 1u << ci;
            if (__ballot(near)) {
                const uint32_t w = k < 32u ? (ilo ? exl[ci] : xbj.x) : (ilo ? exh[ci] : xbj.y);
                if (near && ((w >> (k & 31u)) & 1u) != 0u) xhit |= 1u << ci;
            }
        }
        uint32_t inb = 0, band = 0;
        int r2v[8];  // f32 bits of the listed pairs' r2 estimates (+inf for the others)
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            const int cg = sc * 8 + ci;
            r2v[ci] = 0x7f800000;
            if (a[ci] >= 0 && b >= 0 && (cg < cj || il < jl)) {
                const float dx = (float)wrapdiff(qb.x, qax[ci]) * bp.sq[0];
                const float dy = (float)wrapdiff(qb.y, qay[ci]) * bp.sq[1];
                const float dz = (float)wrapdiff(qb.z, qaz[ci]) * bp.sq[2];
                const float r2f = dx * dx + dy * dy + dz * dz;
                r2v[ci] = __float_as_int(r2f);
                if (fabsf(r2f - bp.r_skin2f) <= bp.band_skin) band |= 1u << ci;
                if (r2f <= bp.r_skin2f) inb |= 1u << ci;
            }
        }
        // pairs near the threshold (rare): the f64 expression decides, one at
        // a time, outside the unrolled loop
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
            inb = (inb & ~(1u << ci)) | (r2 <= bp.r_skin2 ? (1u << ci) : 0u);
        }
        inb &= ~xhit;
        far &= inb;
        while (far) {
            // the register arrays are read through selects, never by a variable index
            const int ci = __ffs(far) - 1;
            far &= far - 1u;
            int aci = a[0];
#pragma unroll
            for (int c = 1; c < 8; c++) aci = ci == c ? a[c] : aci;
            const uint32_t lo = (uint32_t)min(aci, b), hi = (uint32_t)max(aci, b);
            if (excluded(excl_off, excl_nbr, lo, hi)) inb &= ~(1u << ci);
        }
        // per tile, the smallest r2 estimate of its listed pairs (f32 bits of
        // non-negative values order as integers; +inf for a tile without one),
        // in the tile's 8 lanes; the distance byte is a non-decreasing function
        // of r2, so converting the minimum gives the minimum byte
#pragma unroll
        for (int ci = 0; ci < 8; ci++) r2v[ci] = ((inb >> ci) & 1u) ? r2v[ci] : 0x7f800000;
        const int r2min = wave_min8_i32(r2v);
        int db = 255;
        if (r2min != 0x7f800000) {
            const float dl = sqrtf(__int_as_float(r2min)) * (1.0f - 2e-6f) - 1e-5f;
            const float st = floorf((dl - bp.prune_rc) * bp.prune_inv_step);
            db = st <= 0.f ? 0 : (st >= 255.f ? 255 : (int)st);
        }
        uint32_t dlo = 0, dhi = 0;
#pragma unroll
        for (int ci = 0; ci < 4; ci++) {
            dlo |= (uint32_t)__builtin_amdgcn_readlane(db, ci * 8) << (8 * ci);
            dhi |= (uint32_t)__builtin_amdgcn_readlane(db, ci * 8 + 32) << (8 * ci);
        }
        const uint32_t newmask = (uint32_t)__builtin_amdgcn_readfirstlane(wave_or_i32((int)inb));
        ent_lb[(size_t)e * 64 + t] = (uint8_t)inb;
        if (t == 0) {
            ent_imask[e] = newmask;
            ent_meta[e] = make_int2(cj, (int)newmask);
            ent_dmin[e] = make_uint2(dlo, dhi);
        }
    }
}

// List statistics of the last build, on demand: listed pairs, tiles with a
// listed pair, and of those the pairs and tiles within r_cutoff (the f32
// estimate, and the f64 expression within the band of the threshold, as in
// k_masks).  One wavefront per entry; per-wavefront partial counts.
__global__ __launch_bounds__(256) void k_list_counts(const int *__restrict__ meta, const int *__restrict__ ent_sc,
                                                     const int2 *__restrict__ ent_meta,
                                                     const uint8_t *__restrict__ ent_lb,
                                                     const int *__restrict__ slot_atom, const int4 *__restrict__ sq,
                                                     const double *__restrict__ x, const double *__restrict__ y,
                                                     const double *__restrict__ z, BuildParams bp,
                                                     unsigned long long *__restrict__ parts) {
    const int t = threadIdx.x & 63;
    const int il = t & 7, jl = t >> 3;
    const int nent = (meta[1] || meta[kMetaEntOver]) ? 0 : meta[kMetaNent];
    const int wv = __builtin_amdgcn_readfirstlane((int)blockIdx.x * 4 + (int)(threadIdx.x >> 6));
    const int nw = (int)gridDim.x * 4;
    unsigned long long n_pairs = 0, n_tiles = 0, n_ctiles = 0, n_cpairs = 0;
    for (int e = wv; e < nent; e += nw) {
        const int sc = __builtin_amdgcn_readfirstlane(ent_sc[e]);
        const int2 em = ent_meta[e];
        const int cj = __builtin_amdgcn_readfirstlane(em.x);
        const uint32_t im = (uint32_t)__builtin_amdgcn_readfirstlane(em.y);
        const uint32_t lb = ent_lb[(size_t)e * 64 + t];
        const int4 qb = sq[(size_t)cj * 8 + jl];
        const int b = slot_atom[(size_t)cj * 8 + jl];
        // every cluster's slot data is loaded at once (every slot exists)
        int4 qi[8];
        int ai[8];
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            const size_t is = (size_t)sc * 64 + ci * 8 + il;
            qi[ci] = sq[is];
            ai[ci] = slot_atom[is];
        }
        uint32_t cut = 0;
#pragma unroll
        for (int ci = 0; ci < 8; ci++) {
            const int4 q = qi[ci];
            const int aci = ai[ci];
            if ((lb >> ci) & 1u) {
                const float dx = (float)wrapdiff(qb.x, q.x) * bp.sq[0];
                const float dy = (float)wrapdiff(qb.y, q.y) * bp.sq[1];
                const float dz = (float)wrapdiff(qb.z, q.z) * bp.sq[2];
                const float r2f = dx * dx + dy * dy + dz * dz;
                bool c = r2f <= bp.r_cut2f;
                if (fabsf(r2f - bp.r_cut2f) <= bp.band_cut) {
                    const uint32_t lo = (uint32_t)min(aci, b), hi = (uint32_t)max(aci, b);
                    const double ddx = min_image(x[hi] - x[lo], bp.box[0]);
                    const double ddy = min_image(y[hi] - y[lo], bp.box[1]);
                    const double ddz = min_image(z[hi] - z[lo], bp.box[2]);
                    c = ddx * ddx + ddy * ddy + ddz * ddz <= bp.r_cut2;
                }
                if (c) cut |= 1u << ci;
            }
        }
        n_pairs += (unsigned long long)__popc(lb);
        n_cpairs += (unsigned long long)__popc(cut);
        if (t == 0) n_tiles += (unsigned long long)__popc(im);
        const uint32_t cutor = (uint32_t)__builtin_amdgcn_readfirstlane(wave_or_i32((int)cut));
        if (t == 0) n_ctiles += (unsigned long long)__popc(cutor);
    }
    // per-wavefront sums in a fixed order (lane sums by a tree, then lane 0 writes)
    for (int o = 32; o > 0; o >>= 1) {
        n_pairs += __shfl_xor(n_pairs, o);
        n_cpairs += __shfl_xor(n_cpairs, o);
    }
    if (t == 0) {
        unsigned long long *pw = parts + ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 4;
        pw[0] = n_pairs;
        pw[1] = n_tiles;
        pw[2] = n_ctiles;
        pw[3] = n_cpairs;
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

// One thread per work item: a run of `chunk` consecutive entries of one
// supercluster (whole blocks, counted from its first entry), with its first
// block and the j-cluster of its first entry, and its estimated cost in
// half-tile units (tiles plus 0.5 per entry plus 4 per item).  The costs are
// counted per bin in LDS first, so a workgroup adds one global count per
// distinct cost.
__global__ __launch_bounds__(256) void k_wi_make(const int *meta, const int *sc_start, const int *item_off,
                                                 const int *blk_off, const uint32_t *ent_imask,
                                                 const int *ent_cj, int4 *wi_tmp, int *cost, int *bins) {
    __shared__ int lb[kCostBins];
    for (int k = threadIdx.x; k < kCostBins; k += 256) lb[k] = 0;
    __syncthreads();
    const int it = blockIdx.x * 256 + threadIdx.x;
    if (!meta[kMetaEntOver] && it < meta[kMetaNwi]) {
        const int nsc = meta[1] ? 0 : meta[0];
        const int chunk = meta[kMetaChunk];
        int lo = 0, hi = nsc - 1;  // the last supercluster whose first item is <= it
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (item_off[mid] <= it) lo = mid; else hi = mid - 1;
        }
        const int sc = lo, k = it - item_off[sc];
        const int e0 = sc_start[sc] + k * chunk, e1 = min(e0 + chunk, sc_start[sc + 1]);
        // every load issued before any is used: the index is clamped into the run
        int c = 8;
#pragma unroll
        for (int r = 0; r < kMaxRun; r++) {
            const uint32_t m = ent_imask[min(e0 + r, e1 - 1)];
            c += e0 + r < e1 ? 2 * __popc(m & 0xffu) + 1 : 0;
        }
        wi_tmp[it] = make_int4(sc | ((e1 - e0) << kWiEntryShift), e0, blk_off[sc] + k * (chunk / kBlock),
                               ent_cj[e0]);
        cost[it] = c;
        atomicAdd(&lb[c], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kCostBins; k += 256)
        if (lb[k]) atomicAdd(&bins[k], lb[k]);
}

// Items to their place in the heaviest-first order: each workgroup scans the
// cost counts itself (bins[0..kCostBins)) for the start of every cost in
// decreasing order, ranks its own items in LDS, then reserves one range per
// distinct cost from the running counts (bins[kCostBins + cost], cleared by the
// scan).  The order within one cost is whatever the atomics give: results do
// not depend on the item order.
__global__ __launch_bounds__(256) void k_wi_scatter(const int *meta, const int4 *wi_tmp, const int *cost,
                                                    int *bins, int4 *wi) {
    __shared__ int lb[kCostBins];
    __shared__ int st[kCostBins + 1];  // st[kCostBins - 1 - c]: items costlier than c
    __shared__ int sh[8];
    for (int k = threadIdx.x; k < kCostBins; k += 256) lb[k] = 0;
    block_exclusive_scan<256>(kCostBins, [&](int k) { return bins[kCostBins - 1 - k]; }, st, sh);
    __syncthreads();
    const int it = blockIdx.x * 256 + threadIdx.x;
    const bool on = !meta[kMetaEntOver] && it < meta[kMetaNwi];
    int c = 0, r = 0;
    if (on) {
        c = cost[it];
        r = atomicAdd(&lb[c], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kCostBins; k += 256)
        if (lb[k]) lb[k] = st[kCostBins - 1 - k] + atomicAdd(&bins[kCostBins + k], lb[k]);
    __syncthreads();
    if (on) wi[lb[c] + r] = wi_tmp[it];
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

// ---- Partner lists ----------------------------------------------------------
// The per-slot lists the default nonbonded path walks: each slot's listed
// partners (the pairs of neighbor::build, neighbor.rs:31-43, in both
// directions), the partners within r_cutoff + kPlInner at the build first.
// Built from the full pair search (every cluster near a supercluster, not
// only those at or above it), in five steps without atomics, so the order of
// every list is fixed by the data: partners in ascending full-entry order (j
// cluster), then ascending slot within the cluster.

// Full pair search, pass 0: full entries per supercluster.
__global__ __launch_bounds__(256) void k_fpair_count(const int *meta, BuildParams bp, const int *sc_col,
                                                     const int *col_sc, const int4 *cl_ref,
                                                     const float4 *cl_half, const int4 *sc_ref,
                                                     const float4 *sc_half, int *fsc_count) {
    const int sc = blockIdx.x;
    if (meta[1] || sc >= meta[0]) return;  // the same for the whole block
    const int count = pair_search_sc<0, true>(meta, bp, sc_col, col_sc, cl_ref, cl_half, sc_ref, sc_half,
                                              nullptr, nullptr, nullptr, nullptr);
    if (threadIdx.x == 0) fsc_count[sc] = count;
}

// Full-entry offsets per supercluster, their total and whether it exceeds the
// full-entry buffers' capacity.
__global__ __launch_bounds__(1024) void k_fsc_scan(int *meta, const int *fsc_count, int *fsc_start, int fent_cap) {
    __shared__ int sh[2 * (1024 / 64)];
    const int nsc = meta[1] ? 0 : meta[0];
    const int tot = block_exclusive_scan<1024>(nsc, [&](int k) { return fsc_count[k]; }, fsc_start, sh);
    if (threadIdx.x == 0) {
        meta[kMetaNfent] = tot;
        meta[kMetaFentOver] = tot > fent_cap ? 1 : 0;
        meta[kMetaPlWide] = 0;  // (raised by the mask kernel)
    }
}

// Full pair search, pass 1: the full entries (supercluster, j cluster).
__global__ __launch_bounds__(256) void k_fpair_fill(const int *meta, BuildParams bp, const int *sc_col,
                                                    const int *col_sc, const int4 *cl_ref,
                                                    const float4 *cl_half, const int4 *sc_ref,
                                                    const float4 *sc_half, const int *fsc_start, int *fent_sc,
                                                    int *fent_cj) {
    const int sc = blockIdx.x;
    if (meta[1] || sc >= meta[0] || meta[kMetaFentOver]) return;  // the same for the whole block
    (void)pair_search_sc<1, true>(meta, bp, sc_col, col_sc, cl_ref, cl_half, sc_ref, sc_half, fsc_start,
                                  fent_sc, fent_cj, nullptr);
}

// Full entries of a build, in the pair-mask kernels' walk: workgroups dealt to
// the 8 XCDs in turn, each XCD's wavefronts over a contiguous eighth.
struct EntryWalk {
    int e, e_hi, step;
};
__device__ inline EntryWalk entry_walk(int n) {
    const int P = (int)min(gridDim.x, 8u);
    const int xg = (int)(blockIdx.x % (unsigned)P);
    const int nwx = (int)((gridDim.x - (unsigned)xg + (unsigned)P - 1u) / (unsigned)P) * 4;
    const int l = __builtin_amdgcn_readfirstlane((int)(blockIdx.x / (unsigned)P) * 4 + (int)(threadIdx.x >> 6));
    return EntryWalk{(int)((int64_t)n * xg / P) + l, (int)((int64_t)n * (xg + 1) / P), nwx};
}

// A supercluster's full-entry words are stored per quad of its slots: the
// word of slot t (0..63) for the supercluster's f-th full entry (of nf, from
// f0) is here, so that each quad's words over the entries are contiguous
// (8 bytes per entry).
__device__ inline size_t fent_word_at(int f0, int nf, int t, int f) {
    return (size_t)f0 * 64 + (size_t)(t >> 2) * nf * 4 + (size_t)f * 4 + (t & 3);
}

// One wavefront per full entry (sc, cj), four per workgroup: lane t is slot
// i = sc * 64 + t, and the 8 slots j of cluster cj follow one another (the
// same in every lane: scalar loads).  Bit jl of the lane's word: the pair
// {i, j} is listed, by exactly the decision k_masks makes for it (i != j,
// both positions finite, the f32 estimate of r2 or the f64 expression within
// the band of r_skin^2, not excluded); bit 8 + jl: listed and the estimate at
// most r_in2f + band_in (the list's first part; a pair in the second part is
// farther than r_cutoff + kPlInner at the build).  The estimate of a pair does
// not depend on which slot is i, so both directions get the same decision.
__global__ __launch_bounds__(256) void k_pl_masks(int *meta, const int *__restrict__ fent_sc,
                                                  const int *__restrict__ fent_cj, const int *__restrict__ fsc_start,
                                                  uint16_t *__restrict__ fent_w,
                                                  const int *__restrict__ slot_atom, const int4 *__restrict__ sq,
                                                  const double *__restrict__ x, const double *__restrict__ y,
                                                  const double *__restrict__ z, const int64_t *__restrict__ excl_off,
                                                  const uint32_t *__restrict__ excl_nbr,
                                                  const uint4 *__restrict__ sxb, BuildParams bp,
                                                  int *__restrict__ col_count) {
    // the column counters (last read by the column sort) start the next build
    // at zero (the tile list's mask kernel, when it runs, clears them again)
    for (int c = blockIdx.x * 256 + threadIdx.x; c < bp.ncol; c += gridDim.x * 256) col_count[c] = 0;
    const int t = threadIdx.x & 63;
    const int n = (meta[1] || meta[kMetaFentOver]) ? 0 : meta[kMetaNfent];
    bool wide = false;  // (the same in every lane)
    for (EntryWalk it = entry_walk(n); it.e < it.e_hi; it.e += it.step) {
        const int e = it.e;
        const int sc = __builtin_amdgcn_readfirstlane(fent_sc[e]);
        const int cj = __builtin_amdgcn_readfirstlane(fent_cj[e]);
        // a partner farther than the compact codes reach from its group's first
        // slot (any slot of cluster cj against groups sc * 64 .. sc * 64 + 60):
        // the lists hold 32-bit slot numbers (one flag per wavefront, raised
        // after its walk)
        wide = wide || cj * 8 - (sc * 64 + 60) < -kQposGuard || cj * 8 + 7 - sc * 64 > kQposGuard - 1;
        const int f0 = fsc_start[sc], nf = fsc_start[sc + 1] - f0;
        const size_t si = (size_t)sc * 64 + t;
        const int4 qa = sq[si];
        const int a = qa.w ? -1 : slot_atom[si];  // (a non-finite position is never listed)
        const uint4 xa = sxb[si];
        uint32_t inb = 0, first = 0, band = 0, xhit = 0, far = 0;
        int bat[8];
#pragma unroll
        for (int jl = 0; jl < 8; jl++) {
            const size_t sj = (size_t)cj * 8 + jl;
            const int4 qb = sq[sj];
            const int b = qb.w ? -1 : slot_atom[sj];
            const uint4 xb = sxb[sj];
            bat[jl] = b;
            if (a >= 0 && b >= 0 && a != b) {
                const bool ilo = a < b;
                const int lo = ilo ? a : b, hi = ilo ? b : a;
                const uint32_t k = (uint32_t)(hi - lo - 1);
                if (k < 64u) {
                    const uint32_t w = k < 32u ? (ilo ? xa.x : xb.x) : (ilo ? xa.y : xb.y);
                    if ((w >> (k & 31u)) & 1u) xhit |= 1u << jl;
                } else if ((uint32_t)hi <= (ilo ? xa.z : xb.z)) {
                    far |= 1u << jl;
                }
                const float dx = (float)wrapdiff(qb.x, qa.x) * bp.sq[0];
                const float dy = (float)wrapdiff(qb.y, qa.y) * bp.sq[1];
                const float dz = (float)wrapdiff(qb.z, qa.z) * bp.sq[2];
                const float r2f = dx * dx + dy * dy + dz * dz;
                if (fabsf(r2f - bp.r_skin2f) <= bp.band_skin) band |= 1u << jl;
                if (r2f <= bp.r_skin2f) inb |= 1u << jl;
                if (r2f <= bp.r_in2f + bp.band_in) first |= 1u << jl;
            }
        }
        // pairs near the threshold (rare): the f64 expression decides
        while (band) {
            const int jl = __ffs(band) - 1;
            band &= band - 1u;
            int b = bat[0];
#pragma unroll
            for (int c = 1; c < 8; c++) b = jl == c ? bat[c] : b;
            const uint32_t lo = (uint32_t)min(a, b), hi = (uint32_t)max(a, b);
            const double ddx = min_image(x[hi] - x[lo], bp.box[0]);
            const double ddy = min_image(y[hi] - y[lo], bp.box[1]);
            const double ddz = min_image(z[hi] - z[lo], bp.box[2]);
            const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            inb = (inb & ~(1u << jl)) | (r2 <= bp.r_skin2 ? (1u << jl) : 0u);
        }
        inb &= ~xhit;
        far &= inb;
        while (far) {
            const int jl = __ffs(far) - 1;
            far &= far - 1u;
            int b = bat[0];
#pragma unroll
            for (int c = 1; c < 8; c++) b = jl == c ? bat[c] : b;
            const uint32_t lo = (uint32_t)min(a, b), hi = (uint32_t)max(a, b);
            if (excluded(excl_off, excl_nbr, lo, hi)) inb &= ~(1u << jl);
        }
        fent_w[fent_word_at(f0, nf, t, e - f0)] = (uint16_t)(inb | ((inb & first) << 8));
    }
    if (wide && t == 0) atomicOr(&meta[kMetaPlWide], 1);
}

// One workgroup of 16 wavefronts per supercluster, lane t = slot sc * 64 + t:
// wavefront w sums the lane's partners of both parts over the full entries
// f0 + w, f0 + w + 16, ...; the 16 sums are added per slot through LDS in
// order, per group of kPlAtoms slots the parts' lengths are padded to the
// group's longest, rounded up to kPlPad, with each group's list offset in
// the supercluster (gloc), and the supercluster's lists' total length goes to
// sc_len.
__global__ __launch_bounds__(1024) void k_pl_count(const int *__restrict__ meta, const int *__restrict__ fsc_start,
                                                   const uint16_t *__restrict__ fent_w, int2 *__restrict__ pl_cnt,
                                                   int2 *__restrict__ pl_glen, int *__restrict__ gloc,
                                                   int *__restrict__ sc_len) {
    __shared__ int2 part[16][64];
    const int sc = blockIdx.x;
    if (meta[1] || meta[kMetaFentOver] || sc >= meta[0]) return;  // the same for the whole block
    const int t = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int f0 = fsc_start[sc], nf = fsc_start[sc + 1] - f0;
    int ni = 0, no = 0;
#pragma unroll 4
    for (int f = wv; f < nf; f += 16) {
        const uint32_t v = fent_w[fent_word_at(f0, nf, t, f)];
        const uint32_t fi = (v >> 8) & 255u;
        ni += __popc(fi);
        no += __popc(v & 255u & ~fi);
    }
    part[wv][t] = make_int2(ni, no);
    __syncthreads();
    if (wv != 0) return;
    int si1 = 0, so1 = 0;
    for (int u = 0; u < 16; u++) {
        si1 += part[u][t].x;
        so1 += part[u][t].y;
    }
    pl_cnt[(size_t)sc * 64 + t] = make_int2(si1, so1);
    // group maxima over the group's lanes (a group is kPlAtoms consecutive slots)
    static_assert(kPlAtoms == 2 || kPlAtoms == 4, "groups of 2 or 4 slots");
    int mi = si1, mo = so1;
    mi = max(mi, __builtin_amdgcn_update_dpp(0, mi, 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
    mo = max(mo, __builtin_amdgcn_update_dpp(0, mo, 0xB1, 0xF, 0xF, false));
    if (kPlAtoms == 4) {
        mi = max(mi, __builtin_amdgcn_update_dpp(0, mi, 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
        mo = max(mo, __builtin_amdgcn_update_dpp(0, mo, 0x4E, 0xF, 0xF, false));
    }
    const int pi = (mi + kPlPad - 1) / kPlPad * kPlPad, po = (mo + kPlPad - 1) / kPlPad * kPlPad;
    if ((t % kPlAtoms) == 0) pl_glen[(size_t)sc * (64 / kPlAtoms) + t / kPlAtoms] = make_int2(pi, pi + po);
    // groups' offsets in the supercluster and its total: each group's
    // kPlAtoms lists counted once (by its first lane), scanned over the lanes
    const int own = (t % kPlAtoms) == 0 ? (pi + po) * kPlAtoms : 0;
    int inc = own;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(inc, o, 64);
        if (t >= o) inc += u;
    }
    if ((t % kPlAtoms) == 0) gloc[(size_t)sc * (64 / kPlAtoms) + t / kPlAtoms] = inc - own;
    if (t == 63) sc_len[sc] = inc;
}

// One block: list offsets per supercluster (exclusive scan of the
// superclusters' list lengths), the lists' total length and whether it
// exceeds the list buffer (pl_cap entries before its slack; -1 when there is
// no buffer), and the buffer's slack after the lists (valid slots, read ahead
// by the force kernel and never used).
__global__ __launch_bounds__(1024) void k_pl_scan(int *meta, const int *sc_len, int *sc_off, int pl_cap,
                                                  int *pl_nbr) {
    __shared__ int sh[2 * (1024 / 64)];
    __shared__ long long sh64[1024 / 64];
    const bool ok = !(meta[1] || meta[kMetaFentOver]);
    const int nsc = ok ? meta[0] : 0;
    // the total in 64 bits first: lists too long for 32-bit offsets are
    // reported (length INT_MAX, overflow) before any offset is used
    long long part = 0;
    for (int k = threadIdx.x; k < nsc; k += 1024) part += sc_len[k];
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if ((threadIdx.x & 63) == 0) sh64[threadIdx.x >> 6] = part;
    __syncthreads();
    long long tot64 = 0;
    for (int w = 0; w < 1024 / 64; w++) tot64 += sh64[w];
    if (tot64 > 0x7fffffffLL - kPlSlack) {
        if (threadIdx.x == 0) {
            meta[kMetaNpl] = 0x7fffffff;
            meta[kMetaPlOver] = 1;
        }
        return;
    }
    const int tot = block_exclusive_scan<1024>(nsc, [&](int k) { return sc_len[k]; }, sc_off, sh);
    const bool over = ok && tot > pl_cap;
    if (threadIdx.x == 0) {
        meta[kMetaNpl] = tot;
        meta[kMetaPlOver] = over ? 1 : 0;
    }
    if (ok && !over && threadIdx.x < kPlSlack) pl_nbr[tot + threadIdx.x] = 0;
}

// Compact lists pair the rounds: entries p and p + 16 of each block of 32
// (rounds 2b and 2b + 1 of one lane) are the two halves of one 32-bit word,
// so that a lane reads two rounds' codes with one load.  Blocks start at
// multiples of 32 entries (the parts' padding keeps every list there).
__device__ inline int pl_pair_pos(int x) { return (x & ~31) | ((x & 15) << 1) | ((x >> 4) & 1); }
static_assert(kPlLanes == 16 && kPlPad % 32 == 0, "paired rounds: 16 lanes, parts in blocks of 32");
static_assert(kPlAtoms == 4, "a group is one quad of slots (the mask words' layout)");

// Inclusive sum over the wavefront: within rows of 16 lanes by DPP row
// shifts (lanes shifted in from outside the row read 0), then the lower rows'
// totals.
__device__ inline int wave_incl_sum(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);  // row_shr:8
    const int r0 = __builtin_amdgcn_readlane(x, 15), r1 = __builtin_amdgcn_readlane(x, 31),
              r2 = __builtin_amdgcn_readlane(x, 47);
    const int row = (int)(threadIdx.x & 63) >> 4;
    return x + (row >= 1 ? r0 : 0) + (row >= 2 ? r1 : 0) + (row >= 3 ? r2 : 0);
}

// One wavefront per group (four per workgroup), one slot of the group after
// the other: the supercluster's full entries 64 at a time, lane l taking
// entry f = base + l; its word's partners of each part are counted, summed
// over the lanes in entry order (a wavefront scan of both counts at once), and
// written (partner slot cj * 8 + jl, in ascending entry and jl) at their
// places in the slot's list; then the padding of both parts (the slot
// itself).  Compact codes (slot minus the group's first slot plus kQposGuard,
// paired rounds) or slot numbers, as the build word kMetaPlWide says.  Lane 0
// writes the group's list offset for the force kernel.
template <bool C16>
__device__ inline void pl_make_group(int g, int sc, const int *fsc_start, const int *fent_cj,
                                     const uint16_t *fent_w, const int *sc_off, const int2 *pl_glen, const int *gloc,
                                     int *pl_goff, int *pl_nbr) {
    const int lane = (int)(threadIdx.x & 63);
    const int base = sc_off[sc] + gloc[g];
    if (lane == 0) pl_goff[g] = base;
    const int2 gl = pl_glen[g];
    const int f0 = fsc_start[sc], nf = fsc_start[sc + 1] - f0;
    const int q = g % (64 / kPlAtoms);  // the group's quad in the supercluster
    const uint16_t *qw = fent_w + fent_word_at(f0, nf, q * 4, 0);  // slot a, entry f: qw[f * 4 + a]
    const int gb = C16 ? g * kPlAtoms - kQposGuard : 0;  // code = slot - gb
    for (int a = 0; a < kPlAtoms; a++) {
        const int slot = g * kPlAtoms + a;
        const size_t lb = (size_t)base + (size_t)a * gl.y;  // the slot's list
        auto put = [&](int p, int v) {
            if (C16)
                reinterpret_cast<uint16_t *>(pl_nbr)[lb + pl_pair_pos(p)] = (uint16_t)v;
            else
                pl_nbr[lb + p] = v;
        };
        int pi = 0, po = gl.x;  // next place in each part (the same in every lane)
        for (int fb = 0; fb < nf; fb += 64) {
            const int f = fb + lane;
            uint32_t v = 0;
            int c8 = 0;
            if (f < nf) {
                v = qw[(size_t)f * 4 + a];
                c8 = fent_cj[f0 + f] * 8 - gb;
            }
            const uint32_t fi = v & (v >> 8) & 255u, so = v & ~(v >> 8) & 255u;
            const int cnt = __popc(fi) | (__popc(so) << 16);  // (at most 512 per wavefront: no carry)
            const int inc = wave_incl_sum(cnt);
            const int tot = __builtin_amdgcn_readlane(inc, 63);
            int xi = pi + ((inc - cnt) & 0xffff), xo = po + ((inc - cnt) >> 16);
#pragma unroll
            for (int jl = 0; jl < 8; jl++) {
                const bool in_f = (fi >> jl) & 1u, in_s = (so >> jl) & 1u;
                if (in_f || in_s) put(in_f ? xi : xo, c8 + jl);
                xi += in_f ? 1 : 0;
                xo += in_s ? 1 : 0;
            }
            pi += tot & 0xffff;
            po += tot >> 16;
        }
        for (int p = pi + lane; p < gl.x; p += 64) put(p, slot - gb);
        for (int p = po + lane; p < gl.y; p += 64) put(p, slot - gb);
    }
}

__global__ __launch_bounds__(256) void k_pl_make(const int *__restrict__ meta, const int *__restrict__ fsc_start,
                                                  const int *__restrict__ fent_cj, const uint16_t *__restrict__ fent_w,
                                                  const int *__restrict__ sc_off, const int2 *__restrict__ pl_glen,
                                                  const int *__restrict__ gloc, int *__restrict__ pl_goff,
                                                  int *__restrict__ pl_nbr) {
    if (meta[1] || meta[kMetaFentOver] || meta[kMetaPlOver]) return;  // the same for the whole grid
    const int g = __builtin_amdgcn_readfirstlane((int)blockIdx.x * 4 + (int)(threadIdx.x >> 6));
    const int sc = g / (64 / kPlAtoms);
    if (sc >= meta[0]) return;  // (the same for the whole wavefront)
    if (meta[kMetaPlWide])
        pl_make_group<false>(g, sc, fsc_start, fent_cj, fent_w, sc_off, pl_glen, gloc, pl_goff, pl_nbr);
    else
        pl_make_group<true>(g, sc, fsc_start, fent_cj, fent_w, sc_off, pl_glen, gloc, pl_goff, pl_nbr);
}

// Per slot, the range of its atom's bonded force records (empty for an empty
// slot), for the finish kernel of a partner-list evaluation.
__global__ __launch_bounds__(256) void k_slot_rec(const int *__restrict__ meta, const int *__restrict__ slot_atom,
                                                  const int *__restrict__ bc_off, int2 *__restrict__ srec) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (meta[1] || s >= meta[0] * 64) return;
    const int a = slot_atom[s];
    srec[s] = a >= 0 ? make_int2(bc_off[a], bc_off[a + 1]) : make_int2(0, 0);
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
hipError_t launch_extent(gm_sys *s, const double *px, const double *py, const double *pz, double *out) {
    k_extent<<<1, 1024, 0, s->stream>>>((int)s->n, px, py, pz, out);
    return hipGetLastError();
}

hipError_t launch_build(gm_sys *s, bool wrap, double *px, double *py, double *pz, bool clear_cols) {
    hipStream_t st = s->stream;
    const int n = (int)s->n;
    const BuildParams &bp = s->bp;
    hipError_t e;
    if (clear_cols && (e = launch_fill_i32(s->col_count.p, bp.ncol, 0, st)) != hipSuccess) return e;
    const bool lds = bp.ncol <= kLdsColumns;
    if (n > 0) {
        if (lds)
            k_wrap_quantize_lds<<<grid_for(n, 1024), 1024, sizeof(int) * (size_t)bp.ncol, st>>>(
                n, px, py, pz, s->x.p, s->y.p, s->z.p, s->refx.p, s->refy.p, s->refz.p, s->qatom.p,
                s->col_count.p, bp, wrap ? 1 : 0);
        else
            k_wrap_quantize<<<grid_for(n, 256), 256, 0, st>>>(n, px, py, pz, s->x.p, s->y.p, s->z.p, s->refx.p,
                                                                s->refy.p, s->refz.p, s->qatom.p,
                                                                s->col_count.p, bp, wrap ? 1 : 0);
    }
    k_scan_columns<<<1, 1024, 0, st>>>(bp.ncol, s->col_count.p, s->col_start.p, s->col_sc.p, s->meta.p,
                                       kColumnCap, s->col_fill.p);
    if (n > 0) {
        if (lds)
            k_scatter_lds<<<grid_for(n, 1024), 1024, sizeof(int) * (size_t)bp.ncol, st>>>(
                n, s->qatom.p, s->col_start.p, s->col_fill.p, s->bucket.p, bp);
        else
            k_scatter<<<grid_for(n, 256), 256, 0, st>>>(n, s->qatom.p, s->col_start.p, s->col_fill.p,
                                                          s->bucket.p, bp);
    }
    k_column_sort<<<bp.ncol, 1024, kColumnCap * sizeof(unsigned long long), st>>>(
        n, s->qatom.p, s->bucket.p, s->col_start.p, s->col_count.p, s->col_sc.p,
        s->sorted_atoms.p, s->sc_base.p, s->sc_m.p, s->sc_col.p);
    k_sc_setup<<<(unsigned)s->max_sc, 64, 0, st>>>(
        s->meta.p, n, (int)s->max_sc, s->qatom.p, s->sorted_atoms.p, s->sc_base.p, s->sc_m.p, s->slot_atom.p,
        s->atom_par.p, s->excl_bits.p, s->atom_slot.p, s->sq.p, s->sp.p, s->qpos.p, s->qeps.p, s->sxb.p,
        kernel_sigma_scale(s), s->x.p, s->y.p, s->z.p, bp, s->cl_ref.p, s->cl_half.p, s->sc_ref.p, s->sc_half.p,
        s->F.p);
    return hipGetLastError();
}

namespace {
ScScanArgs sc_scan_args(gm_sys *s, int64_t ent_cap, int chunk_req) {
    ScScanArgs A;
    A.sc_start = s->sc_start.p;
    A.item_off = s->item_off.p;
    A.blk_off = s->sc_blk_off.p;
    A.bins = s->cost_bins.p;
    A.ent_cap = (int)std::min<int64_t>(ent_cap, 0x7fffffff);
    A.chunk_req = chunk_req;
    A.slots = 20 * (s->ctx->n_cu > 0 ? s->ctx->n_cu : 256);
    return A;
}
}  // namespace

// Grids of the list kernels cover the capacities (the supercluster bound
// max_sc, the entry buffers' ent_cap); each kernel reads the actual counts
// from meta and leaves the rest alone, so the build needs no host round trip
// between its kernels.
hipError_t launch_pair_count(gm_sys *s, int64_t ent_cap, int chunk_req) {
    if (s->max_sc > 0)
        k_pair_count<<<(unsigned)s->max_sc, 256, 0, s->stream>>>(s->meta.p, s->bp, s->sc_col.p, s->col_sc.p,
                                                                 s->cl_ref.p, s->cl_half.p, s->sc_ref.p,
                                                                 s->sc_half.p, s->sc_count.p);
    return launch_sc_scan(s, ent_cap, chunk_req);
}

hipError_t launch_sc_scan(gm_sys *s, int64_t ent_cap, int chunk_req) {
    k_sc_scan<<<1, 1024, 0, s->stream>>>(s->meta.p, s->sc_count.p, sc_scan_args(s, ent_cap, chunk_req));
    return hipGetLastError();
}

hipError_t launch_pair_fill_and_masks(gm_sys *s) {
    hipStream_t st = s->stream;
    if (s->max_sc > 0)
        k_pair_fill<<<(unsigned)s->max_sc, 256, 0, st>>>(s->meta.p, s->bp, s->sc_col.p, s->col_sc.p,
                                                         s->cl_ref.p, s->cl_half.p, s->sc_ref.p, s->sc_half.p,
                                                         s->sc_start.p, s->ent_sc.p, s->ent_cj.p, s->ent_imask.p);
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((s->ent_cap + 3) / 4, kMaskBlocks));
    k_masks<<<(unsigned)blocks, 256, 0, st>>>(s->meta.p, s->ent_sc.p, s->ent_cj.p, s->ent_imask.p, s->ent_meta.p,
                                              s->ent_dmin.p, s->ent_lb.p, s->slot_atom.p, s->sq.p, s->refx.p,
                                              s->refy.p, s->refz.p, s->excl_off.p, s->excl_nbr.p, s->sxb.p, s->bp,
                                              s->col_count.p);
    return hipGetLastError();
}

hipError_t launch_pl_count(gm_sys *s) {
    hipStream_t st = s->stream;
    if (s->max_sc > 0)
        k_fpair_count<<<(unsigned)s->max_sc, 256, 0, st>>>(s->meta.p, s->bp, s->sc_col.p, s->col_sc.p, s->cl_ref.p,
                                                           s->cl_half.p, s->sc_ref.p, s->sc_half.p, s->fsc_count.p);
    k_fsc_scan<<<1, 1024, 0, st>>>(s->meta.p, s->fsc_count.p, s->fsc_start.p,
                                   (int)std::min<int64_t>(s->fent_cap, 0x7fffffff));
    return hipGetLastError();
}

hipError_t launch_pl_fill(gm_sys *s) {
    hipStream_t st = s->stream;
    if (s->max_sc > 0)
        k_fpair_fill<<<(unsigned)s->max_sc, 256, 0, st>>>(s->meta.p, s->bp, s->sc_col.p, s->col_sc.p, s->cl_ref.p,
                                                          s->cl_half.p, s->sc_ref.p, s->sc_half.p, s->fsc_start.p,
                                                          s->fent_sc.p, s->fent_cj.p);
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((s->fent_cap + 3) / 4, kMaskBlocks));
    k_pl_masks<<<(unsigned)blocks, 256, 0, st>>>(s->meta.p, s->fent_sc.p, s->fent_cj.p, s->fsc_start.p, s->fent_w.p,
                                                 s->slot_atom.p, s->sq.p, s->refx.p, s->refy.p, s->refz.p,
                                                 s->excl_off.p, s->excl_nbr.p, s->sxb.p, s->bp, s->col_count.p);
    if (s->max_sc > 0)
        k_pl_count<<<(unsigned)s->max_sc, 1024, 0, st>>>(s->meta.p, s->fsc_start.p, s->fent_w.p, s->pl_cnt.p,
                                                         s->pl_glen.p, s->pl_gloc.p, s->fsc_count.p);
    // (no list buffer yet: capacity -1, so that even empty lists overflow and
    // the host allocates it, with the slack the force kernel reads)
    // (sc_len and sc_off reuse the full-entry count and scan buffers: the
    // full-entry offsets are in fsc_start, and the counts were only scanned)
    k_pl_scan<<<1, 1024, 0, st>>>(s->meta.p, s->fsc_count.p, s->pl_scoff.p,
                                  s->pl_nbr.p ? (int)std::min<int64_t>(s->pl_cap, 0x7fffffff) : -1, s->pl_nbr.p);
    if (s->max_sc > 0)
        k_pl_make<<<(unsigned)(s->max_sc * (64 / kPlAtoms) / 4), 256, 0, st>>>(
            s->meta.p, s->fsc_start.p, s->fent_cj.p, s->fent_w.p, s->pl_scoff.p, s->pl_glen.p, s->pl_gloc.p,
            s->pl_goff.p, s->pl_nbr.p);
    if (s->max_sc > 0 && s->n_contrib > 0)
        k_slot_rec<<<(unsigned)((s->max_sc * 64 + 255) / 256), 256, 0, st>>>(s->meta.p, s->slot_atom.p, s->bc_off.p,
                                                                           s->srec.p);
    return hipGetLastError();
}

hipError_t launch_work_items(gm_sys *s) {
    hipStream_t st = s->stream;
    const unsigned g = (unsigned)std::max<int64_t>(1, (s->item_cap + 255) / 256);
    k_wi_make<<<g, 256, 0, st>>>(s->meta.p, s->sc_start.p, s->item_off.p, s->sc_blk_off.p, s->ent_imask.p,
                                 s->ent_cj.p, s->wi_tmp.p, s->wi_cost.p, s->cost_bins.p);
    k_wi_scatter<<<g, 256, 0, st>>>(s->meta.p, s->wi_tmp.p, s->wi_cost.p, s->cost_bins.p, s->wi.p);
    return hipGetLastError();
}

hipError_t launch_list_stats(gm_sys *s) {
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((s->nent + 3) / 4, kMaskBlocks));
    hipError_t e = s->mask_parts.ensure((size_t)blocks * 4 * 4);  // 4 counts per wavefront
    if (e != hipSuccess) return e;
    unsigned long long *counters = (unsigned long long *)(s->meta.p + 8);
    k_list_counts<<<(unsigned)blocks, 256, 0, s->stream>>>(s->meta.p, s->ent_sc.p, s->ent_meta.p, s->ent_lb.p,
                                                            s->slot_atom.p, s->sq.p, s->refx.p, s->refy.p, s->refz.p,
                                                            s->bp, s->mask_parts.p);
    k_sum_parts<<<1, 1024, 0, s->stream>>>(s->mask_parts.p, blocks * 4, counters);
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
