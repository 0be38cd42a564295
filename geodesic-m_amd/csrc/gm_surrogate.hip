// Shell-force surrogate kernels for gfx950: radial features, the BF16 MFMA
// network (inference and a hand-written backward pass), the optimiser, and
// the elementwise steps of the multiple-time-step driver.  No reference
// counterpart: the reference has no surrogate.  The model is described in
// gm_surrogate.hpp.
#include "gm_geom.hpp"
#include "gm_surrogate.hpp"

using namespace gm;
using namespace gm::sur;

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (no NaN inputs occur here).
__device__ inline unsigned short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// Quintic fade 0 -> 1 over u in [0, 1] (one minus the reference switch S(u)).
__device__ inline float fade(float u) {
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    return u * u * u * (10.0f + u * (-15.0f + 6.0f * u));
}

// ---- features -------------------------------------------------------------

// Current positions in slot order as fixed-point fractions of the box, with
// each slot's class (-1 for an empty slot).
__global__ void k_feat_prep(int n_slots, const int *slot_atom, const double *x, const double *y,
                            const double *z, const uint8_t *cls, double iq0, double iq1, double iq2,
                            int4 *qs) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    const int a = slot_atom[s];
    if (a < 0) {
        qs[s] = make_int4(0, 0, 0, -1);
        return;
    }
    qs[s] = make_int4(quantize_fixed(x[a], iq0), quantize_fixed(y[a], iq1), quantize_fixed(z[a], iq2),
                      (int)cls[a]);
}

// Quintic switch S(u) = 1 - 10u^3 + 15u^4 - 6u^5 on u in [0, 1] and its first
// two derivatives with respect to u (all constant outside [0, 1]).
__device__ inline void switch_d2(float u, float &S, float &S1, float &S2) {
    u = fminf(fmaxf(u, 0.0f), 1.0f);
    const float om = 1.0f - u;
    S = 1.0f + u * u * u * (-10.0f + u * (15.0f - 6.0f * u));
    S1 = -30.0f * u * u * om * om;
    S2 = -60.0f * u * om * (1.0f - 2.0f * u);
}

constexpr int kRed = kNG + 7;  // LDS row: 24 features, 6 Hessian components, pad

// acc[k] += env exp(-(t - k)^2 / 2), k = 0..23, for t in [0, 23]: two
// recurrences from the middle centres, G(k-1) = G(k) e^{k - t - 1/2} downward
// from k = 11 and G(k+1) = G(k) e^{t - k - 1/2} upward from k = 12, so every
// intermediate value is a Gaussian value itself (no underflow ahead of a large
// one) and a pair costs 4 exponentials instead of 24.
__device__ __forceinline__ void gauss_add(float *acc, float t, float env) {
    static_assert(kNG == 24, "recurrence split at 11 / 12");
    const float L2E = 1.4426950408889634f;
    const float d11 = t - 11.0f, d12 = t - 12.0f;
    float g_dn = env * exp2f(-0.5f * L2E * d11 * d11);
    float g_up = env * exp2f(-0.5f * L2E * d12 * d12);
    const float base_dn = exp2f((-t - 0.5f) * L2E);  // times e^k
    const float base_up = exp2f((t - 0.5f) * L2E);   // times e^-k
    acc[11] += g_dn;
    acc[12] += g_up;
#pragma unroll
    for (int k = 11; k > 0; k--) {
        g_dn *= base_dn * __builtin_expf((float)k);  // constant-folded e^k
        acc[k - 1] += g_dn;
    }
#pragma unroll
    for (int k = 12; k < kNG - 1; k++) {
        g_up *= base_up * __builtin_expf(-(float)k);
        acc[k + 1] += g_up;
    }
}

// Whether atoms a and b form an excluded pair.  rg_a is a's partner range on
// either side (kept in a register by the caller): a partner outside it is
// rejected without a memory access; otherwise a binary search in the CSR of
// partners above min(a, b).
__device__ inline bool excluded_pair(const FeatArgs &A, int a, uint2 rg_a, int b) {
    if (a < 0 || b < 0) return true;
    if ((uint32_t)b < rg_a.x || (uint32_t)b > rg_a.y) return false;
    const uint32_t lo = (uint32_t)min(a, b), hi = (uint32_t)max(a, b);
    int64_t l = A.excl_off[lo], r = A.excl_off[lo + 1];
    while (l < r) {
        const int64_t m = (l + r) >> 1;
        if (A.excl_nbr[m] < hi) l = m + 1; else r = m;
    }
    return l < A.excl_off[lo + 1] && A.excl_nbr[l] == hi;
}

// One workgroup of kFeatWaves wavefronts per supercluster of the exact
// system; lane = slot.  Wavefront w owns atom class w % 4 and every other
// candidate cluster (parity w / 4).  All wavefronts walk the same candidate
// clusters (xy columns in reach, then the box test against the
// supercluster); j atoms are the same for all lanes and a j atom of another
// class is skipped with a uniform branch, so each lane accumulates its own
// atom's 24 features of one class, and its part of the shell Hessian K_i,
// with no cross-lane traffic.  Partial sums are added in a fixed order, so
// the results are reproducible bit for bit.
//
// K_i = sum_j H(d_ij), H = u'' d d^T / r^2 + (u' / r)(I - d d^T / r^2) for the
// shell pair potential u(r) = V_LJ(r) W(r), W = S(u1) - S(u2).
__global__ __launch_bounds__(64 * kFeatWaves) void k_features(FeatArgs A, int nsc) {
    static_assert(kFeatWaves == 2 * kNC, "two wavefronts per class");
    __shared__ float red[kFeatWaves][64][kRed];
    const int sc = blockIdx.x;
    if (sc >= nsc) return;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int cls = w % kNC, half = w / kNC;
    const BuildParams &bp = A.bp;
    const int4 qi = A.qs[(size_t)sc * 64 + lane];
    const float2 spi = A.sp[(size_t)sc * 64 + lane];
    const int ai = A.slot_atom[(size_t)sc * 64 + lane];
    const uint2 rg_i = ai >= 0 ? A.excl_rng_any[ai] : make_uint2(1u, 0u);
    float acc[kNG];
#pragma unroll
    for (int k = 0; k < kNG; k++) acc[k] = 0.0f;
    float kxx = 0.0f, kyy = 0.0f, kzz = 0.0f, kxy = 0.0f, kxz = 0.0f, kyz = 0.0f;
    const float lo2 = fminf(A.r_lo2, A.rk_lo2), hi2 = fmaxf(A.r_hi2, A.rk_hi2);

    const int4 sr = A.sc_ref[sc];
    const float4 sh = A.sc_half[sc];
    const int c0 = A.sc_col[sc];
    const int cx0 = c0 / bp.ncy, cy0 = c0 - (c0 / bp.ncy) * bp.ncy;
    int2 px[2], py[2];
    ring_pieces(cx0, bp.kx, bp.ncx, px[0], px[1]);
    ring_pieces(cy0, bp.ky, bp.ncy, py[0], py[1]);
    const int npx = 2, npy = 2;  // an absent piece is [0, -1]: no columns
    const float rs2 = bp.r_search * bp.r_search;
    int seen = 0;
    for (int a = 0; a < npx; a++) {
        for (int cx = px[a].x; cx <= px[a].y; cx++) {
            const float gx = column_gap(cx - cx0, bp.ncx, bp.wcol[0]);
            for (int b = 0; b < npy; b++) {
                for (int cy = py[b].x; cy <= py[b].y; cy++) {
                    const float gy = column_gap(cy - cy0, bp.ncy, bp.wcol[1]);
                    if (gx * gx + gy * gy > rs2) continue;
                    const int c = cx * bp.ncy + cy;
                    const int lo = A.col_sc[c] * 8, hi = A.col_sc[c + 1] * 8;
                    for (int j0 = lo; j0 < hi; j0 += 64) {
                        const int cj = j0 + lane;
                        bool ok = false;
                        if (cj < hi) {
                            const int4 jr = A.cl_ref[cj];
                            if (jr.w) ok = boxes_close(sr, sh, jr, A.cl_half[cj], bp);
                        }
                        unsigned long long mask = __ballot(ok);
                        while (mask) {
                            // gather up to 8 of this wavefront's clusters: lane l
                            // loads atom l % 8 of the (l / 8)-th of them
                            int myJ = -1, ng = 0;
                            while (mask && ng < 8) {
                                const int bit = __builtin_ctzll(mask);
                                mask &= mask - 1;
                                if (((seen++) & 1) != half) continue;
                                if ((lane >> 3) == ng) myJ = j0 + bit;
                                ng++;
                            }
                            const size_t js = (size_t)(myJ < 0 ? 0 : myJ) * 8 + (lane & 7);
                            const int4 qv = myJ >= 0 ? A.qs[js] : make_int4(0, 0, 0, -1);
                            const float2 sv = myJ >= 0 ? A.sp[js] : make_float2(0.0f, 0.0f);
                            const int av = myJ >= 0 ? A.slot_atom[js] : -1;
                            for (int t = 0; t < ng * 8; t++) {
                                const int cj4 = __builtin_amdgcn_readlane(qv.w, t);
                                if (cj4 != cls) continue;
                                const int qjx = __builtin_amdgcn_readlane(qv.x, t);
                                const int qjy = __builtin_amdgcn_readlane(qv.y, t);
                                const int qjz = __builtin_amdgcn_readlane(qv.z, t);
                                const float dx = (float)wrapdiff(qjx, qi.x) * bp.sq[0];
                                const float dy = (float)wrapdiff(qjy, qi.y) * bp.sq[1];
                                const float dz = (float)wrapdiff(qjz, qi.z) * bp.sq[2];
                                const float r2 = dx * dx + dy * dy + dz * dz;
                                if (qi.w >= 0 && r2 >= lo2 && r2 <= hi2) {
                                    const float r = sqrtf(r2);
                                    if (r2 >= A.r_lo2 && r2 <= A.r_hi2) {
                                        const float env = fade((r - A.r_lo) * A.inv_delta) *
                                                          fade((A.r_hi - r) * A.inv_delta);
                                        const float tt = (r - A.mu0) * A.inv_w;
                                        gauss_add(acc, tt, env);
                                    }
                                    const int aj = __builtin_amdgcn_readlane(av, t);
                                    if (r2 >= A.rk_lo2 && r2 <= A.rk_hi2 && !excluded_pair(A, ai, rg_i, aj)) {
                                        const float sjx = __int_as_float(
                                            __builtin_amdgcn_readlane(__float_as_int(sv.x), t));
                                        const float sjy = __int_as_float(
                                            __builtin_amdgcn_readlane(__float_as_int(sv.y), t));
                                        const float ir = 1.0f / r;
                                        const float sig = spi.x + sjx, eps = spi.y * sjy;
                                        const float q2 = sig * sig * ir * ir, q6 = q2 * q2 * q2, q12 = q6 * q6;
                                        const float V = 4.0f * eps * (q12 - q6);
                                        const float V1 = -24.0f * eps * (2.0f * q12 - q6) * ir;
                                        const float V2 = 24.0f * eps * (26.0f * q12 - 7.0f * q6) * ir * ir;
                                        float Sa, Sa1, Sa2, Sb, Sb1, Sb2;
                                        switch_d2((r - A.a1) * A.inv_w1, Sa, Sa1, Sa2);
                                        switch_d2((r - A.a2) * A.inv_w2, Sb, Sb1, Sb2);
                                        const float W = Sa - Sb;
                                        const float W1 = Sa1 * A.inv_w1 - Sb1 * A.inv_w2;
                                        const float W2 = Sa2 * A.inv_w1 * A.inv_w1 - Sb2 * A.inv_w2 * A.inv_w2;
                                        const float du = V1 * W + V * W1;
                                        const float d2u = V2 * W + 2.0f * V1 * W1 + V * W2;
                                        const float b = du * ir, a = (d2u - b) * ir * ir;
                                        kxx += a * dx * dx + b;
                                        kyy += a * dy * dy + b;
                                        kzz += a * dz * dz + b;
                                        kxy += a * dx * dy;
                                        kxz += a * dx * dz;
                                        kyz += a * dy * dz;
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kNG; k++) red[w][lane][k] = acc[k];
    red[w][lane][kNG + 0] = kxx;
    red[w][lane][kNG + 1] = kyy;
    red[w][lane][kNG + 2] = kzz;
    red[w][lane][kNG + 3] = kxy;
    red[w][lane][kNG + 4] = kxz;
    red[w][lane][kNG + 5] = kyz;
    __syncthreads();
    // 64 slots x 96 features, halves added in a fixed order
    for (int e = threadIdx.x; e < 64 * kNF; e += 64 * kFeatWaves) {
        const int l = e / kNF, f = e - (e / kNF) * kNF;
        const int c = f / kNG, k = f - (f / kNG) * kNG;
        const int atom = A.slot_atom[(size_t)sc * 64 + l];
        if (atom >= 0) A.feat[(size_t)atom * kNF + f] = red[c][l][k] + red[kNC + c][l][k];
    }
    // 64 slots x 6 Hessian components, the 8 wavefronts in order
    for (int e = threadIdx.x; e < 64 * 6; e += 64 * kFeatWaves) {
        const int l = e / 6, m = e - (e / 6) * 6;
        const int atom = A.slot_atom[(size_t)sc * 64 + l];
        float v = red[0][l][kNG + m];
        for (int u = 1; u < kFeatWaves; u++) v += red[u][l][kNG + m];
        if (atom >= 0) A.kmat[(size_t)atom * 6 + m] = v;
    }
}

// ---- network ----------------------------------------------------------------

// One dense layer for the 16 rows of a wavefront on MFMA:
// acc[nt] = X[16][K] (BF16, LDS, row stride ldx) x W[nt*16 .. +16][K]^T (BF16).
// v_mfma_f32_16x16x16_bf16 operand layout: lane l holds A[l%16][4(l/16)..+3]
// and B[4(l/16)..+3][l%16]; result D[4(l/16)+i][l%16] in acc element i.
template <int K, int NT>
__device__ __forceinline__ void mfma_layer(const unsigned short *xs, int ldx,
                                           const unsigned short *__restrict__ w, int lane, f4 *acc) {
    const int r = lane & 15, kq = (lane >> 4) << 2;
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[nt] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kt = 0; kt < K / 16; kt++) {
        const s4 a = *reinterpret_cast<const s4 *>(xs + r * ldx + kt * 16 + kq);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const s4 b = *reinterpret_cast<const s4 *>(w + (size_t)(nt * 16 + r) * K + kt * 16 + kq);
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
        }
    }
}

constexpr int kLdX = kNF + 4;  // padded LDS row strides (BF16 elements)
constexpr int kLdH = kH + 4;

// Standardised features of 16 atoms into LDS as BF16 (zero rows past n).
__device__ inline void load_x16(const NetArgs &N, const int64_t *atoms, unsigned short *xs, int lane) {
    for (int e = lane; e < 16 * kNF; e += 64) {
        const int row = e / kNF, k = e - (e / kNF) * kNF;
        const int64_t a = atoms[row];
        float v = 0.0f;
        if (a >= 0) v = (N.feat[a * kNF + k] - N.fmean[k]) * N.finv[k];
        xs[row * kLdX + k] = f2bf(v);
    }
}

// Forward pass for the 16 atoms whose BF16 inputs are in xs.  Writes the
// hidden activations (BF16) to h1s/h2s, optionally the pre-activations
// (fp32, row stride kH) to z1/z2, and the outputs (fp32) to o[16][3].
template <bool SAVE>
__device__ inline void forward16(const NetArgs &N, const unsigned short *xs, unsigned short *h1s,
                                 unsigned short *h2s, float *z1, float *z2, float *o, int lane) {
    const int rq = (lane >> 4) << 2, col0 = lane & 15;
    f4 acc[8];
    mfma_layer<kNF, 8>(xs, kLdX, N.wbf + kBfW1, lane, acc);
#pragma unroll
    for (int nt = 0; nt < 8; nt++) {
        const int col = nt * 16 + col0;
        const float b = N.param[kOffB1 + col];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float zz = acc[nt][i] + b;
            if (SAVE) z1[(rq + i) * kH + col] = zz;
            h1s[(rq + i) * kLdH + col] = f2bf(silu(zz));
        }
    }
    __syncthreads();
    mfma_layer<kH, 8>(h1s, kLdH, N.wbf + kBfW2, lane, acc);
#pragma unroll
    for (int nt = 0; nt < 8; nt++) {
        const int col = nt * 16 + col0;
        const float b = N.param[kOffB2 + col];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float zz = acc[nt][i] + b;
            if (SAVE) z2[(rq + i) * kH + col] = zz;
            h2s[(rq + i) * kLdH + col] = f2bf(silu(zz));
        }
    }
    __syncthreads();
    mfma_layer<kH, 1>(h2s, kLdH, N.wbf + kBfW3, lane, acc);
    if (col0 < kNO) {
        const float b = N.param[kOffB3 + col0];
#pragma unroll
        for (int i = 0; i < 4; i++) o[(rq + i) * kNO + col0] = acc[0][i] + b;
    }
}

// Inference: 4 wavefronts x 16 atoms per workgroup; out is n x 3.
__global__ __launch_bounds__(256) void k_infer(NetArgs N, float *out) {
    __shared__ unsigned short xs[4][16 * kLdX];
    __shared__ unsigned short h1s[4][16 * kLdH];
    __shared__ unsigned short h2s[4][16 * kLdH];
    __shared__ float os[4][16 * kNO];
    __shared__ int64_t atoms[4][16];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t base = (int64_t)blockIdx.x * kInferAtoms + w * 16;
    if (lane < 16) atoms[w][lane] = (base + lane < N.n) ? base + lane : -1;
    __syncthreads();
    load_x16(N, atoms[w], xs[w], lane);
    __syncthreads();
    forward16<false>(N, xs[w], h1s[w], h2s[w], nullptr, nullptr, os[w], lane);
    __syncthreads();
    if (lane < 16 * kNO) {
        const int row = lane / kNO, j = lane - (lane / kNO) * kNO;
        const int64_t a = atoms[w][row];
        if (a >= 0) out[a * kNO + j] = os[w][row * kNO + j];
    }
}

// One training step's gradient over kTrainAtoms atoms per workgroup (2
// wavefronts x 16 for the forward pass, 128 threads for the backward pass).
// Loss = (1/nb) sum_i |Fhat_i - T_i|^2 with Fhat = (1 + o1) B1 + o2 B2 + o3 B3
// (normalised vectors).  The backward pass runs in fp32 on the values the
// forward pass used (BF16 inputs, weights and activations; rounding is taken
// as the identity).  Each workgroup writes its partial gradient; they are
// added in a fixed order by k_reduce_adam.
__global__ __launch_bounds__(128) void k_train(TrainArgs T) {
    __shared__ unsigned short xs[kTrainAtoms * kLdX];
    __shared__ unsigned short h1s[kTrainAtoms * kLdH];
    __shared__ unsigned short h2s[kTrainAtoms * kLdH];
    __shared__ float z1s[kTrainAtoms * kH];   // z1, then dZ1
    __shared__ float z2s[kTrainAtoms * kH];   // z2, then dZ2
    __shared__ float os[kTrainAtoms * kNO];
    __shared__ float dos[kTrainAtoms * kNO];
    __shared__ double lossv[kTrainAtoms], lossv0[kTrainAtoms], lossvt[kTrainAtoms];
    __shared__ int64_t atoms[kTrainAtoms];
    const NetArgs &N = T.net;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int64_t b0 = (int64_t)blockIdx.x * kTrainAtoms;
    if (t < kTrainAtoms) {
        const int64_t b = b0 + t;
        atoms[t] = (b < T.nb) ? (T.start + b) % N.n : -1;
    }
    __syncthreads();
    load_x16(N, atoms + w * 16, xs + w * 16 * kLdX, lane);
    __syncthreads();
    forward16<true>(N, xs + w * 16 * kLdX, h1s + w * 16 * kLdH, h2s + w * 16 * kLdH, z1s + w * 16 * kH,
                    z2s + w * 16 * kH, os + w * 16 * kNO, lane);
    __syncthreads();

    // residuals and dL/do
    if (t < kTrainAtoms) {
        const int64_t b = b0 + t;
        double l = 0.0, l0 = 0.0, lt = 0.0;
        float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
        if (b < T.nb) {
            const float *B = T.basis + b * 9;
            const float *Tg = T.target + b * 3;
            const float o1 = os[t * kNO], o2 = os[t * kNO + 1], o3 = os[t * kNO + 2];
            const float sc = 2.0f / (float)T.nb;
            for (int d = 0; d < 3; d++) {
                const float fh = (1.0f + o1) * B[d] + o2 * B[3 + d] + o3 * B[6 + d];
                const float r = fh - Tg[d];
                const float r0 = B[d] - Tg[d];
                const float rt = B[d] + T.taylor * B[3 + d] - Tg[d];
                l += (double)r * (double)r;
                l0 += (double)r0 * (double)r0;
                lt += (double)rt * (double)rt;
                d0 += sc * r * B[d];
                d1 += sc * r * B[3 + d];
                d2 += sc * r * B[6 + d];
            }
        }
        dos[t * kNO] = d0;
        dos[t * kNO + 1] = d1;
        dos[t * kNO + 2] = d2;
        lossv[t] = l;
        lossv0[t] = l0;
        lossvt[t] = lt;
    }
    __syncthreads();

    const unsigned short *W2 = N.wbf + kBfW2, *W3 = N.wbf + kBfW3;
    const int n = t;  // hidden unit owned by this thread
    float *P = T.partial + (size_t)blockIdx.x * kNParam;
    // dZ2 = (dO W3) * silu'(z2); dW3, db3
    {
        const float w0 = bf2f(W3[0 * kH + n]), w1 = bf2f(W3[1 * kH + n]), w2 = bf2f(W3[2 * kH + n]);
        float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
        for (int i = 0; i < kTrainAtoms; i++) {
            const float h2 = bf2f(h2s[i * kLdH + n]);
            g0 += dos[i * kNO] * h2;
            g1 += dos[i * kNO + 1] * h2;
            g2 += dos[i * kNO + 2] * h2;
            const float g = dos[i * kNO] * w0 + dos[i * kNO + 1] * w1 + dos[i * kNO + 2] * w2;
            z2s[i * kH + n] = g * dsilu(z2s[i * kH + n]);
        }
        P[kOffW3 + 0 * kH + n] = g0;
        P[kOffW3 + 1 * kH + n] = g1;
        P[kOffW3 + 2 * kH + n] = g2;
        if (n < kNO) {
            float s = 0.0f;
            for (int i = 0; i < kTrainAtoms; i++) s += dos[i * kNO + n];
            P[kOffB3 + n] = s;
        }
    }
    __syncthreads();
    // dW2[n][m] = sum_i dZ2[i][n] h1[i][m]; db2
    {
        float s = 0.0f;
        for (int i = 0; i < kTrainAtoms; i++) s += z2s[i * kH + n];
        P[kOffB2 + n] = s;
        for (int m0 = 0; m0 < kH; m0 += 32) {
            float acc[32];
#pragma unroll
            for (int mm = 0; mm < 32; mm++) acc[mm] = 0.0f;
            for (int i = 0; i < kTrainAtoms; i++) {
                const float d = z2s[i * kH + n];
#pragma unroll
                for (int mm = 0; mm < 32; mm++) acc[mm] += d * bf2f(h1s[i * kLdH + m0 + mm]);
            }
#pragma unroll
            for (int mm = 0; mm < 32; mm++) P[kOffW2 + n * kH + m0 + mm] = acc[mm];
        }
    }
    // dZ1[i][m] = (sum_n dZ2[i][n] W2[n][m]) * silu'(z1[i][m]), m = this thread's unit
    {
        const int m = t;
        float acc[kTrainAtoms];
#pragma unroll
        for (int i = 0; i < kTrainAtoms; i++) acc[i] = 0.0f;
        for (int nn = 0; nn < kH; nn++) {
            const float wv = bf2f(W2[nn * kH + m]);
#pragma unroll
            for (int i = 0; i < kTrainAtoms; i++) acc[i] += z2s[i * kH + nn] * wv;
        }
#pragma unroll
        for (int i = 0; i < kTrainAtoms; i++) z1s[i * kH + m] = acc[i] * dsilu(z1s[i * kH + m]);
    }
    __syncthreads();
    // dW1[n][k] = sum_i dZ1[i][n] x[i][k]; db1
    {
        float s = 0.0f;
        for (int i = 0; i < kTrainAtoms; i++) s += z1s[i * kH + n];
        P[kOffB1 + n] = s;
        for (int k0 = 0; k0 < kNF; k0 += 32) {
            float acc[32];
#pragma unroll
            for (int kk = 0; kk < 32; kk++) acc[kk] = 0.0f;
            for (int i = 0; i < kTrainAtoms; i++) {
                const float d = z1s[i * kH + n];
#pragma unroll
                for (int kk = 0; kk < 32; kk++) acc[kk] += d * bf2f(xs[i * kLdX + k0 + kk]);
            }
#pragma unroll
            for (int kk = 0; kk < 32; kk++) P[kOffW1 + n * kNF + k0 + kk] = acc[kk];
        }
    }
    if (t == 0) {
        double s = 0.0, s0 = 0.0, st = 0.0;
        for (int i = 0; i < kTrainAtoms; i++) {
            s += lossv[i];
            s0 += lossv0[i];
            st += lossvt[i];
        }
        T.loss_partial[3 * blockIdx.x] = s;
        T.loss_partial[3 * blockIdx.x + 1] = s0;
        T.loss_partial[3 * blockIdx.x + 2] = st;
    }
}

__device__ inline int bf_index(int p) {
    if (p < kOffB1) return kBfW1 + p;
    if (p >= kOffW2 && p < kOffB2) return kBfW2 + (p - kOffW2);
    if (p >= kOffW3 && p < kOffB3) return kBfW3 + (p - kOffW3);
    return -1;
}

// Gradient = fixed-order sum of the workgroup partials.  mode 0 writes it to
// grad; mode 1 takes one Adam step on the fp32 master parameters and
// refreshes their BF16 copies; mode 2 only reports the loss.  Thread 0 writes
// the batch loss with the network, with o = 0 and with the Taylor
// correction to loss_out[0..2].
__global__ void k_reduce_adam(int nwg, const float *__restrict__ partial, const double *loss_partial,
                              double inv_nb, float *loss_out, int mode, float *grad, float *param,
                              float *m1, float *m2, unsigned short *wbf, float lr, float beta1,
                              float beta2, float c1, float c2, float eps) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && loss_out) {
        double s = 0.0, s0 = 0.0, st = 0.0;
        for (int w = 0; w < nwg; w++) {
            s += loss_partial[3 * w];
            s0 += loss_partial[3 * w + 1];
            st += loss_partial[3 * w + 2];
        }
        loss_out[0] = (float)(s * inv_nb);
        loss_out[1] = (float)(s0 * inv_nb);
        loss_out[2] = (float)(st * inv_nb);
    }
    if (p >= kNParam || mode == 2) return;
    float g = 0.0f;
    for (int w = 0; w < nwg; w++) g += partial[(size_t)w * kNParam + p];
    if (mode == 0) {
        grad[p] = g;
        return;
    }
    const float a = beta1 * m1[p] + (1.0f - beta1) * g;
    const float b = beta2 * m2[p] + (1.0f - beta2) * g * g;
    m1[p] = a;
    m2[p] = b;
    const float th = param[p] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    param[p] = th;
    const int q = bf_index(p);
    if (q >= 0) wbf[q] = f2bf(th);
}

__global__ void k_refresh_bf16(const float *param, unsigned short *wbf) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNParam) return;
    const int q = bf_index(p);
    if (q >= 0) wbf[q] = f2bf(param[p]);
}

// ---- multiple-time-step driver steps ---------------------------------------

__device__ inline double min_img(double d, double L) { return d - L * round(d / L); }

// Exact step t0: F = F_inner + F_shell (both evaluated directly); remember
// the shell force, the inner force and the positions.
__global__ void k_mts_t0(int64_t n, const double *fin, const double *fsh, double *fin0, double *fsh0,
                         const double *x, const double *y, const double *z, double *x0, double *fx,
                         double *fy, double *fz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *out[3] = {fx, fy, fz};
    const double p[3] = {x[i], y[i], z[i]};
    for (int d = 0; d < 3; d++) {
        const double a = fin[d * n + i], b = fsh[d * n + i];
        fin0[d * n + i] = a;
        fsh0[d * n + i] = b;
        x0[d * n + i] = p[d];
        out[d][i] = a + b;
    }
}

// -K_i dx for atom i (K symmetric, stored xx, yy, zz, xy, xz, yz).
__device__ inline void minus_k_dx(const float *k, const double *dx, double *out) {
    out[0] = -((double)k[0] * dx[0] + (double)k[3] * dx[1] + (double)k[4] * dx[2]);
    out[1] = -((double)k[3] * dx[0] + (double)k[1] * dx[1] + (double)k[5] * dx[2]);
    out[2] = -((double)k[4] * dx[0] + (double)k[5] * dx[1] + (double)k[2] * dx[2]);
}

// Inner step: F = F_inner + (1 + o1) F_shell(t0) - o2 s2 K dx + o3 s3 (F_inner - F_inner(t0)).
// mode 0: o = 0; 1: o from coef (n x 3); 2: o = (0, 1/s2, 0).
__global__ void k_mts_combine(MtsArgs M, int mode, const float *coef, double *fx, double *fy, double *fz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M.n) return;
    const int64_t n = M.n;
    double o1 = 0.0, o2 = 0.0, o3 = 0.0;
    if (mode == 1) {
        o1 = coef[i * 3];
        o2 = coef[i * 3 + 1];
        o3 = coef[i * 3 + 2];
    } else if (mode == 2) {
        o2 = 1.0 / M.s2;
    }
    const double p[3] = {M.x[i], M.y[i], M.z[i]};
    double dx[3], kd[3];
    for (int d = 0; d < 3; d++) dx[d] = min_img(p[d] - M.x0[d * n + i], M.box[d]);
    minus_k_dx(M.kmat + i * 6, dx, kd);
    double *out[3] = {fx, fy, fz};
    for (int d = 0; d < 3; d++) {
        const double fi = M.fin[d * n + i];
        const double dfi = fi - M.fin0[d * n + i];
        out[d][i] = fi + (1.0 + o1) * M.fsh0[d * n + i] + o2 * M.s2 * kd[d] + o3 * M.s3 * dfi;
    }
}

// Training frame: normalised basis and target (the shell force at t) for
// batch atoms (start + b) mod n.
__global__ void k_train_prep(MtsArgs M, const double *fsh, int64_t start, int64_t nb, double inv_sigma,
                             float *basis, float *target) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int64_t n = M.n, i = (start + b) % n;
    const double p[3] = {M.x[i], M.y[i], M.z[i]};
    double dx[3], kd[3];
    for (int d = 0; d < 3; d++) dx[d] = min_img(p[d] - M.x0[d * n + i], M.box[d]);
    minus_k_dx(M.kmat + i * 6, dx, kd);
    for (int d = 0; d < 3; d++) {
        const double fi = M.fin[d * n + i];
        basis[b * 9 + d] = (float)(M.fsh0[d * n + i] * inv_sigma);
        basis[b * 9 + 3 + d] = (float)(M.s2 * kd[d] * inv_sigma);
        basis[b * 9 + 6 + d] = (float)(M.s3 * (fi - M.fin0[d * n + i]) * inv_sigma);
        target[b * 3 + d] = (float)(fsh[d * n + i] * inv_sigma);
    }
}

__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Standard normal number for (seed, step, atom, axis): Box-Muller on a counter hash.
__device__ inline double gauss(uint64_t seed, uint64_t step, int64_t i, int d) {
    const uint64_t h1 = mix64(seed ^ mix64(step * 3 + d) ^ mix64(0x5bd1e995ull + (uint64_t)i));
    const uint64_t h2 = mix64(h1);
    const double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53;
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// Seeded Gaussian displacement of every coordinate.
__global__ void k_random_walk(int64_t n, double *x, double *y, double *z, double sigma, uint64_t seed,
                              uint64_t step) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *p[3] = {x, y, z};
    for (int d = 0; d < 3; d++) p[d][i] += sigma * gauss(seed, step, i, d);
}

// Overdamped (Brownian) step: dx = mu F + sigma xi, each component clamped to
// [-cap, cap].  F = fa (+ fb when given), atom-order SoA arrays.
__global__ void k_bd_step(int64_t n, double *x, double *y, double *z, const double *fax, const double *fay,
                          const double *faz, const double *fb, double mu, double sigma, double cap,
                          uint64_t seed, uint64_t step) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *p[3] = {x, y, z};
    const double fa[3] = {fax[i], fay[i], faz[i]};
    for (int d = 0; d < 3; d++) {
        double f = fa[d] + (fb ? fb[d * n + i] : 0.0);
        double dx = mu * f + (sigma > 0.0 ? sigma * gauss(seed, step, i, d) : 0.0);
        dx = fmin(fmax(dx, -cap), cap);
        p[d][i] += dx;
    }
}

inline unsigned grid_of(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

namespace gm {
namespace sur {

hipError_t launch_feat_prep(gm_sys *s, const double *x, const double *y, const double *z,
                            const uint8_t *cls, int4 *qs) {
    const int n_slots = (int)(s->nsc * 64);
    if (n_slots == 0) return hipSuccess;
    k_feat_prep<<<grid_of(n_slots, 256), 256, 0, s->stream>>>(n_slots, s->slot_atom.p, x, y, z, cls,
                                                               s->bp.inv_q[0], s->bp.inv_q[1],
                                                               s->bp.inv_q[2], qs);
    return hipGetLastError();
}

hipError_t launch_features(gm_sys *s, const FeatArgs &A) {
    if (s->nsc == 0) return hipSuccess;
    k_features<<<(unsigned)s->nsc, 64 * kFeatWaves, 0, s->stream>>>(A, (int)s->nsc);
    return hipGetLastError();
}

hipError_t launch_infer(const NetArgs &N, float *out, hipStream_t st) {
    if (N.n == 0) return hipSuccess;
    k_infer<<<grid_of(N.n, kInferAtoms), 256, 0, st>>>(N, out);
    return hipGetLastError();
}

hipError_t launch_train(const TrainArgs &T, hipStream_t st) {
    if (T.nb == 0) return hipSuccess;
    k_train<<<grid_of(T.nb, kTrainAtoms), 128, 0, st>>>(T);
    return hipGetLastError();
}

int train_workgroups(int64_t nb) { return (int)((nb + kTrainAtoms - 1) / kTrainAtoms); }

hipError_t launch_reduce_adam(int nwg, const float *partial, const double *loss_partial, double inv_nb,
                              float *loss_out, int mode, float *grad, float *param, float *m1,
                              float *m2, unsigned short *wbf, float lr, float beta1, float beta2,
                              float c1, float c2, float eps, hipStream_t st) {
    k_reduce_adam<<<grid_of(kNParam, 256), 256, 0, st>>>(nwg, partial, loss_partial, inv_nb, loss_out,
                                                          mode, grad, param, m1, m2, wbf, lr, beta1,
                                                          beta2, c1, c2, eps);
    return hipGetLastError();
}

hipError_t launch_refresh_bf16(const float *param, unsigned short *wbf, hipStream_t st) {
    k_refresh_bf16<<<grid_of(kNParam, 256), 256, 0, st>>>(param, wbf);
    return hipGetLastError();
}

hipError_t launch_mts_t0(int64_t n, const double *fin, const double *fsh, double *fin0, double *fsh0,
                         const double *x, const double *y, const double *z, double *x0, double *fx,
                         double *fy, double *fz, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_mts_t0<<<grid_of(n, 256), 256, 0, st>>>(n, fin, fsh, fin0, fsh0, x, y, z, x0, fx, fy, fz);
    return hipGetLastError();
}

hipError_t launch_mts_combine(const MtsArgs &M, int mode, const float *coef, double *fx, double *fy, double *fz,
                              hipStream_t st) {
    if (M.n == 0) return hipSuccess;
    k_mts_combine<<<grid_of(M.n, 256), 256, 0, st>>>(M, mode, coef, fx, fy, fz);
    return hipGetLastError();
}

hipError_t launch_train_prep(const MtsArgs &M, const double *fsh, int64_t start, int64_t nb, double inv_sigma,
                             float *basis, float *target, hipStream_t st) {
    if (nb == 0) return hipSuccess;
    k_train_prep<<<grid_of(nb, 256), 256, 0, st>>>(M, fsh, start, nb, inv_sigma, basis, target);
    return hipGetLastError();
}

hipError_t launch_bd_step(int64_t n, double *x, double *y, double *z, const double *fax, const double *fay,
                          const double *faz, const double *fb, double mu, double sigma, double cap,
                          uint64_t seed, uint64_t step, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_bd_step<<<grid_of(n, 256), 256, 0, st>>>(n, x, y, z, fax, fay, faz, fb, mu, sigma, cap, seed, step);
    return hipGetLastError();
}

hipError_t launch_random_walk(int64_t n, double *x, double *y, double *z, double sigma, uint64_t seed,
                              uint64_t step, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_random_walk<<<grid_of(n, 256), 256, 0, st>>>(n, x, y, z, sigma, seed, step);
    return hipGetLastError();
}

}  // namespace sur
}  // namespace gm
