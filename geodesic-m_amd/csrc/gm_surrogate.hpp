// Shell-force surrogate: sizes, parameter layout and kernel arguments shared
// by the kernels (gm_surrogate.hip) and the host driver (gm_surrogate.cpp).
//
// Split.  The exact model's switched LJ V S(r) is the sum of an inner part
// V S_in(r) (the inner system: same atoms, LJ cut at r_in_c, switched from
// r_in_sw) and a shell part V (S(r) - S_in(r)), which the exact system
// evaluates directly in shell mode (no cancellation of large short-range
// forces).  Bonded terms belong to the inner part.
//
// Model.  Per atom i, 96 radial features phi_i[c][k] = sum over atoms j of
// class c with r_lo <= r_ij <= r_hi of env(r_ij) exp(-(r_ij - mu_k)^2 / 2w^2)
// (24 centres mu_k evenly spaced over [r_lo, r_hi], w = spacing; env is a
// quintic fade over delta at both window edges) are standardised and fed to a
// 96 -> 128 -> 128 -> 3 SiLU network (BF16 weights and activations, FP32
// accumulate).  Its three outputs o weight three equivariant vectors:
//   F_shell(t) ~ (1 + o1) F_shell(t0) - o2 s2 K_i dx(t) + o3 s3 dF_inner(t)
// with dx the displacement since the last exact step t0, K_i the shell
// potential's Hessian with respect to atom i's own position at t0 (summed
// over its shell pairs, computed with the features), dF_inner the change of
// the inner force since t0, and s2, s3 fixed scales.  o = 0 is the plain
// multiple-time-step method (shell force held from t0); o = (0, 1/s2, 0) is
// the first-order Taylor correction in the atom's own displacement.
#pragma once

#include "gm_internal.hpp"

namespace gm {
namespace sur {

constexpr int kNG = 24;             // Gaussians per class
constexpr int kNC = 4;              // atom classes
constexpr int kNF = kNG * kNC;      // 96 features
constexpr int kH = 128;             // hidden width
constexpr int kNO = 3;              // outputs

// fp32 master parameters, flat
constexpr int kOffW1 = 0;                    // [kH][kNF]
constexpr int kOffB1 = kOffW1 + kH * kNF;    // [kH]
constexpr int kOffW2 = kOffB1 + kH;          // [kH][kH]
constexpr int kOffB2 = kOffW2 + kH * kH;     // [kH]
constexpr int kOffW3 = kOffB2 + kH;          // [kNO][kH]
constexpr int kOffB3 = kOffW3 + kNO * kH;    // [kNO]
constexpr int kNParam = kOffB3 + kNO;        // 29315

// BF16 weight copies read by the MFMA layers; W3 is padded to 16 rows.
constexpr int kBfW1 = 0;                     // [kH][kNF]
constexpr int kBfW2 = kBfW1 + kH * kNF;      // [kH][kH]
constexpr int kBfW3 = kBfW2 + kH * kH;       // [16][kH], rows >= kNO zero
constexpr int kNBf = kBfW3 + 16 * kH;

constexpr int kFeatWaves = 8;     // features kernel: wavefronts per supercluster (2 per class)
constexpr int kInferAtoms = 64;   // atoms per inference workgroup (4 wavefronts x 16)
constexpr int kTrainAtoms = 32;   // atoms per training workgroup (2 wavefronts x 16)

struct FeatArgs {
    const int4 *qs;                  // per slot: fixed-point x, y, z and class (-1: empty slot)
    const float2 *sp;                // per slot: (sigma/2, sqrt(eps))
    const int *slot_atom;
    const int *sc_col, *col_sc;  // col_sc[c] .. col_sc[c + 1]: column c's superclusters
    const int4 *cl_ref;
    const float4 *cl_half;
    const int4 *sc_ref;
    const float4 *sc_half;
    BuildParams bp;                  // the exact system's build parameters (reach, boxes)
    float r_lo2, r_hi2, r_lo, r_hi;  // feature window
    float inv_delta;                 // 1 / edge fade width
    float mu0, inv_w;                // first centre, 1 / Gaussian width
    float *feat;                     // atom order, n x kNF
    // shell Hessian K_i: pairs with rk_lo <= r <= rk_hi; shell weight
    // W(r) = S((r - a1) / w1) - S((r - a2) / w2), S the quintic switch
    float rk_lo2, rk_hi2;
    float a1, inv_w1, a2, inv_w2;
    const int64_t *excl_off;         // exclusion set of the exact system (pairs skipped in K)
    const uint32_t *excl_nbr;
    const uint2 *excl_rng;
    const uint2 *excl_rng_any;       // per atom: smallest and largest partner on either side
    float *kmat;                     // atom order, n x 6: xx, yy, zz, xy, xz, yz
};

struct NetArgs {
    const float *feat;               // atom order, n x kNF (raw features)
    const float *fmean, *finv;       // standardisation: (phi - mean) * inv
    const unsigned short *wbf;       // BF16 weights (kBf* layout)
    const float *param;              // fp32 master parameters (biases read from here)
    int64_t n;                       // atoms
};

struct TrainArgs {
    NetArgs net;
    int64_t start;                   // batch = atoms (start + b) mod n, b < nb
    int64_t nb;
    const float *basis;              // batch order, nb x 9: B1, B2, B3 (each x, y, z), normalised
    const float *target;             // batch order, nb x 3, normalised
    float taylor;                    // o2 of the Taylor correction (1 / s2)
    float *partial;                  // per workgroup kNParam gradient partials
    double *loss_partial;            // per workgroup: sums of squared residuals (network, o = 0, Taylor)
};

// Elementwise steps of the multiple-time-step driver.  Force and position
// buffers with "3n" are SoA blocks: x [0, n), y [n, 2n), z [2n, 3n).
struct MtsArgs {
    int64_t n;
    const double *fin, *fin0, *fsh0, *x0;   // 3n each
    const float *kmat;                      // n x 6, shell Hessian at t0
    const double *x, *y, *z;                // current positions (atom order)
    double box[3];
    double s2, s3;                          // basis scales
};

// launchers (gm_surrogate.hip)
hipError_t launch_feat_prep(gm_sys *s, const double *x, const double *y, const double *z,
                            const uint8_t *cls, int4 *qs);
hipError_t launch_features(gm_sys *s, const FeatArgs &A);
hipError_t launch_infer(const NetArgs &N, float *out, hipStream_t st);
hipError_t launch_train(const TrainArgs &T, hipStream_t st);
int train_workgroups(int64_t nb);
hipError_t launch_reduce_adam(int nwg, const float *partial, const double *loss_partial, double inv_nb,
                              float *loss_out, int mode, float *grad, float *param, float *m1,
                              float *m2, unsigned short *wbf, float lr, float beta1, float beta2,
                              float c1, float c2, float eps, hipStream_t st);
hipError_t launch_refresh_bf16(const float *param, unsigned short *wbf, hipStream_t st);
hipError_t launch_mts_t0(int64_t n, const double *fin, const double *fsh, double *fin0, double *fsh0,
                         const double *x, const double *y, const double *z, double *x0, double *fx,
                         double *fy, double *fz, hipStream_t st);
// mode 0: shell force held (o = 0); 1: network coefficients; 2: Taylor (o = (0, 1/s2, 0))
hipError_t launch_mts_combine(const MtsArgs &M, int mode, const float *coef, double *fx, double *fy, double *fz,
                              hipStream_t st);
hipError_t launch_train_prep(const MtsArgs &M, const double *fsh, int64_t start, int64_t nb, double inv_sigma,
                             float *basis, float *target, hipStream_t st);
hipError_t launch_bd_step(int64_t n, double *x, double *y, double *z, const double *fax, const double *fay,
                          const double *faz, const double *fb, double mu, double sigma, double cap,
                          uint64_t seed, uint64_t step, hipStream_t st);
hipError_t launch_random_walk(int64_t n, double *x, double *y, double *z, double sigma, uint64_t seed,
                              uint64_t step, hipStream_t st);

}  // namespace sur
}  // namespace gm
