/*
 * gm_sur_oracle.c -- f64 CPU restatement of the shell-force surrogate.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the surrogate's features kernel,
 * network forward pass and hand-written backward pass.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * The surrogate has no counterpart in the reference, so nothing here follows
 * a reference file.  The definitions are the ones stated in
 * include/geodesic_hip.h (gm_surrogate_*):
 *   - features: phi_i[c][k] = sum over j != i of class c with
 *     r_lo <= r_ij <= r_hi of env(r_ij) exp(-(r_ij - mu_k)^2 / (2 w^2)),
 *     mu_k = r_lo + k w, w = (r_hi - r_lo) / 23, env = P((r - r_lo)/fade)
 *     P((r_hi - r)/fade), P(u) = u^3 (10 - 15u + 6u^2) on [0, 1];
 *     r_ij is the minimum-image distance d - L round(d / L);
 *   - network: x = (phi - mean) * inv_std, h1 = SiLU(W1 x + b1),
 *     h2 = SiLU(W2 h1 + b2), o = W3 h2 + b3;
 *   - loss = (1/n) sum_i |(1 + o1) B1 + o2 B2 + o3 B3 - T|^2;
 *   - shell pair potential u(r) = V_LJ(r) (S((r - a1)/w1) - S((r - a2)/w2))
 *     (exact switch minus inner switch, S the quintic switch), its force and
 *     its Hessian summed per atom (K).
 * With `round_bf16` set, x, the weights W1..W3 and the activations h1, h2
 * are rounded to BF16 (round to nearest even, via f32) where the device
 * rounds them; the backward pass then uses those rounded values and treats
 * the rounding as the identity, as the device does.
 *
 * Pins (tests/test_oracle_surrogate.py): the gradient against central finite
 * differences of the loss with rounding off; features against a direct numpy
 * evaluation on small systems, a two-atom closed form, permutation and
 * translation invariance; the forward pass against a numpy restatement; the
 * shell force against the golden-pinned pair routine run with the exact and
 * the inner switch (difference); K against central differences of the shell
 * force.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NG 24
#define NC 4
#define NF 96
#define NH 128
#define NO 3
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + NH * NF)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NO * NH)
#define NPARAM (OFF_B3 + NO)

static double min_image(double d, double L) { return d - L * round(d / L); }

static double fade(double u) {
    if (u < 0.0) u = 0.0;
    if (u > 1.0) u = 1.0;
    return u * u * u * (10.0 + u * (-15.0 + 6.0 * u));
}

int64_t or_sur_nparam(void) { return NPARAM; }

/* Brute force over all ordered pairs, O(n^2). */
void or_sur_features(int64_t n, const double *x, const double *y, const double *z, const double *box,
                     const uint8_t *cls, double r_lo, double r_hi, double fade_w, double *out) {
    const double w = (r_hi - r_lo) / (NG - 1);
    memset(out, 0, sizeof(double) * (size_t)n * NF);
    for (int64_t i = 0; i < n; i++) {
        double *f = out + (size_t)i * NF;
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            const double dx = min_image(x[j] - x[i], box[0]);
            const double dy = min_image(y[j] - y[i], box[1]);
            const double dz = min_image(z[j] - z[i], box[2]);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < r_lo * r_lo || r2 > r_hi * r_hi) continue;
            const double r = sqrt(r2);
            const double env = fade((r - r_lo) / fade_w) * fade((r_hi - r) / fade_w);
            const double t = (r - r_lo) / w;
            for (int k = 0; k < NG; k++) {
                const double d = t - k;
                f[cls[j] * NG + k] += env * exp(-0.5 * d * d);
            }
        }
    }
}

static double bf16(double v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double silu(double z) { return z / (1.0 + exp(-z)); }
static double dsilu(double z) {
    const double s = 1.0 / (1.0 + exp(-z));
    return s * (1.0 + z * (1.0 - s));
}

typedef struct {
    double x[NF], z1[NH], h1[NH], z2[NH], h2[NH], o[NO];
} fwd_t;

/* Weights as the device uses them (rounded or not). */
static void weights(const float *param, int round_bf16, double *w) {
    for (int p = 0; p < NPARAM; p++) {
        const int is_w = p < OFF_B1 || (p >= OFF_W2 && p < OFF_B2) || (p >= OFF_W3 && p < OFF_B3);
        w[p] = (round_bf16 && is_w) ? bf16(param[p]) : (double)param[p];
    }
}

static void forward_one(const float *feat, const float *fmean, const float *finv, const double *w,
                        int round_bf16, fwd_t *s) {
    for (int k = 0; k < NF; k++) {
        /* the device standardises in f32 */
        const float v = (feat[k] - fmean[k]) * finv[k];
        s->x[k] = round_bf16 ? bf16(v) : (double)v;
    }
    for (int n = 0; n < NH; n++) {
        double a = 0.0;
        for (int k = 0; k < NF; k++) a += w[OFF_W1 + n * NF + k] * s->x[k];
        s->z1[n] = a + w[OFF_B1 + n];
        const double h = silu(s->z1[n]);
        s->h1[n] = round_bf16 ? bf16(h) : h;
    }
    for (int n = 0; n < NH; n++) {
        double a = 0.0;
        for (int m = 0; m < NH; m++) a += w[OFF_W2 + n * NH + m] * s->h1[m];
        s->z2[n] = a + w[OFF_B2 + n];
        const double h = silu(s->z2[n]);
        s->h2[n] = round_bf16 ? bf16(h) : h;
    }
    for (int j = 0; j < NO; j++) {
        double a = 0.0;
        for (int m = 0; m < NH; m++) a += w[OFF_W3 + j * NH + m] * s->h2[m];
        s->o[j] = a + w[OFF_B3 + j];
    }
}

void or_sur_forward(int64_t n, const float *feat, const float *fmean, const float *finv, const float *param,
                    int round_bf16, double *out) {
    double *w = (double *)malloc(sizeof(double) * NPARAM);
    fwd_t s;
    weights(param, round_bf16, w);
    for (int64_t i = 0; i < n; i++) {
        forward_one(feat + (size_t)i * NF, fmean, finv, w, round_bf16, &s);
        for (int j = 0; j < NO; j++) out[(size_t)i * NO + j] = s.o[j];
    }
    free(w);
}

/* Loss and its gradient with respect to the fp32 parameters. */
double or_sur_loss_grad(int64_t n, const float *feat, const float *fmean, const float *finv, const float *param,
                        const float *basis, const float *target, int round_bf16, double *grad) {
    double *w = (double *)malloc(sizeof(double) * NPARAM);
    weights(param, round_bf16, w);
    memset(grad, 0, sizeof(double) * NPARAM);
    double loss = 0.0;
    fwd_t s;
    double d_h2[NH], d_z2[NH], d_h1[NH], d_z1[NH];
    for (int64_t i = 0; i < n; i++) {
        forward_one(feat + (size_t)i * NF, fmean, finv, w, round_bf16, &s);
        const float *B = basis + (size_t)i * 9, *T = target + (size_t)i * 3;
        double d_o[NO] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; d++) {
            const double fh = (1.0 + s.o[0]) * B[d] + s.o[1] * B[3 + d] + s.o[2] * B[6 + d];
            const double r = fh - T[d];
            loss += r * r;
            d_o[0] += 2.0 * r * B[d] / (double)n;
            d_o[1] += 2.0 * r * B[3 + d] / (double)n;
            d_o[2] += 2.0 * r * B[6 + d] / (double)n;
        }
        for (int j = 0; j < NO; j++) {
            grad[OFF_B3 + j] += d_o[j];
            for (int m = 0; m < NH; m++) grad[OFF_W3 + j * NH + m] += d_o[j] * s.h2[m];
        }
        for (int m = 0; m < NH; m++) {
            double g = 0.0;
            for (int j = 0; j < NO; j++) g += d_o[j] * w[OFF_W3 + j * NH + m];
            d_h2[m] = g;
            d_z2[m] = g * dsilu(s.z2[m]);
        }
        for (int nn = 0; nn < NH; nn++) {
            grad[OFF_B2 + nn] += d_z2[nn];
            for (int m = 0; m < NH; m++) grad[OFF_W2 + nn * NH + m] += d_z2[nn] * s.h1[m];
        }
        for (int m = 0; m < NH; m++) {
            double g = 0.0;
            for (int nn = 0; nn < NH; nn++) g += d_z2[nn] * w[OFF_W2 + nn * NH + m];
            d_h1[m] = g;
            d_z1[m] = g * dsilu(s.z1[m]);
        }
        for (int nn = 0; nn < NH; nn++) {
            grad[OFF_B1 + nn] += d_z1[nn];
            for (int k = 0; k < NF; k++) grad[OFF_W1 + nn * NF + k] += d_z1[nn] * s.x[k];
        }
    }
    free(w);
    (void)d_h2;
    (void)d_h1;
    return loss / (double)n;
}

/* Quintic switch S(u) and its derivatives with respect to r for
 * u = (r - a) / w, constant outside [0, 1]. */
static void sw_r(double r, double a, double w, double *S, double *S1, double *S2) {
    double u = (r - a) / w;
    if (u <= 0.0) { *S = 1.0; *S1 = 0.0; *S2 = 0.0; return; }
    if (u >= 1.0) { *S = 0.0; *S1 = 0.0; *S2 = 0.0; return; }
    const double om = 1.0 - u;
    *S = 1.0 - 10.0 * u * u * u + 15.0 * u * u * u * u - 6.0 * u * u * u * u * u;
    *S1 = -30.0 * u * u * om * om / w;
    *S2 = -60.0 * u * om * (1.0 - 2.0 * u) / (w * w);
}

/* Shell force and shell Hessian over a pair list (unordered, non-excluded
 * pairs): pair potential u(r) = V_LJ(r) (S((r - a1)/w1) - S((r - a2)/w2)) for
 * r <= rc, Lorentz-Berthelot sigma/epsilon, minimum-image d = x_j - x_i.
 * F_i += u'(r) d / r, F_j -= u'(r) d / r; K_i and K_j += u'' dd^T/r^2 +
 * (u'/r)(I - dd^T/r^2) (xx, yy, zz, xy, xz, yz). */
void or_sur_shell(int64_t n, const double *x, const double *y, const double *z, const double *sigma,
                  const double *eps, int64_t npairs, const uint32_t *pi, const uint32_t *pj, const double *box,
                  double a1, double w1, double a2, double w2, double rc, double *fx, double *fy, double *fz,
                  double *K) {
    memset(fx, 0, sizeof(double) * (size_t)n);
    memset(fy, 0, sizeof(double) * (size_t)n);
    memset(fz, 0, sizeof(double) * (size_t)n);
    memset(K, 0, sizeof(double) * 6 * (size_t)n);
    for (int64_t p = 0; p < npairs; p++) {
        const uint32_t i = pi[p], j = pj[p];
        const double dx = min_image(x[j] - x[i], box[0]);
        const double dy = min_image(y[j] - y[i], box[1]);
        const double dz = min_image(z[j] - z[i], box[2]);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 == 0.0 || r2 > rc * rc) continue;
        const double r = sqrt(r2);
        const double sig = 0.5 * (sigma[i] + sigma[j]), ep = sqrt(eps[i] * eps[j]);
        const double q = sig / r, q6 = q * q * q * q * q * q, q12 = q6 * q6;
        const double V = 4.0 * ep * (q12 - q6);
        const double V1 = -24.0 * ep * (2.0 * q12 - q6) / r;
        const double V2 = 24.0 * ep * (26.0 * q12 - 7.0 * q6) / r2;
        double Sa, Sa1, Sa2, Sb, Sb1, Sb2;
        sw_r(r, a1, w1, &Sa, &Sa1, &Sa2);
        sw_r(r, a2, w2, &Sb, &Sb1, &Sb2);
        const double W = Sa - Sb, W1 = Sa1 - Sb1, W2 = Sa2 - Sb2;
        const double du = V1 * W + V * W1;
        const double d2u = V2 * W + 2.0 * V1 * W1 + V * W2;
        fx[i] += du * dx / r; fy[i] += du * dy / r; fz[i] += du * dz / r;
        fx[j] -= du * dx / r; fy[j] -= du * dy / r; fz[j] -= du * dz / r;
        const double b = du / r, a = (d2u - b) / r2;
        const double h[6] = {a * dx * dx + b, a * dy * dy + b, a * dz * dz + b, a * dx * dy, a * dx * dz, a * dy * dz};
        for (int m = 0; m < 6; m++) {
            K[6 * (size_t)i + m] += h[m];
            K[6 * (size_t)j + m] += h[m];
        }
    }
}
