// Shell-force surrogate and 4:1 multiple-time-step driver: host side.  No
// reference counterpart.  The model and the parameter layout are described in
// gm_surrogate.hpp and include/geodesic_hip.h.
//
// The surrogate owns two systems on one stream: the exact model (as the
// caller describes it) and the inner model (same atoms and topology, short
// LJ cutoff, own list).  The shell force is the exact system evaluated in
// shell mode (each pair weighted by S - S_in), so exact = inner + shell holds
// by construction and no large short-range force is ever subtracted.
// Everything is ordered on the one stream, so the driver needs no
// cross-stream synchronisation.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "gm_surrogate.hpp"

using namespace gm;
using namespace gm::sur;

struct gm_surrogate {
    gm_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    gm_sys *ex = nullptr, *in = nullptr;
    int64_t n = 0;
    double box[3] = {0, 0, 0};
    double r_lo = 0, r_hi = 0, fade = 0;
    double ex_rsw = 0, ex_rc = 0, in_rsw = 0, in_rc = 0;  // switch windows of the split
    DevBuf<uint8_t> cls;
    DevBuf<int4> qs;
    DevBuf<float> feat, kmat;
    DevBuf<float> param, m1, m2, grad;
    DevBuf<unsigned short> wbf;
    DevBuf<float> fmean, finv;
    std::vector<float> h_fmean, h_finv;
    double s2 = 1.0, s3 = 1.0, sigma = 1.0;
    bool norm_set = false, scales_set = false;
    int64_t adam_t = 0;
    // driver state (3n SoA blocks)
    DevBuf<double> fsh, fin, fin0, fsh0, x0, ftot;
    DevBuf<float> coef;
    DevBuf<float> basis, target, partial, loss_hist;
    DevBuf<double> lossp;
    DevBuf<float> h_in;  // staging for host-array entry points
    int64_t mts_k = 0;
};

#define GM_REQUIRE(cond, msg)                                   \
    do {                                                        \
        if (!(cond)) return set_error(GM_ERR_INVALID_ARG, msg); \
    } while (0)

namespace {

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform(-a, a) weights per layer (a = sqrt(6 / (fan_in + fan_out)); the
// output layer is scaled by 0.1 so an untrained network starts near o = 0);
// zero biases.
std::vector<float> init_params(uint64_t seed) {
    std::vector<float> p((size_t)kNParam, 0.0f);
    uint64_t st = seed;
    auto fill = [&](int off, int rows, int cols, double scale) {
        const double a = scale * std::sqrt(6.0 / (double)(rows + cols));
        for (int k = 0; k < rows * cols; k++) {
            const double u = (double)(splitmix(st) >> 11) * 0x1.0p-53;
            p[(size_t)off + k] = (float)((2.0 * u - 1.0) * a);
        }
    };
    fill(kOffW1, kH, kNF, 1.0);
    fill(kOffW2, kH, kH, 1.0);
    fill(kOffW3, kNO, kH, 0.1);
    return p;
}

gm_status alloc_state(gm_surrogate *s) {
    const size_t n = (size_t)std::max<int64_t>(s->n, 1);
    GM_HIP(s->feat.ensure(n * kNF));
    GM_HIP(s->kmat.ensure(n * 6));
    for (DevBuf<double> *b : {&s->fsh, &s->fin, &s->fin0, &s->fsh0, &s->x0, &s->ftot}) GM_HIP(b->ensure(3 * n));
    GM_HIP(s->coef.ensure(3 * n));
    GM_HIP(s->param.ensure(kNParam));
    GM_HIP(s->m1.ensure(kNParam));
    GM_HIP(s->m2.ensure(kNParam));
    GM_HIP(s->grad.ensure(kNParam));
    GM_HIP(s->wbf.ensure(kNBf));
    GM_HIP(s->fmean.ensure(kNF));
    GM_HIP(s->finv.ensure(kNF));
    GM_HIP(s->loss_hist.ensure(3));
    return GM_OK;
}

gm_status upload_params(gm_surrogate *s, const float *p) {
    GM_HIP(hipMemcpyAsync(s->param.p, p, sizeof(float) * kNParam, hipMemcpyHostToDevice, s->stream));
    // padded W3 rows stay zero; the refresh writes every real weight
    GM_HIP(hipMemsetAsync(s->wbf.p, 0, sizeof(unsigned short) * kNBf, s->stream));
    GM_HIP(launch_refresh_bf16(s->param.p, s->wbf.p, s->stream));
    GM_HIP(hipMemsetAsync(s->m1.p, 0, sizeof(float) * kNParam, s->stream));
    GM_HIP(hipMemsetAsync(s->m2.p, 0, sizeof(float) * kNParam, s->stream));
    s->adam_t = 0;
    return GM_OK;
}

gm_status upload_norm(gm_surrogate *s) {
    GM_HIP(hipMemcpyAsync(s->fmean.p, s->h_fmean.data(), sizeof(float) * kNF, hipMemcpyHostToDevice, s->stream));
    GM_HIP(hipMemcpyAsync(s->finv.p, s->h_finv.data(), sizeof(float) * kNF, hipMemcpyHostToDevice, s->stream));
    return GM_OK;
}

NetArgs net_args(gm_surrogate *s, const float *feat, int64_t n) {
    NetArgs N{};
    N.feat = feat;
    N.fmean = s->fmean.p;
    N.finv = s->finv.p;
    N.wbf = s->wbf.p;
    N.param = s->param.p;
    N.n = n;
    return N;
}

MtsArgs mts_args(gm_surrogate *s, const double *x, const double *y, const double *z) {
    MtsArgs M{};
    M.n = s->n;
    M.fin = s->fin.p;
    M.fin0 = s->fin0.p;
    M.fsh0 = s->fsh0.p;
    M.x0 = s->x0.p;
    M.kmat = s->kmat.p;
    M.x = x;
    M.y = y;
    M.z = z;
    for (int a = 0; a < 3; a++) M.box[a] = s->box[a];
    M.s2 = s->s2;
    M.s3 = s->s3;
    return M;
}

gm_status features_on_device(gm_surrogate *s, const double *x, const double *y, const double *z) {
    gm_sys *e = s->ex;
    if (!e->built) return set_error(GM_ERR_NOT_BUILT, "surrogate features before the exact list is built");
    GM_HIP(s->qs.ensure((size_t)std::max<int64_t>(e->nsc * 64, 1)));
    GM_HIP(launch_feat_prep(e, x, y, z, s->cls.p, s->qs.p));
    FeatArgs A{};
    A.qs = s->qs.p;
    A.slot_atom = e->slot_atom.p;
    A.sc_col = e->sc_col.p;
    A.col_sc = e->col_sc.p;
    A.cl_ref = e->cl_ref.p;
    A.cl_half = e->cl_half.p;
    A.sc_ref = e->sc_ref.p;
    A.sc_half = e->sc_half.p;
    A.bp = e->bp;
    A.r_lo = (float)s->r_lo;
    A.r_hi = (float)s->r_hi;
    A.r_lo2 = (float)(s->r_lo * s->r_lo);
    A.r_hi2 = (float)(s->r_hi * s->r_hi);
    A.inv_delta = (float)(1.0 / s->fade);
    const double w = (s->r_hi - s->r_lo) / (kNG - 1);
    A.mu0 = (float)s->r_lo;
    A.inv_w = (float)(1.0 / w);
    A.feat = s->feat.p;
    A.sp = e->sp.p;
    A.rk_lo2 = (float)(s->in_rsw * s->in_rsw);
    A.rk_hi2 = (float)(s->ex_rc * s->ex_rc);
    A.a1 = (float)s->ex_rsw;
    A.inv_w1 = (float)(1.0 / (s->ex_rc - s->ex_rsw));
    A.a2 = (float)s->in_rsw;
    A.inv_w2 = (float)(1.0 / (s->in_rc - s->in_rsw));
    A.excl_off = e->excl_off.p;
    A.excl_nbr = e->excl_nbr.p;
    A.excl_rng = e->excl_rng.p;
    A.excl_rng_any = e->excl_rng_any.p;
    A.kmat = s->kmat.p;
    GM_HIP(launch_features(e, A));
    return GM_OK;
}

// Feature standardisation from one set of features (host): mean and
// 1/sqrt(var + 1e-4) per feature.
gm_status set_norm_from_features(gm_surrogate *s) {
    std::vector<float> f((size_t)s->n * kNF);
    GM_HIP(hipMemcpyAsync(f.data(), s->feat.p, sizeof(float) * f.size(), hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    s->h_fmean.assign(kNF, 0.0f);
    s->h_finv.assign(kNF, 1.0f);
    for (int k = 0; k < kNF; k++) {
        double m = 0.0, v = 0.0;
        for (int64_t i = 0; i < s->n; i++) m += f[(size_t)i * kNF + k];
        m /= (double)std::max<int64_t>(s->n, 1);
        for (int64_t i = 0; i < s->n; i++) {
            const double d = f[(size_t)i * kNF + k] - m;
            v += d * d;
        }
        v /= (double)std::max<int64_t>(s->n, 1);
        s->h_fmean[(size_t)k] = (float)m;
        s->h_finv[(size_t)k] = (float)(1.0 / std::sqrt(v + 1e-4));
    }
    s->norm_set = true;
    return upload_norm(s);
}

double min_img(double d, double L) { return d - L * std::round(d / L); }

// Basis scales from one inner frame (host): sigma = rms(F_shell(t0)),
// s2 = sigma / rms(K dx), s3 = sigma / rms(dF_inner).
gm_status set_scales_from_frame(gm_surrogate *s, const double *x, const double *y, const double *z) {
    const size_t n = (size_t)s->n;
    std::vector<double> fsh0(3 * n), fin(3 * n), fin0(3 * n), x0(3 * n), p(3 * n);
    std::vector<float> km(6 * n);
    GM_HIP(hipMemcpyAsync(fsh0.data(), s->fsh0.p, 24 * n, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipMemcpyAsync(fin.data(), s->fin.p, 24 * n, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipMemcpyAsync(fin0.data(), s->fin0.p, 24 * n, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipMemcpyAsync(x0.data(), s->x0.p, 24 * n, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipMemcpyAsync(km.data(), s->kmat.p, 24 * n, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipMemcpyAsync(p.data(), x, 8 * n, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipMemcpyAsync(p.data() + n, y, 8 * n, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipMemcpyAsync(p.data() + 2 * n, z, 8 * n, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    double a = 0.0, b = 0.0, c = 0.0;
    for (size_t i = 0; i < n; i++) {
        double dx[3];
        for (size_t d = 0; d < 3; d++) {
            const size_t k = d * n + i;
            a += fsh0[k] * fsh0[k];
            dx[d] = min_img(p[k] - x0[k], s->box[d]);
            const double df = fin[k] - fin0[k];
            c += df * df;
        }
        const float *K = km.data() + 6 * i;
        const double kx = K[0] * dx[0] + K[3] * dx[1] + K[4] * dx[2];
        const double ky = K[3] * dx[0] + K[1] * dx[1] + K[5] * dx[2];
        const double kz = K[4] * dx[0] + K[5] * dx[1] + K[2] * dx[2];
        b += kx * kx + ky * ky + kz * kz;
    }
    const double m = (double)std::max<size_t>(3 * n, 1);
    s->sigma = std::sqrt(a / m);
    if (!(s->sigma > 0.0)) s->sigma = 1.0;
    const double rb = std::sqrt(b / m), rc = std::sqrt(c / m);
    s->s2 = rb > 0.0 ? s->sigma / rb : 1.0;
    s->s3 = rc > 0.0 ? s->sigma / rc : 1.0;
    s->scales_set = true;
    return GM_OK;
}

gm_status ensure_train_buffers(gm_surrogate *s, int64_t nb) {
    const size_t b = (size_t)std::max<int64_t>(nb, 1);
    const int nwg = train_workgroups(nb);
    GM_HIP(s->basis.ensure(b * 9));
    GM_HIP(s->target.ensure(b * 3));
    GM_HIP(s->partial.ensure((size_t)std::max(nwg, 1) * kNParam));
    GM_HIP(s->lossp.ensure((size_t)std::max(nwg, 1) * 3));
    return GM_OK;
}

// Gradient (mode 0), Adam step (mode 1) or loss only (mode 2) over the
// staged batch; the two losses go to loss_out (device).
gm_status train_batch(gm_surrogate *s, const float *feat, int64_t n_feat, int64_t start, int64_t nb,
                      int mode, float lr, float *loss_out) {
    TrainArgs T{};
    T.net = net_args(s, feat, n_feat);
    T.start = start;
    T.nb = nb;
    T.basis = s->basis.p;
    T.target = s->target.p;
    T.taylor = (float)(1.0 / s->s2);
    T.partial = s->partial.p;
    T.loss_partial = s->lossp.p;
    GM_HIP(launch_train(T, s->stream));
    const float beta1 = 0.9f, beta2 = 0.999f;
    float c1 = 1.0f, c2 = 1.0f;
    if (mode == 1) {
        s->adam_t++;
        c1 = (float)(1.0 - std::pow((double)beta1, (double)s->adam_t));
        c2 = (float)(1.0 - std::pow((double)beta2, (double)s->adam_t));
    }
    GM_HIP(launch_reduce_adam(train_workgroups(nb), s->partial.p, s->lossp.p, 1.0 / (double)nb, loss_out,
                              mode, s->grad.p, s->param.p, s->m1.p, s->m2.p, s->wbf.p, lr, beta1, beta2,
                              c1, c2, 1e-8f, s->stream));
    return GM_OK;
}

// Exact step of the driver at positions x: inner and shell forces, their sum
// into (fx, fy, fz), and the remembered shell force, inner force and positions.
gm_status exact_step(gm_surrogate *s, const double *x, const double *y, const double *z, double *fx,
                     double *fy, double *fz) {
    const int64_t n = s->n;
    gm_status rc = compute_on_device(s->in, x, y, z, s->fin.p, s->fin.p + n, s->fin.p + 2 * n);
    if (rc != GM_OK) return rc;
    rc = compute_on_device(s->ex, x, y, z, s->fsh.p, s->fsh.p + n, s->fsh.p + 2 * n, true);
    if (rc != GM_OK) return rc;
    GM_HIP(launch_mts_t0(n, s->fin.p, s->fsh.p, s->fin0.p, s->fsh0.p, x, y, z, s->x0.p, fx, fy, fz, s->stream));
    return GM_OK;
}

gm_status maybe_rebuild(gm_surrogate *s, double *x, double *y, double *z, bool *rebuilt) {
    int32_t a = 0, b = 0;
    gm_status rc = gm_needs_rebuild_device(s->ex, x, y, z, &a);
    if (rc != GM_OK) return rc;
    if ((rc = gm_needs_rebuild_device(s->in, x, y, z, &b)) != GM_OK) return rc;
    *rebuilt = a || b;
    if (*rebuilt) return gm_surrogate_build(s, x, y, z);
    return GM_OK;
}

}  // namespace

extern "C" {

gm_status gm_surrogate_create(gm_ctx *ctx, const gm_system_desc *exact, const gm_surrogate_desc *d,
                              gm_surrogate **out) {
    GM_REQUIRE(ctx && exact && d && out, "gm_surrogate_create: NULL argument");
    *out = nullptr;
    GM_REQUIRE(exact->n_atoms == 0 || d->atom_class, "gm_surrogate_create: atom_class missing");
    for (int64_t i = 0; i < exact->n_atoms; i++)
        GM_REQUIRE(d->atom_class[i] < kNC, "gm_surrogate_create: atom class must be 0..3");
    GM_REQUIRE(d->r_lo > 0.0 && d->r_lo < d->r_hi && d->fade > 0.0 && 2.0 * d->fade <= d->r_hi - d->r_lo,
               "gm_surrogate_create: need 0 < r_lo < r_hi and 0 < 2 fade <= r_hi - r_lo");
    GM_REQUIRE(d->r_hi <= exact->r_cutoff,
               "gm_surrogate_create: the feature window must lie inside the exact model's cutoff");
    GM_REQUIRE(d->r_inner_switch < d->r_inner_cutoff && d->r_inner_cutoff <= exact->r_switch,
               "gm_surrogate_create: need r_inner_switch < r_inner_cutoff <= the exact model's r_switch");
    GM_HIP(hipSetDevice(ctx->device));
    gm_surrogate *s = new gm_surrogate();
    s->ctx = ctx;
    s->n = exact->n_atoms;
    for (int a = 0; a < 3; a++) s->box[a] = exact->box[a];
    s->r_lo = d->r_lo;
    s->r_hi = d->r_hi;
    s->fade = d->fade;
    s->ex_rsw = exact->r_switch;
    s->ex_rc = exact->r_cutoff;
    s->in_rsw = d->r_inner_switch;
    s->in_rc = d->r_inner_cutoff;
    auto fail = [&](gm_status st) {
        gm_surrogate_destroy(s);
        return st;
    };
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(from_hip(e, "hipStreamCreateWithFlags"));
    gm_status rc = system_create_impl(ctx, exact, s->stream, &s->ex);
    if (rc != GM_OK) return fail(rc);
    gm_system_desc in = *exact;
    in.r_cutoff = d->r_inner_cutoff;
    in.r_switch = d->r_inner_switch;
    in.r_skin = d->r_inner_skin;
    rc = system_create_impl(ctx, &in, s->stream, &s->in);
    if (rc != GM_OK) return fail(rc);
    // shell mode of the exact system: weight S(r) - S_in(r), rebuilt with its
    // step-unit constants whenever the system's fixed-point span changes
    s->ex->has_shell = true;
    s->ex->use_pl = false;  // (shell-mode evaluations walk the tile list)
    s->ex->shell_rsw = d->r_inner_switch;
    s->ex->shell_w = d->r_inner_cutoff - d->r_inner_switch;
    refresh_shell(s->ex);
    rc = alloc_state(s);
    if (rc != GM_OK) return fail(rc);
    if ((e = s->cls.ensure((size_t)std::max<int64_t>(s->n, 1))) != hipSuccess) return fail(from_hip(e, "hipMalloc"));
    if (s->n > 0 &&
        (e = hipMemcpy(s->cls.p, d->atom_class, (size_t)s->n, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(from_hip(e, "hipMemcpy"));
    const std::vector<float> p0 = init_params(d->seed);
    if ((rc = upload_params(s, p0.data())) != GM_OK) return fail(rc);
    s->h_fmean.assign(kNF, 0.0f);
    s->h_finv.assign(kNF, 1.0f);
    if ((rc = upload_norm(s)) != GM_OK) return fail(rc);
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return fail(from_hip(e, "hipStreamSynchronize"));
    *out = s;
    return GM_OK;
}

gm_status gm_surrogate_destroy(gm_surrogate *s) {
    if (!s) return GM_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->ex) gm_system_destroy(s->ex);
    if (s->in) gm_system_destroy(s->in);
    s->cls.release();
    s->qs.release();
    for (DevBuf<float> *b : {&s->feat, &s->kmat, &s->param, &s->m1, &s->m2, &s->grad, &s->fmean, &s->finv, &s->coef,
                             &s->basis, &s->target, &s->partial, &s->loss_hist, &s->h_in})
        b->release();
    s->wbf.release();
    for (DevBuf<double> *b : {&s->fsh, &s->fin, &s->fin0, &s->fsh0, &s->x0, &s->ftot, &s->lossp}) b->release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return GM_OK;
}

gm_status gm_surrogate_system(gm_surrogate *s, int32_t which, gm_sys **out) {
    GM_REQUIRE(s && out && (which == 0 || which == 1), "gm_surrogate_system: bad argument");
    *out = which == 0 ? s->ex : s->in;
    return GM_OK;
}

gm_status gm_surrogate_build(gm_surrogate *s, double *d_x, double *d_y, double *d_z) {
    GM_REQUIRE(s && (s->n == 0 || (d_x && d_y && d_z)), "gm_surrogate_build: NULL argument");
    gm_status rc = gm_build_neighbor_list_device(s->ex, d_x, d_y, d_z);
    if (rc != GM_OK) return rc;
    return gm_build_neighbor_list_device(s->in, d_x, d_y, d_z);
}

gm_status gm_surrogate_features(gm_surrogate *s, const double *d_x, const double *d_y, const double *d_z,
                                float *h_out) {
    GM_REQUIRE(s && (s->n == 0 || (d_x && d_y && d_z)), "gm_surrogate_features: NULL argument");
    gm_status rc = features_on_device(s, d_x, d_y, d_z);
    if (rc != GM_OK) return rc;
    if (h_out && s->n > 0)
        GM_HIP(hipMemcpyAsync(h_out, s->feat.p, sizeof(float) * (size_t)s->n * kNF, hipMemcpyDeviceToHost,
                              s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

gm_status gm_surrogate_get_hessian(gm_surrogate *s, float *h_out) {
    GM_REQUIRE(s && (s->n == 0 || h_out), "gm_surrogate_get_hessian: NULL argument");
    if (s->n > 0)
        GM_HIP(hipMemcpyAsync(h_out, s->kmat.p, sizeof(float) * 6 * (size_t)s->n, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

gm_status gm_surrogate_set_params(gm_surrogate *s, const float *params, const float *fmean, const float *finv,
                                  const double *scales) {
    GM_REQUIRE(s, "gm_surrogate_set_params: NULL surrogate");
    gm_status rc;
    if (params && (rc = upload_params(s, params)) != GM_OK) return rc;
    if (fmean) {
        s->h_fmean.assign(fmean, fmean + kNF);
        s->norm_set = true;
    }
    if (finv) {
        s->h_finv.assign(finv, finv + kNF);
        s->norm_set = true;
    }
    if ((fmean || finv) && (rc = upload_norm(s)) != GM_OK) return rc;
    if (scales) {
        GM_REQUIRE(scales[0] > 0.0 && scales[1] > 0.0 && scales[2] > 0.0, "gm_surrogate_set_params: scales must be positive");
        s->s2 = scales[0];
        s->s3 = scales[1];
        s->sigma = scales[2];
        s->scales_set = true;
    }
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

gm_status gm_surrogate_get_params(gm_surrogate *s, float *params, float *fmean, float *finv, double *scales) {
    GM_REQUIRE(s, "gm_surrogate_get_params: NULL surrogate");
    if (params)
        GM_HIP(hipMemcpyAsync(params, s->param.p, sizeof(float) * kNParam, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    if (fmean) std::memcpy(fmean, s->h_fmean.data(), sizeof(float) * kNF);
    if (finv) std::memcpy(finv, s->h_finv.data(), sizeof(float) * kNF);
    if (scales) {
        scales[0] = s->s2;
        scales[1] = s->s3;
        scales[2] = s->sigma;
    }
    return GM_OK;
}

gm_status gm_surrogate_infer(gm_surrogate *s, const float *h_feat, int64_t n, float *h_out) {
    GM_REQUIRE(s && n >= 0 && (n == 0 || (h_feat && h_out)), "gm_surrogate_infer: bad argument");
    if (n == 0) return GM_OK;
    DevBuf<float> out;
    GM_HIP(s->h_in.ensure((size_t)n * kNF));
    GM_HIP(out.ensure((size_t)n * kNO));
    GM_HIP(hipMemcpyAsync(s->h_in.p, h_feat, sizeof(float) * (size_t)n * kNF, hipMemcpyHostToDevice, s->stream));
    hipError_t e = launch_infer(net_args(s, s->h_in.p, n), out.p, s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_out, out.p, sizeof(float) * (size_t)n * kNO, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    out.release();
    GM_HIP(e);
    return GM_OK;
}

gm_status gm_surrogate_loss_grad(gm_surrogate *s, const float *h_feat, const float *h_basis, const float *h_target,
                                 int64_t n, double *loss, float *h_grad) {
    GM_REQUIRE(s && n > 0 && h_feat && h_basis && h_target && loss && h_grad, "gm_surrogate_loss_grad: bad argument");
    gm_status rc = ensure_train_buffers(s, n);
    if (rc != GM_OK) return rc;
    GM_HIP(s->h_in.ensure((size_t)n * kNF));
    GM_HIP(hipMemcpyAsync(s->h_in.p, h_feat, sizeof(float) * (size_t)n * kNF, hipMemcpyHostToDevice, s->stream));
    GM_HIP(hipMemcpyAsync(s->basis.p, h_basis, sizeof(float) * (size_t)n * 9, hipMemcpyHostToDevice, s->stream));
    GM_HIP(hipMemcpyAsync(s->target.p, h_target, sizeof(float) * (size_t)n * 3, hipMemcpyHostToDevice, s->stream));
    if ((rc = train_batch(s, s->h_in.p, n, 0, n, 0, 0.0f, s->loss_hist.p)) != GM_OK) return rc;
    float l[3] = {0.0f, 0.0f, 0.0f};
    GM_HIP(hipMemcpyAsync(h_grad, s->grad.p, sizeof(float) * kNParam, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipMemcpyAsync(l, s->loss_hist.p, sizeof(l), hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    *loss = l[0];
    return GM_OK;
}

gm_status gm_surrogate_random_walk(gm_surrogate *s, double *d_x, double *d_y, double *d_z, double sigma,
                                   uint64_t seed, uint64_t step) {
    GM_REQUIRE(s && (s->n == 0 || (d_x && d_y && d_z)), "gm_surrogate_random_walk: NULL argument");
    GM_HIP(launch_random_walk(s->n, d_x, d_y, d_z, sigma, seed, step, s->stream));
    return GM_OK;
}

gm_status gm_surrogate_train(gm_surrogate *s, double *d_x, double *d_y, double *d_z, const gm_train_desc *td,
                             float *h_loss, gm_train_report *rep) {
    GM_REQUIRE(s && td && (s->n == 0 || (d_x && d_y && d_z)), "gm_surrogate_train: NULL argument");
    GM_REQUIRE(td->n_frames >= 0 && td->batch > 0 && td->mu >= 0.0 && td->sigma >= 0.0 && td->cap > 0.0 &&
                   td->lr >= 0.0f,
               "gm_surrogate_train: need frames >= 0, batch > 0, mu, sigma >= 0, cap > 0");
    GM_REQUIRE(s->n > 0, "gm_surrogate_train: empty system");
    const int64_t n = s->n, nb = std::min<int64_t>(td->batch, n);
    gm_status rc = ensure_train_buffers(s, nb);
    if (rc != GM_OK) return rc;
    GM_HIP(s->loss_hist.ensure((size_t)std::max<int64_t>(td->n_frames, 1) * 3));
    const auto t0 = std::chrono::steady_clock::now();
    if ((rc = gm_surrogate_build(s, d_x, d_y, d_z)) != GM_OK) return rc;
    int64_t frame = 0, k = 0;
    const double *tx = s->ftot.p, *ty = s->ftot.p + n, *tz = s->ftot.p + 2 * n;
    while (frame < td->n_frames) {
        if (k % 4 == 0) {
            bool rebuilt = false;
            if (k > 0 && (rc = maybe_rebuild(s, d_x, d_y, d_z, &rebuilt)) != GM_OK) return rc;
            if ((rc = exact_step(s, d_x, d_y, d_z, s->ftot.p, s->ftot.p + n, s->ftot.p + 2 * n)) != GM_OK) return rc;
            if ((rc = features_on_device(s, d_x, d_y, d_z)) != GM_OK) return rc;
            if (!s->norm_set && (rc = set_norm_from_features(s)) != GM_OK) return rc;
            GM_HIP(launch_bd_step(n, d_x, d_y, d_z, tx, ty, tz, nullptr, td->mu, td->sigma, td->cap, td->seed,
                                  (uint64_t)k, s->stream));
        } else {
            if ((rc = compute_on_device(s->in, d_x, d_y, d_z, s->fin.p, s->fin.p + n, s->fin.p + 2 * n)) != GM_OK)
                return rc;
            if ((rc = compute_on_device(s->ex, d_x, d_y, d_z, s->fsh.p, s->fsh.p + n, s->fsh.p + 2 * n, true)) !=
                GM_OK)
                return rc;
            if (!s->scales_set && (rc = set_scales_from_frame(s, d_x, d_y, d_z)) != GM_OK) return rc;
            const int64_t start = (frame * nb) % n;
            GM_HIP(launch_train_prep(mts_args(s, d_x, d_y, d_z), s->fsh.p, start, nb, 1.0 / s->sigma, s->basis.p,
                                     s->target.p, s->stream));
            if ((rc = train_batch(s, s->feat.p, n, start, nb, td->update ? 1 : 2, td->lr,
                                  s->loss_hist.p + 3 * frame)) != GM_OK)
                return rc;
            frame++;
            // trajectory driven by the exact force: inner + shell
            GM_HIP(launch_bd_step(n, d_x, d_y, d_z, s->fin.p, s->fin.p + n, s->fin.p + 2 * n, s->fsh.p, td->mu,
                                  td->sigma, td->cap, td->seed, (uint64_t)k, s->stream));
        }
        k++;
    }
    GM_HIP(hipStreamSynchronize(s->stream));
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::vector<float> hist((size_t)td->n_frames * 3);
    if (td->n_frames > 0)
        GM_HIP(hipMemcpy(hist.data(), s->loss_hist.p, sizeof(float) * hist.size(), hipMemcpyDeviceToHost));
    if (h_loss && td->n_frames > 0) std::memcpy(h_loss, hist.data(), sizeof(float) * hist.size());
    if (rep) {
        gm_train_report r{};
        r.frames = td->n_frames;
        r.steps = k;
        r.ms_total = ms;
        if (td->n_frames > 0) {
            r.loss_first = hist[0];
            r.loss_last = hist[hist.size() - 3];
            double a = 0.0, b = 0.0, c = 0.0;
            for (int64_t f = 0; f < td->n_frames; f++) {
                a += hist[(size_t)f * 3];
                b += hist[(size_t)f * 3 + 1];
                c += hist[(size_t)f * 3 + 2];
            }
            r.loss_net = a / (double)td->n_frames;
            r.loss_stale = b / (double)td->n_frames;
            r.loss_taylor = c / (double)td->n_frames;
        }
        *rep = r;
    }
    return GM_OK;
}

gm_status gm_surrogate_bd_step(gm_surrogate *s, double *d_x, double *d_y, double *d_z, const double *d_fx,
                               const double *d_fy, const double *d_fz, double mu, double sigma, double cap,
                               uint64_t seed, uint64_t step) {
    GM_REQUIRE(s && (s->n == 0 || (d_x && d_y && d_z && d_fx && d_fy && d_fz)), "gm_surrogate_bd_step: NULL argument");
    GM_REQUIRE(cap > 0.0 && mu >= 0.0 && sigma >= 0.0, "gm_surrogate_bd_step: need mu, sigma >= 0, cap > 0");
    GM_HIP(launch_bd_step(s->n, d_x, d_y, d_z, d_fx, d_fy, d_fz, nullptr, mu, sigma, cap, seed, step, s->stream));
    return GM_OK;
}

gm_status gm_surrogate_shell_forces(gm_surrogate *s, const double *d_x, const double *d_y, const double *d_z,
                                    double *d_fx, double *d_fy, double *d_fz) {
    GM_REQUIRE(s && (s->n == 0 || (d_x && d_y && d_z && d_fx && d_fy && d_fz)),
               "gm_surrogate_shell_forces: NULL argument");
    return compute_on_device(s->ex, d_x, d_y, d_z, d_fx, d_fy, d_fz, true);
}

gm_status gm_surrogate_maybe_rebuild(gm_surrogate *s, double *d_x, double *d_y, double *d_z, int32_t *rebuilt) {
    GM_REQUIRE(s && (s->n == 0 || (d_x && d_y && d_z)), "gm_surrogate_maybe_rebuild: NULL argument");
    bool r = false;
    gm_status rc = maybe_rebuild(s, d_x, d_y, d_z, &r);
    if (rc != GM_OK) return rc;
    if (rebuilt) *rebuilt = r ? 1 : 0;
    return GM_OK;
}

gm_status gm_surrogate_minimize(gm_surrogate *s, double *d_x, double *d_y, double *d_z, int32_t steps, double mu,
                                double cap, double *max_force) {
    GM_REQUIRE(s && (s->n == 0 || (d_x && d_y && d_z)) && steps >= 0 && mu >= 0.0 && cap > 0.0,
               "gm_surrogate_minimize: bad argument");
    const int64_t n = s->n;
    gm_status rc = gm_surrogate_build(s, d_x, d_y, d_z);
    if (rc != GM_OK) return rc;
    double *fx = s->ftot.p, *fy = s->ftot.p + n, *fz = s->ftot.p + 2 * n;
    for (int32_t it = 0; it < steps; it++) {
        bool rebuilt = false;
        if (it > 0 && (rc = maybe_rebuild(s, d_x, d_y, d_z, &rebuilt)) != GM_OK) return rc;
        if ((rc = compute_on_device(s->ex, d_x, d_y, d_z, fx, fy, fz)) != GM_OK) return rc;
        GM_HIP(launch_bd_step(n, d_x, d_y, d_z, fx, fy, fz, nullptr, mu, 0.0, cap, 0, 0, s->stream));
    }
    if ((rc = gm_surrogate_build(s, d_x, d_y, d_z)) != GM_OK) return rc;
    if (max_force) {
        if ((rc = compute_on_device(s->ex, d_x, d_y, d_z, fx, fy, fz)) != GM_OK) return rc;
        std::vector<double> f((size_t)(3 * n));
        GM_HIP(hipMemcpyAsync(f.data(), s->ftot.p, sizeof(double) * f.size(), hipMemcpyDeviceToHost, s->stream));
        GM_HIP(hipStreamSynchronize(s->stream));
        double m = 0.0;
        for (double v : f) m = std::max(m, std::fabs(v));
        *max_force = m;
    }
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

gm_status gm_mts_reset(gm_surrogate *s) {
    GM_REQUIRE(s, "gm_mts_reset: NULL surrogate");
    s->mts_k = 0;
    return GM_OK;
}

gm_status gm_mts_step(gm_surrogate *s, const double *d_x, const double *d_y, const double *d_z, double *d_fx,
                      double *d_fy, double *d_fz, int32_t mode, int32_t *was_exact) {
    GM_REQUIRE(s && (s->n == 0 || (d_x && d_y && d_z && d_fx && d_fy && d_fz)), "gm_mts_step: NULL argument");
    GM_REQUIRE(mode >= GM_MTS_HELD && mode <= GM_MTS_TAYLOR, "gm_mts_step: unknown mode");
    GM_REQUIRE(mode == GM_MTS_HELD || s->scales_set,
               "gm_mts_step: no basis scales yet (train or set them first)");
    const bool exact = (s->mts_k % 4) == 0;
    gm_status rc;
    if (exact) {
        if ((rc = exact_step(s, d_x, d_y, d_z, d_fx, d_fy, d_fz)) != GM_OK) return rc;
        if (mode != GM_MTS_HELD) {
            if ((rc = features_on_device(s, d_x, d_y, d_z)) != GM_OK) return rc;
            if (mode == GM_MTS_NET) GM_HIP(launch_infer(net_args(s, s->feat.p, s->n), s->coef.p, s->stream));
        }
    } else {
        if ((rc = compute_on_device(s->in, d_x, d_y, d_z, s->fin.p, s->fin.p + s->n, s->fin.p + 2 * s->n)) != GM_OK)
            return rc;
        GM_HIP(launch_mts_combine(mts_args(s, d_x, d_y, d_z), mode, s->coef.p, d_fx, d_fy, d_fz, s->stream));
    }
    if (was_exact) *was_exact = exact ? 1 : 0;
    s->mts_k++;
    return GM_OK;
}

}  // extern "C"
