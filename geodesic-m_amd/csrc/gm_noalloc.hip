// gm_compute_forces_device_noalloc: a force evaluation that allocates nothing.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid it).  It is gm_compute_forces_device
// with the evaluation's scratch (fixed-point accumulator, non-finite flags,
// displacement words, bonded and per-block energies, i-force records, bonded
// force records) taken
// from `scratch`, a static device array of kScratchBytes the library holds on
// every GPU it runs on until the process ends, instead of the system's own
// buffers.  The system (list, slot positions, energies, outputs) is the
// caller's, built beforehand.
//
// In host code `scratch` names a host-side stand-in, not the device array:
// gm_init (scratch_init, with the device current) asks HIP for this device's
// address of it once, and every use below goes through that address.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two host threads) race.  Each call clears what must start at
// zero with a fill kernel (no memset node), so it may be captured in a graph.
#include "gm_internal.hpp"

constexpr size_t kScratchBytes = (size_t)32 << 20;  // 32 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

namespace gm {

gm_status scratch_init(gm_ctx *c) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return from_hip(e, "hipGetSymbolAddress(scratch)");
    c->scratch = p;
    c->scratch_bytes = kScratchBytes;
    return GM_OK;
}

}  // namespace gm

using namespace gm;

extern "C" gm_status gm_compute_forces_device_noalloc(gm_sys *s, const double *dx, const double *dy,
                                                      const double *dz, double *dfx, double *dfy, double *dfz) {
    if (!s) return set_error(GM_ERR_INVALID_ARG, "NULL system");
    if (!(s->n == 0 || (dx && dy && dz && dfx && dfy && dfz))) return set_error(GM_ERR_INVALID_ARG, "NULL array");
    if (!s->built) return set_error(GM_ERR_NOT_BUILT, "compute_forces before build_neighbor_list");
    if (!s->ctx->scratch) return set_error(GM_ERR_KERNEL, "static scratch array has no device address");
    // layout: [F 3 x slots][bad slots][disp]  (zeroed by every call)  [e_items][e_blk][fi_blk][brec]
    const size_t n_pad = (size_t)s->nsc * 64;
    const size_t b_F = align256(sizeof(unsigned long long) * 3 * n_pad);
    const size_t b_bad = align256(sizeof(unsigned) * n_pad);
    const size_t b_disp = align256(sizeof(unsigned) * kDispSlots * kDispStride);
    const size_t b_items = align256(sizeof(double) * (size_t)(s->nb + s->na + s->nd + s->n14));
    const bool recs = s->fi_records && !pl_active(s);  // i-force records (tile path, small work items)
    const size_t b_blk = align256(sizeof(double) * (size_t)lj_records(s));
    const size_t b_rec = recs ? align256(sizeof(float) * 192 * (size_t)s->n_blk) : 0;
    const bool brecs = pl_active(s) && s->n_contrib > 0;  // bonded force records (partner lists)
    const size_t b_brec = brecs ? align256(sizeof(long long) * 3 * (size_t)s->n_contrib) : 0;
    const size_t need = b_F + b_bad + b_disp + b_items + b_blk + b_rec + b_brec;
    if (need > s->ctx->scratch_bytes)
        return set_error(GM_ERR_CAPACITY, "gm_compute_forces_device_noalloc: this list needs " +
                                              std::to_string(need) + " bytes of scratch, more than the " +
                                              std::to_string(s->ctx->scratch_bytes) +
                                              " of the static array (larger work items need no i-force records: "
                                              "gm_set_work_chunk 16 or more)");
    unsigned char *base = static_cast<unsigned char *>(s->ctx->scratch);
    EvalBufs b{};
    size_t o = 0;
    b.F = reinterpret_cast<unsigned long long *>(base + o);
    o += b_F;
    b.bad = reinterpret_cast<unsigned *>(base + o);
    o += b_bad;
    b.disp = reinterpret_cast<unsigned *>(base + o);
    o += b_disp;
    const size_t zeroed = o;  // the accumulator, flags and displacement words start at zero
    b.e_items = reinterpret_cast<double *>(base + o);
    o += b_items;
    b.e_blk = reinterpret_cast<double *>(base + o);
    o += b_blk;
    b.fi_blk = recs ? reinterpret_cast<float *>(base + o) : nullptr;
    o += b_rec;
    b.brec = brecs ? reinterpret_cast<long long *>(base + o) : nullptr;  // (every record written each call)
    GM_HIP(launch_fill_u64(b.F, zeroed / sizeof(unsigned long long), s->stream));
    // the system's own accumulator is untouched (still zero for its next evaluation)
    return compute_on_device(s, dx, dy, dz, dfx, dfy, dfz, false, &b);
}
