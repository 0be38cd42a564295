// C-ABI of the gfx950 force engine (include/geodesic_hip.h).
//
// Host orchestration of the device path that replaces the reference's
// GpuBackend (geodesic-gpu/src/gpu_backend.rs:11-100; paths relative to the
// reference repository root).  Everything numeric runs on the GPU; the host
// only validates descriptors, sizes buffers, scans small per-supercluster
// counts between the list-build kernels and decodes the list for export.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "gm_internal.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace gm {

gm_status nonfinite_status(const gm_sys *s) {
    if (s->last_bad == 0x7fffffff) return GM_OK;
    return set_error(GM_ERR_NUMERICAL, "force not finite at atom " + std::to_string(s->last_bad) +
                                           " (forces written; non-finite atoms hold NaN)");
}

gm_status set_error(gm_status s, const std::string &msg) {
    g_last_error = msg;
    return s;
}

gm_status from_hip(hipError_t e, const char *what) {
    gm_status s = GM_ERR_KERNEL;
    switch (e) {
        case hipErrorOutOfMemory: s = GM_ERR_OOM; break;
        case hipErrorNoDevice:
        case hipErrorInvalidDevice: s = GM_ERR_NO_DEVICE; break;
        case hipErrorNoBinaryForGpu:
        case hipErrorInvalidImage:
        case hipErrorInvalidKernelFile:
        case hipErrorSharedObjectInitFailed: s = GM_ERR_KERNEL_LOAD; break;
        case hipErrorLaunchFailure:
        case hipErrorIllegalAddress:
        case hipErrorECCNotCorrectable:
        case hipErrorAssert: s = GM_ERR_DEVICE_LOST; break;
        default: break;
    }
    return set_error(s, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace gm

using namespace gm;

#define GM_REQUIRE(cond, msg)                                   \
    do {                                                        \
        if (!(cond)) return set_error(GM_ERR_INVALID_ARG, msg); \
    } while (0)

namespace gm {

template <typename T>
hipError_t upload(DevBuf<T> &b, const T *src, size_t n, hipStream_t st) {
    hipError_t e = b.ensure(n);
    if (e != hipSuccess || n == 0) return e;
    return hipMemcpyAsync(b.p, src, sizeof(T) * n, hipMemcpyHostToDevice, st);
}

bool indices_ok(const uint32_t *a, int64_t cnt, int64_t n) {
    for (int64_t k = 0; k < cnt; k++)
        if ((int64_t)a[k] >= n) return false;
    return true;
}

// Work items: runs of at most `chunk` consecutive list entries of one
// supercluster (whole fixed-point blocks of kBlock entries, counted from the
// supercluster's first entry), one wavefront each, built on the device
// (k_sc_scan, k_wi_make, k_cost_scan, k_wi_scatter).  Automatic chunk: about
// 1.5 work items per resident wavefront slot (5 per SIMD), rounded up to
// whole blocks, 8 to 32 entries.  The chunk only groups blocks into items:
// every block's i forces and energy are rounded on their own, so forces and
// energies do not depend on it.  Items are issued heaviest first (estimated
// cost: tiles plus per-entry and per-item overhead): workgroups are
// dispatched in index order as slots free up, so this is longest-processing-
// time-first list scheduling and the light items fill the tail.
//
// The host learns the counts from the build words the device wrote
// (meta): work items, blocks, entries per item.
gm_status finish_work_items(gm_sys *s, const int *m) {
    s->n_wi = m[kMetaNwi];
    s->n_blk = m[kMetaNblk];
    s->fi_records = m[kMetaChunk] < 16;
    const void *old_fi = s->fi_blk.p, *old_e = s->e_blk.p;
    if (s->fi_records) GM_HIP(s->fi_blk.ensure((size_t)std::max<int64_t>(s->n_blk, 1) * 192));
    GM_HIP(s->e_blk.ensure((size_t)std::max<int64_t>(s->n_blk, 1)));  // LJ energy records, per block
    if (s->fi_blk.p != old_fi || s->e_blk.p != old_e) s->buf_gen++;
    if (!s->use_pl) s->stats.n_work_items = s->n_wi;
    return GM_OK;
}

// New work items for a built tile list (gm_set_work_chunk, a batch's shape).
gm_status make_work_items(gm_sys *s) {
    GM_HIP(launch_sc_scan(s, s->ent_cap, s->chunk));
    GM_HIP(launch_work_items(s));
    int m[kMetaWords];
    GM_HIP(hipMemcpyAsync(m, s->meta.p, sizeof(m), hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    return finish_work_items(s, m);
}

// Entry and work-item buffers for about nent entries (a quarter more, so
// that later builds of a moving system fit without resizing).
gm_status size_list_buffers(gm_sys *s, int64_t nent) {
    const int64_t cap = nent + nent / 4 + 1024;
    if (cap > (int64_t)0x7fffffff) return set_error(GM_ERR_CAPACITY, "list build: too many entries");
    GM_HIP(s->ent_sc.ensure((size_t)cap));
    GM_HIP(s->ent_cj.ensure((size_t)cap));
    GM_HIP(s->ent_imask.ensure((size_t)cap));
    GM_HIP(s->ent_meta.ensure((size_t)cap));
    GM_HIP(s->ent_dmin.ensure((size_t)cap));
    GM_HIP(s->ent_lb.ensure((size_t)cap * 64));
    const int64_t icap = cap / kBlock + s->max_sc + 1;  // items of at least kBlock entries
    GM_HIP(s->wi.ensure((size_t)icap));
    GM_HIP(s->wi_tmp.ensure((size_t)icap));
    GM_HIP(s->wi_cost.ensure((size_t)icap));
    s->ent_cap = cap;
    s->item_cap = icap;
    return GM_OK;
}

// Partner-list buffers: full entries for about nfent (a quarter more), list
// entries for about npl (a quarter more, plus one group's worth).
gm_status size_fent_buffers(gm_sys *s, int64_t nfent) {
    const int64_t cap = nfent + nfent / 4 + 1024;
    if (cap > (int64_t)0x7fffffff / 64) return set_error(GM_ERR_CAPACITY, "list build: too many full entries");
    GM_HIP(s->fent_sc.ensure((size_t)cap));
    GM_HIP(s->fent_cj.ensure((size_t)cap));
    GM_HIP(s->fent_w.ensure((size_t)cap * 64));
    s->fent_cap = cap;
    return GM_OK;
}

gm_status size_pl_buffer(gm_sys *s, int64_t npl) {
    const int64_t cap = npl + npl / 4 + 1024;
    if (cap + kPlSlack > (int64_t)0x7fffffff) return set_error(GM_ERR_CAPACITY, "list build: partner lists too long");
    GM_HIP(s->pl_nbr.ensure((size_t)(cap + kPlSlack)));
    s->pl_cap = cap;
    return GM_OK;
}

constexpr int64_t kProfMax = 8192;
constexpr int kProfStride = 4;  // events per profiled evaluation

void refresh_shell(gm_sys *s) {
    NbParams q = s->nbp;
    const double w_in = s->shell_w;
    const double step = s->qbox[0] / 4294967296.0;
    q.rin_sw = (float)s->shell_rsw;
    q.inv_w_in = (float)(1.0 / w_in);
    q.ds_coef_in = (float)(-30.0 / w_in);
    q.rin_sw_u = (float)(s->shell_rsw / step);
    q.inv_w_in_u = (float)(step / w_in);
    q.ds_coef_in_u = (float)(-30.0 * step / w_in);
    q.nrw_in = (float)(-s->shell_rsw / w_in);
    s->nbp_shell = q;
}

void set_quant_span(gm_sys *s, const double L[3], int ncx, int ncy) {
    const double two32 = 4294967296.0;
    NbParams &p = s->nbp;
    BuildParams &bp = s->bp;
    for (int a = 0; a < 3; a++) {
        s->qbox[a] = L[a];
        bp.inv_q[a] = two32 / L[a];
        bp.sq[a] = (float)(L[a] / two32);
    }
    p.sx = (float)(L[0] / two32);
    p.sy = (float)(L[1] / two32);
    p.sz = (float)(L[2] / two32);
    {
        const double step = L[0] / two32;  // fixed-point step, A (cubic spans)
        p.inv_s = (float)(1.0 / step);
        p.rc2_u = (float)((s->rc / step) * (s->rc / step));
        p.rsw2_u = (float)((s->rsw / step) * (s->rsw / step));
        p.rsw_u = (float)(s->rsw / step);
        p.inv_w_u = (float)(step / (s->rc - s->rsw));
        p.ds_coef_u = (float)(-30.0 * step / (s->rc - s->rsw));
        p.fscale_u = (float)(two32 / step);
    }
    bp.ncx = ncx;
    bp.ncy = ncy;
    bp.ncol = ncx * ncy;
    {
        // The f32 estimate of r2 from fixed-point displacements errs by a few
        // ulp plus 2 r (quantisation step) sqrt(3); the band is far wider.
        const double qmax = std::max(L[0], std::max(L[1], L[2])) / two32;
        bp.band_skin = (float)(1e-4 * bp.r_skin2 + 16.0 * s->rskin * qmax);
        bp.band_cut = (float)(1e-4 * bp.r_cut2 + 16.0 * s->rc * qmax);
        const double r_in = s->rc + kPlInner;
        bp.band_in = (float)(1e-4 * r_in * r_in + 16.0 * r_in * qmax);
    }
    bp.wcol[0] = (float)(L[0] / bp.ncx);
    bp.wcol[1] = (float)(L[1] / bp.ncy);
    bp.kx = (int)std::min<double>(bp.ncx, std::floor(bp.r_search / bp.wcol[0]) + 1.0);
    bp.ky = (int)std::min<double>(bp.ncy, std::floor(bp.r_search / bp.wcol[1]) + 1.0);
    // the pruning bound needs every pair's nearest image in the span to stay
    // the nearest one under displacements below (r_skin - r_cutoff) / 2: edges >= 2 r_skin
    s->prune_geom = std::min(L[0], std::min(L[1], L[2])) >= 2.0 * s->rskin * (1.0 + 1e-6);
    if (s->has_shell) refresh_shell(s);
}

// Non-periodic systems: the fixed-point span is a cube of edge the power of
// two at or above 2 (extent + 2 r_skin), so that no two atoms of the build
// nor any listed pair until the next rebuild (each atom moves less than
// (r_skin - r_cutoff) / 2) are half a span apart: the wrapped 32-bit
// difference is the plain difference.  When that cube does not fit in the
// smallest box edge, or a position is not finite, the span is the box.
gm_status fit_span_to_atoms(gm_sys *s, const double *px, const double *py, const double *pz) {
    double ext[8];
    GM_HIP(launch_extent(s, px, py, pz, s->ext.p));
    GM_HIP(hipMemcpyAsync(ext, s->ext.p, sizeof(ext), hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    const double e[3] = {ext[3] - ext[0], ext[4] - ext[1], ext[5] - ext[2]};
    const double emax = std::max(e[0], std::max(e[1], e[2]));
    const double need = 2.0 * (emax + 2.0 * s->rskin);
    const double bmin = std::min(s->box[0], std::min(s->box[1], s->box[2]));
    const bool finite = ext[6] == 0.0 && std::isfinite(emax);
    if (!finite || !(need <= bmin)) {
        set_quant_span(s, s->box, s->box_ncx, s->box_ncy);
        return GM_OK;
    }
    const double L = std::min(std::exp2(std::ceil(std::log2(need))), bmin);
    // columns about the width of a cube holding 64 atoms at the density of the
    // atoms' bounding box (each edge at least r_skin)
    const double vol = std::max(e[0], s->rskin) * std::max(e[1], s->rskin) * std::max(e[2], s->rskin);
    const double side = std::cbrt(64.0 * vol / (double)s->n);
    const int nc = (int)std::max(1.0, std::min((double)s->ncx_cap, std::floor(L / side + 0.5)));
    const double Lc[3] = {L, L, L};
    set_quant_span(s, Lc, nc, nc);
    return GM_OK;
}

// Host arrays to and from device arrays through the system's pinned stage
// (3 n doubles each way, allocated at the first use).  Pageable transfers ran
// at about 2.4 GB/s here; a copy into pinned memory and one DMA of all three
// arrays (the system's position and force triples are one allocation each)
// take a fraction of that.  Three separate device arrays go one by one, the
// copy of array k + 1 overlapping the transfer of array k.  The stage is
// reused only after the previous call's transfers out of it ended.
gm_status stage_ensure(gm_sys *s) {
    const size_t n = (size_t)s->n;
    if (s->stage && s->stage_n >= n) return GM_OK;
    if (s->stage) {
        if (s->stage_in_pending) GM_HIP(hipEventSynchronize(s->stage_ev[3]));
        GM_HIP(hipHostFree(s->stage));
    }
    s->stage = nullptr;
    s->stage_n = 0;
    s->stage_in_pending = false;
    GM_HIP(hipHostMalloc((void **)&s->stage, sizeof(double) * 6 * n, hipHostMallocDefault));
    s->stage_n = n;
    for (hipEvent_t &ev : s->stage_ev)
        if (!ev) GM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    return GM_OK;
}

gm_status host_to_device(gm_sys *s, const double *const src[3], double *const dst[3]) {
    const size_t n = (size_t)s->n;
    if (n == 0) return GM_OK;
    gm_status rc = stage_ensure(s);
    if (rc != GM_OK) return rc;
    if (s->stage_in_pending) GM_HIP(hipEventSynchronize(s->stage_ev[3]));  // the previous call's transfers
    if (dst[1] == dst[0] + n && dst[2] == dst[0] + 2 * n) {  // one allocation: one transfer
        for (int k = 0; k < 3; k++) std::memcpy(s->stage + k * n, src[k], sizeof(double) * n);
        GM_HIP(hipMemcpyAsync(dst[0], s->stage, sizeof(double) * 3 * n, hipMemcpyHostToDevice, s->stream));
    } else {
        for (int k = 0; k < 3; k++) {
            std::memcpy(s->stage + k * n, src[k], sizeof(double) * n);
            GM_HIP(hipMemcpyAsync(dst[k], s->stage + k * n, sizeof(double) * n, hipMemcpyHostToDevice,
                                  s->stream));
        }
    }
    GM_HIP(hipEventRecord(s->stage_ev[3], s->stream));
    s->stage_in_pending = true;
    return GM_OK;
}

// Device arrays to host arrays through the stage's second half (one transfer
// for a contiguous triple; otherwise array k is copied out as soon as its own
// transfer ended).  Returns after the stream reached the last transfer.
gm_status device_to_host(gm_sys *s, const double *const src[3], double *const dst[3]) {
    const size_t n = (size_t)s->n;
    if (n == 0) return GM_OK;
    gm_status rc = stage_ensure(s);
    if (rc != GM_OK) return rc;
    double *out = s->stage + 3 * n;
    if (src[1] == src[0] + n && src[2] == src[0] + 2 * n) {  // one allocation: one transfer
        GM_HIP(hipMemcpyAsync(out, src[0], sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s->stream));
        GM_HIP(hipStreamSynchronize(s->stream));
        for (int k = 0; k < 3; k++) std::memcpy(dst[k], out + k * n, sizeof(double) * n);
        return GM_OK;
    }
    for (int k = 0; k < 3; k++) {
        GM_HIP(hipMemcpyAsync(out + k * n, src[k], sizeof(double) * n, hipMemcpyDeviceToHost, s->stream));
        GM_HIP(hipEventRecord(s->stage_ev[k], s->stream));
    }
    for (int k = 0; k < 3; k++) {
        GM_HIP(hipEventSynchronize(s->stage_ev[k]));
        std::memcpy(dst[k], out + k * n, sizeof(double) * n);
    }
    return GM_OK;
}

// The kernels of a build whose buffers are sized (all but the first build of
// a system), in stream order.
hipError_t launch_sized_build(gm_sys *s, bool wrap, double *px, double *py, double *pz, bool clear_cols) {
    hipError_t e;
    if ((e = launch_build(s, wrap, px, py, pz, clear_cols)) != hipSuccess) return e;
    if (s->use_pl) {
        if ((e = launch_pl_count(s)) != hipSuccess) return e;
        return launch_pl_fill(s);
    }
    if ((e = launch_pair_count(s, s->ent_cap, s->chunk)) != hipSuccess) return e;
    if ((e = launch_pair_fill_and_masks(s)) != hipSuccess) return e;
    return launch_work_items(s);
}

// A build's list buffers are sized (the entry buffers of the layout in use).
bool build_sized(const gm_sys *s) { return s->use_pl ? s->fent_cap > 0 : s->ent_cap > 0; }

// What the captured build depends on: its launch arguments (positions,
// buffers that are resized, the work chunk, the span parameters).
std::vector<unsigned char> build_graph_key(gm_sys *s, bool wrap, double *px, double *py, double *pz) {
    const bool pl = s->use_pl;  // (the other layout's buffers are not launch arguments)
    const void *ptrs[] = {px, py, pz, pl ? nullptr : s->ent_sc.p, pl ? nullptr : s->ent_cj.p,
                          pl ? nullptr : s->ent_imask.p, pl ? nullptr : s->ent_meta.p, pl ? nullptr : s->ent_dmin.p,
                          pl ? nullptr : s->ent_lb.p, pl ? nullptr : s->wi.p, pl ? nullptr : s->wi_tmp.p,
                          pl ? nullptr : s->wi_cost.p, pl ? s->fent_sc.p : nullptr, pl ? s->fent_cj.p : nullptr,
                          pl ? s->fent_w.p : nullptr, pl ? s->pl_nbr.p : nullptr};
    const int64_t ints[] = {wrap ? 1 : 0, pl ? 0 : s->chunk, pl ? 0 : s->ent_cap, pl ? 0 : s->item_cap, s->n,
                            s->max_sc, pl ? 1 : 0, pl ? s->fent_cap : 0, pl ? s->pl_cap : 0};
    std::vector<unsigned char> k(sizeof(ptrs) + sizeof(ints) + sizeof(BuildParams));
    std::memcpy(k.data(), ptrs, sizeof(ptrs));
    std::memcpy(k.data() + sizeof(ptrs), ints, sizeof(ints));
    std::memcpy(k.data() + sizeof(ptrs) + sizeof(ints), &s->bp, sizeof(BuildParams));
    return k;
}

// A sized build of a system with a fixed span (periodic) runs as one captured
// graph (the same kernels with the same arguments as the stream launches),
// captured again when anything it depends on changes: the kernels go to the
// device at once instead of one host launch each.  A system whose arguments
// change on two builds in a row (positions from alternating arrays, say), or
// whose capture fails, goes back to stream launches for good.
void drop_build_graph(gm_sys *s) {
    if (s->bexec) (void)hipGraphExecDestroy(s->bexec);
    if (s->bgraph) (void)hipGraphDestroy(s->bgraph);
    s->bexec = nullptr;
    s->bgraph = nullptr;
    s->bkey.clear();
}

gm_status launch_build_graph(gm_sys *s, bool wrap, double *px, double *py, double *pz) {
    std::vector<unsigned char> key = build_graph_key(s, wrap, px, py, pz);
    if (s->bexec && key == s->bkey) {
        s->bgraph_misses = 0;
    } else {
        if (s->bexec && ++s->bgraph_misses >= 2) s->bgraph_off = true;
        drop_build_graph(s);
        if (!s->bgraph_off) {
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                const hipError_t e1 = launch_sized_build(s, wrap, px, py, pz, false);
                const hipError_t e2 = hipStreamEndCapture(s->stream, &g);
                e = e1 != hipSuccess ? e1 : e2;
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&s->bexec, g, nullptr, nullptr, 0);
            if (e == hipSuccess) {
                s->bgraph = g;
                s->bkey = std::move(key);
            } else {
                (void)hipGetLastError();  // clear a capture error
                if (g) (void)hipGraphDestroy(g);
                s->bexec = nullptr;
                s->bgraph_off = true;
            }
        }
        if (s->bgraph_off) {
            GM_HIP(launch_sized_build(s, wrap, px, py, pz, false));
            return GM_OK;
        }
    }
    GM_HIP(hipGraphLaunch(s->bexec, s->stream));
    return GM_OK;
}

gm_status read_meta(gm_sys *s, int *m) {
    GM_HIP(hipMemcpyAsync(m, s->meta.p, sizeof(int) * kMetaWords, hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

// Partner lists longer than their buffers (the kernels then skip the rest):
// size the buffers from the build words of the last pass (m) and make the
// lists again.  After a full-entry overflow the list length is not known
// yet, so a second pass may size the list buffer.
gm_status fit_partner_lists(gm_sys *s, int *m) {
    for (int pass = 0; pass < 3 && (m[kMetaFentOver] || m[kMetaPlOver]); pass++) {
        gm_status rc = GM_OK;
        if (m[kMetaFentOver] && (rc = size_fent_buffers(s, m[kMetaNfent])) != GM_OK) return rc;
        if (m[kMetaPlOver] && (rc = size_pl_buffer(s, m[kMetaNpl])) != GM_OK) return rc;
        GM_HIP(launch_pl_count(s));
        GM_HIP(launch_pl_fill(s));
        if ((rc = read_meta(s, m)) != GM_OK) return rc;
    }
    if (m[kMetaFentOver] || m[kMetaPlOver]) return set_error(GM_ERR_KERNEL, "list build: partner lists overflow");
    const void *old_e = s->e_blk.p;
    GM_HIP(s->e_blk.ensure((size_t)std::max<int64_t>(s->max_sc * (64 / kPlAtoms), 1)));  // per group
    if (s->e_blk.p != old_e) s->buf_gen++;
    s->nfent = m[kMetaNfent];
    s->npl = m[kMetaNpl];
    s->pl_wide = m[kMetaPlWide] != 0;
    s->pl_built = true;
    return GM_OK;
}

// Partner lists for the current build (the partner layout chosen after it).
// The kernels read the build's binning and build-time positions: the lists
// are the ones the build would have made.
gm_status make_partner_lists(gm_sys *s) {
    int m[kMetaWords];
    gm_status rc;
    if (s->fent_cap == 0) {
        GM_HIP(launch_pl_count(s));
        if ((rc = read_meta(s, m)) != GM_OK) return rc;
        if ((rc = size_fent_buffers(s, m[kMetaNfent])) != GM_OK) return rc;
    }
    GM_HIP(launch_pl_count(s));
    GM_HIP(launch_pl_fill(s));
    if ((rc = read_meta(s, m)) != GM_OK) return rc;
    return fit_partner_lists(s, m);
}

// The tile list and work items of the current build.  A build with partner
// lists leaves them out; they are made here when something asks for them
// (pair-list export, list statistics, shell-mode evaluations, the tile
// layout chosen after the build), from the build's binning and build-time
// positions: the same list the build would have made.
gm_status ensure_tiles(gm_sys *s) {
    if (s->tiles_built) return GM_OK;
    if (!s->built) return set_error(GM_ERR_NOT_BUILT, "tile list before build_neighbor_list");
    int m[kMetaWords];
    gm_status rc;
    for (int pass = 0; pass < 2; pass++) {
        if (s->ent_cap == 0) {
            GM_HIP(launch_pair_count(s, 0x7fffffff, s->chunk));
            if ((rc = read_meta(s, m)) != GM_OK) return rc;
            if ((rc = size_list_buffers(s, m[kMetaNent])) != GM_OK) return rc;
            GM_HIP(launch_sc_scan(s, s->ent_cap, s->chunk));
        } else {
            GM_HIP(launch_pair_count(s, s->ent_cap, s->chunk));
        }
        GM_HIP(launch_pair_fill_and_masks(s));
        GM_HIP(launch_work_items(s));
        if ((rc = read_meta(s, m)) != GM_OK) return rc;
        if (!m[kMetaEntOver]) break;
        s->ent_cap = 0;  // more entries than the buffers hold: size them for this list
    }
    if (m[kMetaEntOver]) return set_error(GM_ERR_KERNEL, "list build: tile entries overflow");
    s->nent = m[kMetaNent];
    s->stats.n_entries = s->nent;
    if ((rc = finish_work_items(s, m)) != GM_OK) return rc;
    s->tiles_built = true;
    return GM_OK;
}

// The pair layout of a system (gm_set_pair_layout, also for a batch's
// replicas): lists of the other layout are made from the current build.
gm_status set_pair_layout(gm_sys *s, int32_t layout) {
    s->layout = layout;
    const bool pl = s->n_types > 0 && (layout == GM_LAYOUT_PARTNERS || (layout == GM_LAYOUT_AUTO && !s->auto_wide));
    if (pl == s->use_pl) return GM_OK;
    s->use_pl = pl;
    drop_build_graph(s);
    if (!s->built) return GM_OK;
    s->stats.pair_layout = pl ? GM_LAYOUT_PARTNERS : GM_LAYOUT_TILES;
    if (!pl) return ensure_tiles(s);
    gm_status rc = make_partner_lists(s);
    if (rc != GM_OK) return rc;
    s->stats.n_partner_entries = s->npl;
    return GM_OK;
}

// The list build on the device: binning, sort, split, pair search, scans,
// then the partner lists (use_pl) or the tile list's pair masks and work
// items, with one read-back at the end (the build words and counts).  The
// first build of a system also reads the entry count after the pair search,
// to size the entry buffers; later builds use that size, and a build whose
// entries overflow it (the kernels then skip the rest) sizes the buffers
// anew and runs again.
gm_status build_on_device(gm_sys *s, bool wrap, double *px, double *py, double *pz) {
    hipStream_t st = s->stream;
    s->built = false;
    s->pl_built = false;
    s->tiles_built = false;
    const bool clean = s->cols_clean;
    s->cols_clean = false;
    if (!px) { px = s->x.p; py = s->y.p; pz = s->z.p; }
    int m[kMetaWords];
    gm_status rc;
    auto column_overflow = [&]() {
        return set_error(GM_ERR_CAPACITY, "list build: an xy column holds " + std::to_string(m[2]) +
                                              " atoms, more than the LDS sort capacity " +
                                              std::to_string(kColumnCap));
    };
    GM_HIP(hipEventRecord(s->ev_b0, st));
    if (build_sized(s) && clean && !s->fit_span && !s->batch_owned) {  // (stream launches if graphs are off)
        if ((rc = launch_build_graph(s, wrap, px, py, pz)) != GM_OK) return rc;
    } else {
        if (s->fit_span && s->n > 0 && (rc = fit_span_to_atoms(s, px, py, pz)) != GM_OK) return rc;
        if (!build_sized(s)) {
            GM_HIP(launch_build(s, wrap, px, py, pz, !clean));
            if (s->use_pl) GM_HIP(launch_pl_count(s));
            else GM_HIP(launch_pair_count(s, 0x7fffffff, s->chunk));
            if ((rc = read_meta(s, m)) != GM_OK) return rc;
            if (m[1]) return column_overflow();
            if (s->use_pl) {
                if ((rc = size_fent_buffers(s, m[kMetaNfent])) != GM_OK) return rc;
                GM_HIP(launch_pl_count(s));
                GM_HIP(launch_pl_fill(s));
            } else {
                if ((rc = size_list_buffers(s, m[kMetaNent])) != GM_OK) return rc;
                GM_HIP(launch_sc_scan(s, s->ent_cap, s->chunk));
                GM_HIP(launch_pair_fill_and_masks(s));
                GM_HIP(launch_work_items(s));
            }
        } else {
            GM_HIP(launch_sized_build(s, wrap, px, py, pz, !clean));
        }
    }
    s->cols_clean = true;  // the pair-mask kernel is on the stream: it clears them
    GM_HIP(hipEventRecord(s->ev_b1, st));
    if ((rc = read_meta(s, m)) != GM_OK) return rc;
    if (m[1]) return column_overflow();
    if (!s->use_pl && m[kMetaEntOver]) {
        // more entries than the buffers hold: size them for this list and build
        // again from the system's copy of the (already wrapped) positions
        s->ent_cap = 0;
        return build_on_device(s, false);
    }
    s->nsc = m[0];
    s->ncl = s->nsc * kClustersPerSc;
    if (s->nsc > s->max_sc) return set_error(GM_ERR_KERNEL, "list build: supercluster count overflow");
    s->built = true;  // (the binning is complete: the on-demand list kernels may run)
    if (s->use_pl) {
        const bool over = m[kMetaFentOver] || m[kMetaPlOver];
        if ((rc = fit_partner_lists(s, m)) != GM_OK) { s->built = false; return rc; }
        if (over) drop_build_graph(s);  // its buffers were resized
        s->nent = 0;
        s->stats.n_entries = 0;
        s->stats.n_work_items = s->nsc * (64 / kPlAtoms);
        if (s->layout == GM_LAYOUT_AUTO && s->pl_wide) {
            // automatic layout: lists that need 32-bit slot numbers cost 4 bytes
            // per entry, and the tiles' gradual pruning does better as soon as
            // the atoms move; this system evaluates tiles from now on
            s->auto_wide = true;
            s->use_pl = false;
            s->pl_built = false;
            drop_build_graph(s);
            if ((rc = ensure_tiles(s)) != GM_OK) { s->built = false; return rc; }
        }
    } else {
        s->nent = m[kMetaNent];
        s->stats.n_entries = s->nent;
        if ((rc = finish_work_items(s, m)) != GM_OK) { s->built = false; return rc; }
        s->tiles_built = true;
    }
    s->F_dirty = false;
    s->stats_pending = true;  // pair and tile counts: summed when asked for (gm_get_stats)
    float ms = 0.f;
    GM_HIP(hipEventElapsedTime(&ms, s->ev_b0, s->ev_b1));
    s->stats.n_atoms = s->n;
    s->stats.n_superclusters = s->nsc;
    s->stats.n_clusters = s->ncl;
    s->stats.ms_build = ms;
    s->stats.pair_layout = s->use_pl ? GM_LAYOUT_PARTNERS : GM_LAYOUT_TILES;
    s->stats.n_lj_types = s->n_types;
    s->stats.n_partner_entries = s->use_pl ? s->npl : 0;
    return GM_OK;
}

// The list statistics of the last build (pairs, tiles, pairs and tiles within
// r_cutoff), summed from the pair-mask kernel's partial counts on demand.
gm_status list_stats(gm_sys *s) {
    if (!s->stats_pending) return GM_OK;
    gm_status rc = ensure_tiles(s);
    if (rc != GM_OK) return rc;
    GM_HIP(launch_list_stats(s));
    unsigned long long counters[4] = {0, 0, 0, 0};
    GM_HIP(hipMemcpyAsync(counters, s->meta.p + 8, sizeof(counters), hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    s->stats.n_list_pairs = (int64_t)counters[0];
    s->stats.n_tiles = (int64_t)counters[1];
    s->stats.n_tiles_cutoff = (int64_t)counters[2];
    s->stats.n_pairs_cutoff = (int64_t)counters[3];
    s->stats_pending = false;
    return GM_OK;
}

EvalBufs own_bufs(gm_sys *s) {
    EvalBufs b{};
    b.F = s->F.p;
    b.bad = s->bad.p;
    b.disp = s->disp.p;
    b.e_items = s->e_items.p;
    b.e_blk = s->e_blk.p;
    b.fi_blk = (s->fi_records && !pl_active(s)) ? s->fi_blk.p : nullptr;
    b.brec = pl_active(s) && s->n_contrib > 0 ? s->brec.p : nullptr;
    return b;
}

gm_status compute_on_device(gm_sys *s, const double *x, const double *y, const double *z,
                            double *fx, double *fy, double *fz, bool shell, const EvalBufs *bufs) {
    if (!s->built) return set_error(GM_ERR_NOT_BUILT, "compute_forces before build_neighbor_list");
    if (shell || !pl_active(s)) {
        gm_status rc = ensure_tiles(s);
        if (rc != GM_OK) return rc;
    }
    hipStream_t st = s->stream;
    hipEvent_t ev[kProfStride] = {nullptr, nullptr, nullptr, nullptr};
    s->ev = bufs ? *bufs : own_bufs(s);
    if (!bufs && s->profiling && (s->prof_seen++ % s->prof_every) == 0) {
        if (s->prof_used < kProfMax) {
            const size_t need = kProfStride * (size_t)(s->prof_used + 1);
            while (s->prof_pool.size() < need) {
                hipEvent_t e;
                GM_HIP(hipEventCreate(&e));
                s->prof_pool.push_back(e);
            }
            const size_t b = kProfStride * (size_t)s->prof_used;
            for (int k = 0; k < kProfStride; k++) ev[k] = s->prof_pool[b + (size_t)k];
            s->prof_used++;
        } else {
            s->prof_untimed++;
        }
    }
    // Unprofiled evaluations record no events: the stream holds only kernels.
    // Profiled ones carry no separate event packets either: the dispatches
    // stamp them, ev[0] at the prep kernel's start, ev[1] / ev[2] at the force
    // kernel's start and end, ev[3] at the finish kernel's end.
    if (!bufs) {
        if (s->F_dirty) GM_HIP(launch_fill_u64(s->F.p, (size_t)s->nsc * 64 * 3, st));
        s->F_dirty = true;
    }
    GM_HIP(launch_prep(s, x, y, z, ev[0], !shell && pl_active(s)));
    GM_HIP(launch_forces(s, x, y, z, shell, ev[1], ev[2]));
    GM_HIP(launch_finish(s, fx, fy, fz, shell, ev[3]));  // also clears the accumulator
    if (!bufs) s->F_dirty = false;
    s->last_timed = ev[0] ? s->prof_used - 1 : -1;
    return GM_OK;
}

gm_status collect(gm_sys *s, gm_energy_terms *e) {
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    GM_HIP(hipMemcpyAsync(s->pinned_e, s->e_out.p, sizeof(t), hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    std::memcpy(t, s->pinned_e, sizeof(t));
    std::memcpy(&s->last_bad, &t[6], sizeof(int));
    if (s->last_timed >= 0) {
        const size_t k = kProfStride * (size_t)s->last_timed;
        float a = 0.f, b = 0.f;
        GM_HIP(hipEventElapsedTime(&a, s->prof_pool[k + 1], s->prof_pool[k + 2]));
        GM_HIP(hipEventElapsedTime(&b, s->prof_pool[k], s->prof_pool[k + 3]));
        s->stats.ms_nonbonded = a;
        s->stats.ms_eval = b;
    }
    if (e) {
        e->bond = t[0];
        e->angle = t[1];
        e->dihedral = t[2];
        e->lj14 = t[3];
        e->lj = t[4];
        // Reference order (cpu_backend.rs:93-166): bonded terms, then the LJ sum.
        double tot = 0.0;
        tot += t[0];
        tot += t[1];
        tot += t[2];
        tot += t[3];
        tot += t[4];
        e->total = tot;
    }
    return GM_OK;
}

}  // namespace gm

extern "C" {

const char *gm_last_error(void) { return g_last_error.c_str(); }

int gm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

gm_status gm_init(int device_ordinal, gm_ctx **out) {
    GM_REQUIRE(out != nullptr, "gm_init: out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return set_error(GM_ERR_NO_DEVICE, "gm_init: no HIP device visible");
    if (device_ordinal < 0 || device_ordinal >= n)
        return set_error(GM_ERR_NO_DEVICE, "gm_init: device ordinal out of range");
    GM_HIP(hipSetDevice(device_ordinal));
    hipDeviceProp_t prop;
    GM_HIP(hipGetDeviceProperties(&prop, device_ordinal));
    std::string arch(prop.gcnArchName);
    if (arch.rfind("gfx950", 0) != 0)
        return set_error(GM_ERR_KERNEL_LOAD, "gm_init: code objects are built for gfx950, device is " + arch);
    gm_ctx *c = new gm_ctx();
    c->device = device_ordinal;
    c->n_cu = prop.multiProcessorCount;
    c->arch = arch;
    const gm_status rc = scratch_init(c);  // the device is current
    if (rc != GM_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return GM_OK;
}

gm_status gm_device_alloc(gm_ctx *ctx, size_t bytes, void **out) {
    GM_REQUIRE(ctx && out, "NULL argument");
    *out = nullptr;
    GM_HIP(hipSetDevice(ctx->device));
    GM_HIP(hipMalloc(out, bytes > 0 ? bytes : 1));
    return GM_OK;
}

gm_status gm_device_free(gm_ctx *ctx, void *p) {
    GM_REQUIRE(ctx, "NULL context");
    if (p) GM_HIP(hipFree(p));
    return GM_OK;
}

gm_status gm_copy_to_device(gm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    GM_REQUIRE(ctx && (bytes == 0 || (dst && src)), "NULL argument");
    if (bytes) GM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GM_OK;
}

gm_status gm_copy_to_host(gm_ctx *ctx, void *dst, const void *src, size_t bytes) {
    GM_REQUIRE(ctx && (bytes == 0 || (dst && src)), "NULL argument");
    if (bytes) GM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GM_OK;
}

gm_status gm_ctx_destroy(gm_ctx *ctx) {
    delete ctx;
    return GM_OK;
}

gm_status gm_system_create(gm_ctx *ctx, const gm_system_desc *d, gm_sys **out) {
    return gm::system_create_impl(ctx, d, nullptr, out);
}

}  // extern "C"

gm_status gm::system_create_impl(gm_ctx *ctx, const gm_system_desc *d, hipStream_t ext_stream,
                                 gm_sys **out) {
    GM_REQUIRE(ctx && d && out, "gm_system_create: NULL argument");
    *out = nullptr;
    const int64_t n = d->n_atoms;
    GM_REQUIRE(n >= 0 && n < (int64_t)1 << 31, "n_atoms out of range");
    GM_REQUIRE(n == 0 || (d->sigma && d->epsilon), "sigma/epsilon missing");
    for (int a = 0; a < 3; a++)
        GM_REQUIRE(std::isfinite(d->box[a]) && d->box[a] > 0.0, "box lengths must be positive");
    // config.rs:229-236: r_switch < r_cutoff < r_skin
    GM_REQUIRE(d->r_switch < d->r_cutoff && d->r_cutoff < d->r_skin && d->r_switch >= 0.0,
               "require 0 <= r_switch < r_cutoff < r_skin");
    GM_REQUIRE(d->n_bonds >= 0 && d->n_angles >= 0 && d->n_dihedrals >= 0 && d->n_excl >= 0 &&
                   d->n_pairs14 >= 0, "negative term count");
    GM_REQUIRE(d->n_bonds == 0 || (d->bond_i && d->bond_j && d->bond_k && d->bond_r0), "bond arrays missing");
    GM_REQUIRE(d->n_angles == 0 || (d->angle_i && d->angle_j && d->angle_k && d->angle_kth && d->angle_th0),
               "angle arrays missing");
    GM_REQUIRE(d->n_dihedrals == 0 || (d->dihed_i && d->dihed_j && d->dihed_k && d->dihed_l &&
                                       d->dihed_kphi && d->dihed_n && d->dihed_delta),
               "dihedral arrays missing");
    GM_REQUIRE(d->n_excl == 0 || (d->excl_i && d->excl_j), "exclusion arrays missing");
    GM_REQUIRE(d->n_pairs14 == 0 || (d->pair14_i && d->pair14_j), "1-4 arrays missing");
    GM_REQUIRE(indices_ok(d->bond_i, d->n_bonds, n) && indices_ok(d->bond_j, d->n_bonds, n),
               "bond index out of range");
    GM_REQUIRE(indices_ok(d->angle_i, d->n_angles, n) && indices_ok(d->angle_j, d->n_angles, n) &&
                   indices_ok(d->angle_k, d->n_angles, n), "angle index out of range");
    GM_REQUIRE(indices_ok(d->dihed_i, d->n_dihedrals, n) && indices_ok(d->dihed_j, d->n_dihedrals, n) &&
                   indices_ok(d->dihed_k, d->n_dihedrals, n) && indices_ok(d->dihed_l, d->n_dihedrals, n),
               "dihedral index out of range");
    GM_REQUIRE(indices_ok(d->excl_i, d->n_excl, n) && indices_ok(d->excl_j, d->n_excl, n),
               "exclusion index out of range");
    GM_REQUIRE(indices_ok(d->pair14_i, d->n_pairs14, n) && indices_ok(d->pair14_j, d->n_pairs14, n),
               "1-4 index out of range");
    for (int64_t k = 0; k < n; k++)
        GM_REQUIRE(std::isfinite(d->sigma[k]) && std::isfinite(d->epsilon[k]) && d->epsilon[k] >= 0.0,
                   "sigma/epsilon must be finite, epsilon >= 0");

    GM_HIP(hipSetDevice(ctx->device));
    gm_sys *s = new gm_sys();
    s->ctx = ctx;
    auto fail = [&](gm_status st) { gm_system_destroy(s); return st; };
#define TRY(call)                                         \
    do {                                                  \
        hipError_t _e = (call);                           \
        if (_e != hipSuccess) return fail(from_hip(_e, #call)); \
    } while (0)
    if (ext_stream) {
        s->stream = ext_stream;
        s->owns_stream = false;
    } else {
        TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    }
    for (hipEvent_t *ev : {&s->ev_b0, &s->ev_b1})
        TRY(hipEventCreate(ev));
    hipStream_t st = s->stream;
    s->n = n;
    for (int a = 0; a < 3; a++) s->box[a] = d->box[a];
    s->periodic = d->periodic ? 1 : 0;
    s->rc = d->r_cutoff;
    s->rsw = d->r_switch;
    s->rskin = d->r_skin;
    s->scale14 = d->scale_14;
    s->bonded_mi = d->bonded_min_image ? 1 : 0;
    s->nb = d->n_bonds;
    s->na = d->n_angles;
    s->nd = d->n_dihedrals;
    s->n14 = d->n_pairs14;

    s->nbp.rc2 = (float)(d->r_cutoff * d->r_cutoff);
    s->nbp.rsw2 = (float)(d->r_switch * d->r_switch);
    s->nbp.rsw = (float)d->r_switch;
    s->nbp.inv_w = (float)(1.0 / (d->r_cutoff - d->r_switch));
    s->nbp.ds_coef = (float)(-30.0 / (d->r_cutoff - d->r_switch));
    s->nbp.nrw = (float)(-d->r_switch / (d->r_cutoff - d->r_switch));
    s->nbp.nrw_in = 0.f;

    BuildParams &bp = s->bp;
    for (int a = 0; a < 3; a++) bp.box[a] = d->box[a];
    bp.r_skin2 = d->r_skin * d->r_skin;
    bp.r_cut2 = d->r_cutoff * d->r_cutoff;
    bp.r_search = (float)(d->r_skin * (1.0 + 1e-5) + 1e-3);
    bp.r_skin2f = (float)bp.r_skin2;
    bp.r_cut2f = (float)bp.r_cut2;
    bp.prune_rc = (float)d->r_cutoff;
    bp.prune_inv_step = (float)(kPruneSteps / (d->r_skin - d->r_cutoff));
    bp.r_in2f = (float)((d->r_cutoff + kPlInner) * (d->r_cutoff + kPlInner));
    bp.periodic = s->periodic;
    bp.n_atoms = (int)n;
    // xy columns about the width of a cube holding 64 atoms at the mean density
    const double vol = d->box[0] * d->box[1] * d->box[2];
    const double rho = n > 0 ? (double)n / vol : 1.0;
    const double side = std::cbrt(64.0 / rho);
    const int ncx = (int)std::max(1.0, std::min(512.0, std::floor(d->box[0] / side + 0.5)));
    const int ncy = (int)std::max(1.0, std::min(512.0, std::floor(d->box[1] / side + 0.5)));
    int ncol_max = ncx * ncy;
    s->box_ncx = ncx;
    s->box_ncy = ncy;
    s->fit_span = !s->periodic;
    if (s->fit_span) {
        // a fitted span holds the atoms in about half its columns per axis
        s->ncx_cap = (int)std::min(512.0, 4.0 * std::ceil(std::cbrt((double)n / 64.0)) + 4.0);
        ncol_max = std::max(ncol_max, s->ncx_cap * s->ncx_cap);
    }
    set_quant_span(s, d->box, ncx, ncy);
    s->max_sc = (n + 63) / 64 + ncol_max;
    bp.max_sc = (int)s->max_sc;

    const size_t N = (size_t)std::max<int64_t>(n, 1);
    const size_t slots = (size_t)s->max_sc * 64;
    TRY(upload(s->sigma, d->sigma, (size_t)n, st));
    TRY(upload(s->eps, d->epsilon, (size_t)n, st));
    std::vector<float2> par((size_t)n);
    for (int64_t k = 0; k < n; k++)
        par[(size_t)k] = make_float2((float)(0.5 * d->sigma[k]), (float)std::sqrt(d->epsilon[k]));
    TRY(upload(s->atom_par, par.data(), par.size(), st));
    {
        // distinct (sigma/2, sqrt eps) pairs in f32, the values the kernels use
        std::map<std::pair<uint32_t, uint32_t>, int> ids;
        std::vector<int> wt((size_t)n);
        std::vector<float2> tp;
        for (int64_t k = 0; k < n && (int)ids.size() <= kMaxLjTypes; k++) {
            uint32_t a, b;
            std::memcpy(&a, &par[(size_t)k].x, 4);
            std::memcpy(&b, &par[(size_t)k].y, 4);
            auto it = ids.emplace(std::make_pair(a, b), (int)ids.size()).first;
            if (it->second == (int)tp.size()) tp.push_back(make_float2(par[(size_t)k].x, 2.0f * par[(size_t)k].y));
            wt[(size_t)k] = it->second * 8;
        }
        s->n_types = (int)ids.size() <= kMaxLjTypes ? (int)tp.size() : 0;
        if (s->n_types > 0) {
            TRY(upload(s->atom_wtype, wt.data(), wt.size(), st));
            TRY(upload(s->type_par, tp.data(), tp.size(), st));
        }
        if (n == 0) {  // (a one-entry table, never read)
            tp.assign(1, make_float2(0.f, 0.f));
            TRY(upload(s->type_par, tp.data(), tp.size(), st));
            s->n_types = 1;
        }
        if (s->n_types == 0) s->use_pl = false;
    }
    TRY(upload(s->bi, d->bond_i, (size_t)s->nb, st));
    TRY(upload(s->bj, d->bond_j, (size_t)s->nb, st));
    TRY(upload(s->bk, d->bond_k, (size_t)s->nb, st));
    TRY(upload(s->br0, d->bond_r0, (size_t)s->nb, st));
    TRY(upload(s->ai, d->angle_i, (size_t)s->na, st));
    TRY(upload(s->aj, d->angle_j, (size_t)s->na, st));
    TRY(upload(s->ak, d->angle_k, (size_t)s->na, st));
    TRY(upload(s->akth, d->angle_kth, (size_t)s->na, st));
    TRY(upload(s->ath0, d->angle_th0, (size_t)s->na, st));
    TRY(upload(s->di, d->dihed_i, (size_t)s->nd, st));
    TRY(upload(s->dj, d->dihed_j, (size_t)s->nd, st));
    TRY(upload(s->dk, d->dihed_k, (size_t)s->nd, st));
    TRY(upload(s->dl, d->dihed_l, (size_t)s->nd, st));
    TRY(upload(s->dkphi, d->dihed_kphi, (size_t)s->nd, st));
    TRY(upload(s->dn, d->dihed_n, (size_t)s->nd, st));
    {
        std::vector<double2> ph((size_t)s->nd);
        for (int64_t k = 0; k < s->nd; k++) ph[(size_t)k] = make_double2(std::cos(d->dihed_delta[k]),
                                                                          std::sin(d->dihed_delta[k]));
        TRY(upload(s->dphase, ph.data(), ph.size(), st));
        TRY(hipStreamSynchronize(st));  // (ph ends with this scope)
    }
    TRY(upload(s->p14i, d->pair14_i, (size_t)s->n14, st));
    TRY(upload(s->p14j, d->pair14_j, (size_t)s->n14, st));

    // Bonded force records (a partner-list evaluation's bonded terms store each
    // contribution as its own record; the finish kernel adds an atom's records,
    // which lie together, atoms in index order).  Contributions are numbered as
    // the force kernel numbers them: 2 per bond, 3 per angle (i, k, j), 4 per
    // dihedral, 2 per scaled 1-4 pair (only when scale_14 is not 0), in term
    // order; bc_pos maps each to its record.
    {
        const int64_t n14a = (s->scale14 != 0.0) ? s->n14 : 0;
        const int64_t nc = 2 * s->nb + 3 * s->na + 4 * s->nd + 2 * n14a;
        if (nc >= ((int64_t)1 << 31)) return fail(set_error(GM_ERR_INVALID_ARG, "too many bonded terms"));
        std::vector<uint32_t> atom_of;
        atom_of.reserve((size_t)nc);
        for (int64_t k = 0; k < s->nb; k++) {
            atom_of.push_back(d->bond_i[k]);
            atom_of.push_back(d->bond_j[k]);
        }
        for (int64_t k = 0; k < s->na; k++) {
            atom_of.push_back(d->angle_i[k]);
            atom_of.push_back(d->angle_k[k]);
            atom_of.push_back(d->angle_j[k]);
        }
        for (int64_t k = 0; k < s->nd; k++) {
            atom_of.push_back(d->dihed_i[k]);
            atom_of.push_back(d->dihed_j[k]);
            atom_of.push_back(d->dihed_k[k]);
            atom_of.push_back(d->dihed_l[k]);
        }
        for (int64_t k = 0; k < n14a; k++) {
            atom_of.push_back(d->pair14_i[k]);
            atom_of.push_back(d->pair14_j[k]);
        }
        std::vector<int> boff((size_t)n + 1, 0), bpos((size_t)nc);
        for (uint32_t a : atom_of) boff[(size_t)a + 1]++;
        for (int64_t i = 0; i < n; i++) boff[(size_t)i + 1] += boff[(size_t)i];
        std::vector<int> fill(boff.begin(), boff.end() - 1);
        for (int64_t c = 0; c < nc; c++) bpos[(size_t)c] = fill[atom_of[(size_t)c]]++;
        TRY(upload(s->bc_off, boff.data(), boff.size(), st));
        TRY(upload(s->bc_pos, bpos.data(), bpos.size(), st));
        if (nc > 0) TRY(s->brec.ensure((size_t)(3 * nc)));
        s->n_contrib = nc;
        TRY(hipStreamSynchronize(st));  // (the host arrays end with this scope)
    }

    // Exclusions as a canonical set (neighbor.rs:18-23): partners > i, sorted.
    std::vector<uint64_t> keys;
    keys.reserve((size_t)d->n_excl);
    for (int64_t k = 0; k < d->n_excl; k++) {
        uint32_t a = d->excl_i[k], b = d->excl_j[k];
        if (a == b) continue;
        uint32_t lo = std::min(a, b), hi = std::max(a, b);
        keys.push_back(((uint64_t)lo << 32) | hi);
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    std::vector<int64_t> off((size_t)n + 1, 0);
    std::vector<uint32_t> nbr(keys.size());
    for (size_t k = 0; k < keys.size(); k++) {
        off[(size_t)(keys[k] >> 32) + 1]++;
        nbr[k] = (uint32_t)(keys[k] & 0xffffffffu);
    }
    for (int64_t i = 0; i < n; i++) off[(size_t)i + 1] += off[(size_t)i];
    std::vector<uint2> rng((size_t)N, make_uint2(1u, 0u));
    for (int64_t i = 0; i < n; i++)
        if (off[(size_t)i + 1] > off[(size_t)i])
            rng[(size_t)i] = make_uint2(nbr[(size_t)off[(size_t)i]], nbr[(size_t)off[(size_t)i + 1] - 1]);
    TRY(upload(s->excl_off, off.data(), off.size(), st));
    TRY(upload(s->excl_nbr, nbr.data(), nbr.size(), st));
    TRY(upload(s->excl_rng, rng.data(), rng.size(), st));
    // per atom, a bit mask of its partners i+1 .. i+64 (bit k: partner i+1+k)
    // and its largest partner when that is farther up (0 if none)
    std::vector<uint4> bits((size_t)N, make_uint4(0u, 0u, 0u, 0u));
    for (int64_t i = 0; i < n; i++) {
        uint4 &v = bits[(size_t)i];
        if (rng[(size_t)i].y > (uint64_t)i + 64u) v.z = rng[(size_t)i].y;
        for (int64_t k = off[(size_t)i]; k < off[(size_t)i + 1]; k++) {
            const uint64_t d = (uint64_t)nbr[(size_t)k] - (uint64_t)i - 1u;
            if (d < 32u) v.x |= 1u << d;
            else if (d < 64u) v.y |= 1u << (d - 32u);
        }
    }
    TRY(upload(s->excl_bits, bits.data(), bits.size(), st));
    std::vector<uint2> rng_any((size_t)N, make_uint2(0xffffffffu, 0u));
    for (uint64_t k : keys) {
        const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)(k & 0xffffffffu);
        rng_any[a] = make_uint2(std::min(rng_any[a].x, b), std::max(rng_any[a].y, b));
        rng_any[b] = make_uint2(std::min(rng_any[b].x, a), std::max(rng_any[b].y, a));
    }
    for (uint2 &r : rng_any)
        if (r.x > r.y) r = make_uint2(1u, 0u);
    TRY(upload(s->excl_rng_any, rng_any.data(), rng_any.size(), st));

    for (DevBuf<double> *b : {&s->refx, &s->refy, &s->refz, &s->sx, &s->sy, &s->sz}) TRY(b->ensure(N));
    TRY(s->pos3.ensure(3 * N));
    TRY(s->frc3.ensure(3 * N));
    DevBuf<double> *xyz[3] = {&s->x, &s->y, &s->z}, *fxyz[3] = {&s->fx, &s->fy, &s->fz};
    for (int k = 0; k < 3; k++) {
        xyz[k]->p = s->pos3.p + k * N;
        xyz[k]->n = N;
        fxyz[k]->p = s->frc3.p + k * N;
        fxyz[k]->n = N;
    }
    TRY(hipHostMalloc((void **)&s->pinned_e, 8 * sizeof(double), hipHostMallocDefault));
    TRY(s->qatom.ensure(3 * N));
    for (DevBuf<int> *b : {&s->col_count, &s->col_start, &s->col_fill}) TRY(b->ensure((size_t)ncol_max));
    TRY(s->col_sc.ensure((size_t)ncol_max + 1));
    TRY(s->ext.ensure(8));
    TRY(s->bucket.ensure(N));
    TRY(s->sorted_atoms.ensure(N));
    TRY(s->atom_slot.ensure(N));
    for (DevBuf<int> *b : {&s->sc_base, &s->sc_m, &s->sc_col, &s->sc_count, &s->sc_start, &s->item_off,
                           &s->sc_blk_off, &s->fsc_count, &s->fsc_start, &s->pl_scoff})
        TRY(b->ensure((size_t)s->max_sc + 1));
    TRY(s->pl_cnt.ensure(slots));
    TRY(s->srec.ensure(slots));
    TRY(s->pl_glen.ensure(slots / kPlAtoms));
    TRY(s->pl_goff.ensure(slots / kPlAtoms + 1));
    TRY(s->pl_gloc.ensure(slots / kPlAtoms));
    TRY(s->cost_bins.ensure(2 * (size_t)kCostBins + 2));
    if (s->max_sc > (int64_t)kWiScMask + 1) return fail(set_error(GM_ERR_CAPACITY, "more than 2^26 superclusters"));
    TRY(s->meta.ensure(kMetaWords));
    TRY(launch_fill_i32(s->meta.p, kMetaWords, 0, st));
    TRY(s->slot_atom.ensure(slots));
    TRY(s->sq.ensure(slots));
    TRY(s->sxb.ensure(slots));
    TRY(s->sp.ensure(slots));
    TRY(s->qpos_mem.ensure(slots + 2 * (size_t)kQposGuard));
    s->qpos.p = s->qpos_mem.p + kQposGuard;  // (a view: see kQposGuard)
    s->qpos.n = slots;
    TRY(s->qeps.ensure(slots));
    TRY(s->disp.ensure((size_t)kDispSlots * kDispStride));
    TRY(launch_fill_i32(reinterpret_cast<int *>(s->disp.p), (size_t)kDispSlots * kDispStride, 0, st));
    TRY(s->cl_ref.ensure((size_t)s->max_sc * 8));
    TRY(s->cl_half.ensure((size_t)s->max_sc * 8));
    TRY(s->sc_ref.ensure((size_t)s->max_sc));
    TRY(s->sc_half.ensure((size_t)s->max_sc));
    TRY(s->F.ensure(3 * slots));
    TRY(s->bad.ensure(slots));
    TRY(launch_fill_i32(reinterpret_cast<int *>(s->bad.p), slots, 0, st));
    TRY(s->e_items.ensure((size_t)(s->nb + s->na + s->nd + s->n14)));
    TRY(s->e_out.ensure(8));
    TRY(launch_fill_i32(bad_first_word(s), 1, 0x7fffffff, st));
    TRY(s->e_blk.ensure(1));
    TRY(hipStreamSynchronize(st));
#undef TRY
    *out = s;
    return GM_OK;
}

extern "C" {

gm_status gm_system_destroy(gm_sys *s) {
    if (!s) return GM_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    gm_dyn_free(s->dyn);
    s->dyn = nullptr;
    for (DevBuf<double> *v : {&s->x, &s->y, &s->z, &s->fx, &s->fy, &s->fz}) {  // views of pos3, frc3
        v->p = nullptr;
        v->n = 0;
    }
    DevBuf<double> *dbl[] = {&s->sigma, &s->eps, &s->bk, &s->br0, &s->akth, &s->ath0, &s->dkphi,
                             &s->pos3, &s->frc3, &s->refx, &s->refy, &s->refz,
                             &s->sx, &s->sy, &s->sz, &s->e_items, &s->e_blk, &s->e_out, &s->ext};
    for (auto *b : dbl) b->release();
    s->dphase.release();
    if (s->pinned_e) (void)hipHostFree(s->pinned_e);
    s->pinned_e = nullptr;
    DevBuf<uint32_t> *u32[] = {&s->bi, &s->bj, &s->ai, &s->aj, &s->ak, &s->di, &s->dj, &s->dk,
                               &s->dl, &s->dn, &s->p14i, &s->p14j, &s->excl_nbr, &s->qatom,
                               &s->ent_imask};
    for (auto *b : u32) b->release();
    DevBuf<int> *i32[] = {&s->col_count, &s->col_start, &s->col_sc, &s->col_fill, &s->bucket,
                          &s->sorted_atoms, &s->sc_base, &s->sc_m, &s->sc_col, &s->meta, &s->slot_atom,
                          &s->atom_slot, &s->sc_count, &s->sc_start, &s->ent_sc, &s->ent_cj};
    for (auto *b : i32) b->release();
    s->excl_off.release();
    s->bc_off.release();
    s->bc_pos.release();
    s->brec.release();
    s->excl_rng.release();
    s->excl_bits.release();
    s->mask_parts.release();
    s->excl_rng_any.release();
    s->atom_par.release();
    s->sp.release();
    s->sq.release();
    s->sxb.release();
    s->wi.release();
    s->fi_blk.release();
    s->sc_blk_off.release();
    s->bad.release();
    s->cl_ref.release();
    s->sc_ref.release();
    s->cl_half.release();
    s->sc_half.release();
    s->ent_meta.release();
    s->ent_dmin.release();
    // partner lists and the LJ type table
    for (DevBuf<int> *b : {&s->fsc_count, &s->fsc_start, &s->fent_sc, &s->fent_cj, &s->pl_goff, &s->pl_scoff,
                           &s->pl_gloc, &s->pl_nbr, &s->atom_wtype})
        b->release();
    s->fent_w.release();
    s->pl_cnt.release();
    s->srec.release();
    s->pl_glen.release();
    s->type_par.release();
    drop_build_graph(s);
    if (s->stage) (void)hipHostFree(s->stage);
    s->stage = nullptr;
    for (hipEvent_t &ev : s->stage_ev) {
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
    }
    s->item_off.release();
    s->wi_tmp.release();
    s->wi_cost.release();
    s->cost_bins.release();
    s->qpos.p = nullptr;  // (a view into qpos_mem)
    s->qpos.n = 0;
    s->qpos_mem.release();
    s->qeps.release();
    s->disp.release();
    s->ent_lb.release();
    s->F.release();
    for (hipEvent_t ev : {s->ev_b0, s->ev_b1})
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : s->prof_pool) (void)hipEventDestroy(ev);
    if (s->stream && s->owns_stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return GM_OK;
}

gm_status gm_build_neighbor_list(gm_sys *s, double *x, double *y, double *z) {
    GM_REQUIRE(s, "NULL system");
    GM_REQUIRE(!s->batch_owned, "gm_build_neighbor_list: a batch replica's list is built by gm_batch_build_neighbor_lists");
    GM_REQUIRE(s->n == 0 || (x && y && z), "NULL positions");
    const double *const in[3] = {x, y, z};
    double *const dev[3] = {s->x.p, s->y.p, s->z.p};
    gm_status rc = host_to_device(s, in, dev);
    if (rc != GM_OK) return rc;
    rc = build_on_device(s, s->periodic != 0);
    if (rc != GM_OK) return rc;
    if (s->periodic) {  // the wrapped positions back into the caller's arrays (neighbor.rs:13-15)
        double *const out[3] = {x, y, z};
        rc = device_to_host(s, dev, out);
        if (rc != GM_OK) return rc;
    }
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

gm_status gm_build_neighbor_list_device(gm_sys *s, double *dx, double *dy, double *dz) {
    GM_REQUIRE(s, "NULL system");
    GM_REQUIRE(!s->batch_owned,
               "gm_build_neighbor_list_device: a batch replica's list is built by gm_batch_build_neighbor_lists");
    GM_REQUIRE(s->n == 0 || (dx && dy && dz), "NULL positions");
    // the build reads the caller's arrays, wraps them in place (periodic) and
    // keeps its own copy; it ends with a synchronisation of the stream
    if (s->n == 0) return build_on_device(s, s->periodic != 0);
    return build_on_device(s, s->periodic != 0, dx, dy, dz);
}

gm_status gm_needs_rebuild(gm_sys *s, const double *x, const double *y, const double *z, int32_t *out) {
    GM_REQUIRE(s && out, "NULL argument");
    GM_REQUIRE(s->n == 0 || (x && y && z), "NULL positions");
    if (!s->built) return set_error(GM_ERR_NOT_BUILT, "needs_rebuild before build_neighbor_list");
    hipStream_t st = s->stream;
    const double *const in[3] = {x, y, z};
    double *const dev[3] = {s->sx.p, s->sy.p, s->sz.p};
    gm_status rc = host_to_device(s, in, dev);
    if (rc != GM_OK) return rc;
    GM_HIP(launch_needs_rebuild(s, s->sx.p, s->sy.p, s->sz.p));
    int flag = 0;
    GM_HIP(hipMemcpyAsync(&flag, s->meta.p + 5, sizeof(int), hipMemcpyDeviceToHost, st));
    GM_HIP(hipStreamSynchronize(st));
    *out = flag ? 1 : 0;
    return GM_OK;
}

gm_status gm_needs_rebuild_device(gm_sys *s, const double *dx, const double *dy, const double *dz,
                                  int32_t *out) {
    GM_REQUIRE(s && out, "NULL argument");
    GM_REQUIRE(s->n == 0 || (dx && dy && dz), "NULL positions");
    if (!s->built) return set_error(GM_ERR_NOT_BUILT, "needs_rebuild before build_neighbor_list");
    hipStream_t st = s->stream;
    GM_HIP(launch_needs_rebuild(s, dx, dy, dz));
    int flag = 0;
    GM_HIP(hipMemcpyAsync(&flag, s->meta.p + 5, sizeof(int), hipMemcpyDeviceToHost, st));
    GM_HIP(hipStreamSynchronize(st));
    *out = flag ? 1 : 0;
    return GM_OK;
}

gm_status gm_compute_forces(gm_sys *s, const double *x, const double *y, const double *z,
                            double *fx, double *fy, double *fz, gm_energy_terms *e) {
    GM_REQUIRE(s, "NULL system");
    GM_REQUIRE(s->n == 0 || (x && y && z && fx && fy && fz), "NULL array");
    const double *const in[3] = {x, y, z};
    double *const dev[3] = {s->x.p, s->y.p, s->z.p};
    gm_status rc = host_to_device(s, in, dev);
    if (rc != GM_OK) return rc;
    rc = compute_on_device(s, s->x.p, s->y.p, s->z.p, s->fx.p, s->fy.p, s->fz.p);
    if (rc != GM_OK) return rc;
    const double *const fdev[3] = {s->fx.p, s->fy.p, s->fz.p};
    double *const out[3] = {fx, fy, fz};
    rc = device_to_host(s, fdev, out);
    if (rc != GM_OK) return rc;
    rc = collect(s, e);
    return rc != GM_OK ? rc : nonfinite_status(s);
}

gm_status gm_compute_forces_device(gm_sys *s, const double *dx, const double *dy, const double *dz,
                                   double *dfx, double *dfy, double *dfz) {
    GM_REQUIRE(s, "NULL system");
    GM_REQUIRE(s->n == 0 || (dx && dy && dz && dfx && dfy && dfz), "NULL array");
    return compute_on_device(s, dx, dy, dz, dfx, dfy, dfz);
}

gm_status gm_set_positions(gm_sys *s, const double *x, const double *y, const double *z) {
    GM_REQUIRE(s, "NULL system");
    GM_REQUIRE(s->n == 0 || (x && y && z), "NULL positions");
    const double *const in[3] = {x, y, z};
    double *const dev[3] = {s->x.p, s->y.p, s->z.p};
    gm_status rc = host_to_device(s, in, dev);
    if (rc != GM_OK) return rc;
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

gm_status gm_compute_forces_resident(gm_sys *s) {
    GM_REQUIRE(s, "NULL system");
    return compute_on_device(s, s->x.p, s->y.p, s->z.p, s->fx.p, s->fy.p, s->fz.p);
}

gm_status gm_get_forces(gm_sys *s, double *fx, double *fy, double *fz) {
    GM_REQUIRE(s, "NULL system");
    GM_REQUIRE(s->n == 0 || (fx && fy && fz), "NULL array");
    const double *const dev[3] = {s->fx.p, s->fy.p, s->fz.p};
    double *const out[3] = {fx, fy, fz};
    return device_to_host(s, dev, out);
}

gm_status gm_sync(gm_sys *s, gm_energy_terms *e) {
    GM_REQUIRE(s, "NULL system");
    gm_status rc = collect(s, e);
    return rc != GM_OK ? rc : nonfinite_status(s);
}

gm_status gm_get_pair_list(gm_sys *s, uint32_t *pi, uint32_t *pj, int64_t cap, int64_t *nout) {
    GM_REQUIRE(s && nout, "NULL argument");
    GM_REQUIRE(cap == 0 || (pi && pj), "NULL output with nonzero capacity");
    if (!s->built) return set_error(GM_ERR_NOT_BUILT, "get_pair_list before build_neighbor_list");
    gm_status rt = ensure_tiles(s);
    if (rt != GM_OK) return rt;
    hipStream_t st = s->stream;
    const size_t nslots = (size_t)s->nsc * 64, ne = (size_t)s->nent;
    std::vector<int> slot(nslots), esc(ne), ecj(ne);
    std::vector<uint32_t> eim(ne);
    std::vector<uint8_t> lb(ne * 64);
    if (nslots) GM_HIP(hipMemcpyAsync(slot.data(), s->slot_atom.p, nslots * sizeof(int), hipMemcpyDeviceToHost, st));
    if (ne) {
        GM_HIP(hipMemcpyAsync(esc.data(), s->ent_sc.p, ne * sizeof(int), hipMemcpyDeviceToHost, st));
        GM_HIP(hipMemcpyAsync(ecj.data(), s->ent_cj.p, ne * sizeof(int), hipMemcpyDeviceToHost, st));
        GM_HIP(hipMemcpyAsync(eim.data(), s->ent_imask.p, ne * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        GM_HIP(hipMemcpyAsync(lb.data(), s->ent_lb.p, ne * 64, hipMemcpyDeviceToHost, st));
    }
    GM_HIP(hipStreamSynchronize(st));
    gm_status rs = list_stats(s);
    if (rs != GM_OK) return rs;
    std::vector<uint64_t> pairs;
    pairs.reserve((size_t)s->stats.n_list_pairs);
    for (size_t e = 0; e < ne; e++) {
        for (int ci = 0; ci < 8; ci++) {
            if (!(eim[e] & (1u << ci))) continue;
            for (int bit = 0; bit < 64; bit++) {  // lane = (jl, il): bit ci of its mask byte
                if (!((lb[e * 64 + (size_t)bit] >> ci) & 1u)) continue;
                const int il = bit & 7, jl = bit >> 3;
                const int a = slot[(size_t)esc[e] * 64 + ci * 8 + il];
                const int b = slot[(size_t)ecj[e] * 8 + jl];
                const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
                pairs.push_back(((uint64_t)lo << 32) | hi);
            }
        }
    }
    std::sort(pairs.begin(), pairs.end());
    *nout = (int64_t)pairs.size();
    const int64_t w = std::min<int64_t>(cap, (int64_t)pairs.size());
    for (int64_t k = 0; k < w; k++) {
        pi[k] = (uint32_t)(pairs[(size_t)k] >> 32);
        pj[k] = (uint32_t)(pairs[(size_t)k] & 0xffffffffu);
    }
    return GM_OK;
}

gm_status gm_profile_begin(gm_sys *s, int32_t every) {
    GM_REQUIRE(s, "NULL system");
    GM_REQUIRE(every >= 1, "gm_profile_begin: every must be >= 1");
    s->profiling = true;
    s->prof_every = every;
    s->prof_seen = 0;
    s->prof_used = 0;
    s->prof_untimed = 0;
    s->last_timed = -1;
    return GM_OK;
}

gm_status gm_profile_end(gm_sys *s, gm_profile *out) {
    GM_REQUIRE(s && out, "NULL argument");
    GM_HIP(hipStreamSynchronize(s->stream));
    gm_profile p{};
    p.n_evals = s->prof_used;
    p.n_untimed = s->prof_untimed;
    p.ms_nonbonded_min = 0.0;
    p.ms_nonbonded_max = 0.0;
    for (int64_t k = 0; k < s->prof_used; k++) {
        float nb = 0.f, ev = 0.f, pr = 0.f;
        const size_t b = kProfStride * (size_t)k;
        GM_HIP(hipEventElapsedTime(&pr, s->prof_pool[b], s->prof_pool[b + 1]));
        GM_HIP(hipEventElapsedTime(&nb, s->prof_pool[b + 1], s->prof_pool[b + 2]));
        GM_HIP(hipEventElapsedTime(&ev, s->prof_pool[b], s->prof_pool[b + 3]));
        p.ms_prep_total += pr;
        p.ms_nonbonded_total += nb;
        p.ms_eval_total += ev;
        if (k == 0 || nb < p.ms_nonbonded_min) p.ms_nonbonded_min = nb;
        if (k == 0 || nb > p.ms_nonbonded_max) p.ms_nonbonded_max = nb;
    }
    s->profiling = false;
    *out = p;
    return GM_OK;
}

gm_status gm_set_work_chunk(gm_sys *s, int32_t entries_per_item) {
    GM_REQUIRE(s && entries_per_item >= 0, "bad argument");
    GM_REQUIRE(!s->batch_owned, "gm_set_work_chunk: a batch replica's work shape is set by gm_batch_set_shape");
    s->chunk = entries_per_item;
    if (!s->built || !s->tiles_built) return GM_OK;  // (applies when the tile list is made)
    return make_work_items(s);
}

gm_status gm_set_pair_layout(gm_sys *s, int32_t layout) {
    GM_REQUIRE(s && (layout == GM_LAYOUT_TILES || layout == GM_LAYOUT_PARTNERS || layout == GM_LAYOUT_AUTO),
               "bad argument");
    GM_REQUIRE(!s->batch_owned, "gm_set_pair_layout: a batch sets its replicas' layout (gm_batch_set_pair_layout)");
    return set_pair_layout(s, layout);
}


gm_status gm_set_tile_pruning(gm_sys *s, int32_t enable) {
    GM_REQUIRE(s && (enable == 0 || enable == 1), "bad argument");
    s->prune = enable != 0;
    return GM_OK;
}

gm_status gm_get_stats(gm_sys *s, gm_stats *out) {
    GM_REQUIRE(s && out, "NULL argument");
    gm_status rc = list_stats(s);
    if (rc != GM_OK) return rc;
    *out = s->stats;
    return GM_OK;
}

}  // extern "C"
