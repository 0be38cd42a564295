// Batched replicas: many independent systems evaluated by one persistent
// launch, replayed from a captured hipGraph.  No reference counterpart (the
// reference evaluates one system per CpuBackend / GpuBackend).
//
// Each replica is an ordinary gm_sys (own box, topology, Verlet list) sharing
// the batch's stream.  After the lists are built, a device table of replica
// descriptors and two item lists (replica, bonded group) and (replica, tile
// work item or partner-list workgroup) are written.  k_batch_prep writes every
// replica's slot-order fixed-point positions and displacement bound; then
// k_batch_bonded runs the bonded terms on a side branch while the nonbonded
// kernel runs: k_batch_pl over every replica's partner-list groups (the
// default, when every replica's lists hold 16-bit codes), or the persistent
// k_batch_tiles walking the tile items with a fixed grid of wavefronts
// (per-XCD work counters); k_batch_finish then reduces per-replica f64
// energies and writes the concatenated f64 forces.  The layout is decided by
// the lists alone, never by timing, so a batch's bytes do not depend on it
// being tuned.  The three kernels are captured once per list build into a
// hipGraph, so a batched evaluation is one graph launch (the four kernels).  Every item runs the
// single-system arithmetic, so each replica's forces and energies are
// bit-identical to evaluating it alone with the same work shape.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gm_internal.hpp"
#include "gm_terms.hpp"

using namespace gm;

struct gm_batch {
    gm_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    std::vector<gm_sys *> reps;
    std::vector<int64_t> off;  // replica atom offsets, size R+1
    DevBuf<ReplicaDesc> d_reps;
    DevBuf<int2> d_items;  // bonded items, then tile items
    int n_bonded_items = 0;
    hipStream_t side = nullptr;  // second branch of the captured graph
    hipEvent_t fork = nullptr, join = nullptr;
    DevBuf<int> atom_rep;
    DevBuf<int64_t> atom_off;
    DevBuf<int2> scl;      // every (replica, supercluster) pair, for the finish kernel
    int64_t n_scl = 0;
    int64_t n_scl_rec = 0;  // pairs finished by four wavefronts each (the first n_scl_rec of scl)
    DevBuf<double> fx, fy, fz, E;
    DevBuf<int> bad_first;  // per replica, the first atom whose force is not finite (INT_MAX if none)
    DevBuf<unsigned> ctr;   // the tile kernel's per-XCD work counters
    int n_items = 0;
    int chunk = 0;         // entries per tile work item (0 = automatic)
    int waves_per_cu = 16; // persistent wavefronts per compute unit
    int layout = GM_LAYOUT_AUTO;  // the caller's choice (gm_batch_set_pair_layout)
    bool pl = false;              // the current plan evaluates partner lists
    bool pair_table = false;      // ... with type-pair tables (every replica has at most kPlPairTypes types)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> plan_gen;  // each replica's buffer generation when the plan was made
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool built = false;
};

#define GM_REQUIRE(cond, msg)                                   \
    do {                                                        \
        if (!(cond)) return set_error(GM_ERR_INVALID_ARG, msg); \
    } while (0)

namespace {

void drop_graph(gm_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// Descriptor table, item queue and graph for the current lists and shape.
gm_status make_plan(gm_batch *b) {
    hipStream_t st = b->stream;
    const int R = (int)b->reps.size();
    std::vector<ReplicaDesc> desc((size_t)R);
    std::vector<int2> bonded_items, tile_items;
    b->plan_gen.assign((size_t)R, 0);
    bool all_pt = true;
    for (int r = 0; r < R; r++) {
        gm_sys *s = b->reps[(size_t)r];
        ReplicaDesc &d = desc[(size_t)r];
        s->ev = own_bufs(s);
        b->plan_gen[(size_t)r] = s->buf_gen;
        d.tile = tile_args(s);
        d.prep = prep_args(s, s->x.p, s->y.p, s->z.p);
        d.terms = term_args(s, s->x.p, s->y.p, s->z.p);
        d.n_bonded_groups = (int)((d.terms.nb + d.terms.na + d.terms.nd + d.terms.n14 + 63) / 64);
        d.tile.bad_first = d.prep.bad_first = b->bad_first.p + r;
        for (int g = 0; g < d.n_bonded_groups; g++) bonded_items.push_back(make_int2(r, g));
        if (b->pl) {
            // partner lists: type-table offsets in the slot words, bonded forces
            // as records (term_args), one item per workgroup of kPlWaves groups
            d.pl = pl_args(s);
            all_pt = all_pt && d.pl.pt_shift >= 0;
            d.srec = s->srec.p;
            d.prep.wtype = s->atom_wtype.p;
            d.n_wi = 0;
            d.n_blk = (int)lj_records(s);
            for (int g = 0; g < d.pl.n_grp; g += kPlWaves) tile_items.push_back(make_int2(r, g));
        } else {
            d.terms.rec = nullptr;  // (tile evaluations: bonded forces go to the accumulator)
            d.n_wi = (int)s->n_wi;
            d.n_blk = (int)s->n_blk;
            for (int w = 0; w < d.n_wi; w++) tile_items.push_back(make_int2(r, w));
        }
    }
    std::vector<int2> items(bonded_items);
    items.insert(items.end(), tile_items.begin(), tile_items.end());
    // finish: the superclusters of replicas that keep i-force records first
    // (four wavefronts each share the records), then the others (one each)
    std::vector<int2> scl;
    for (int pass = 0; pass < 2; pass++)
        for (int r = 0; r < R; r++)
            if ((b->reps[(size_t)r]->fi_records && !b->pl) == (pass == 0))
                for (int64_t k = 0; k < b->reps[(size_t)r]->nsc; k++) scl.push_back(make_int2(r, (int)k));
    b->n_scl = (int64_t)scl.size();
    b->n_scl_rec = 0;
    for (gm_sys *s : b->reps)
        if (s->fi_records && !b->pl) b->n_scl_rec += s->nsc;
    GM_HIP(b->scl.ensure(std::max<size_t>(scl.size(), 1)));
    if (!scl.empty())
        GM_HIP(hipMemcpyAsync(b->scl.p, scl.data(), sizeof(int2) * scl.size(), hipMemcpyHostToDevice, st));
    b->pair_table = b->pl && all_pt;
    if (!b->pair_table)  // one kernel form for the whole launch: by type in every replica
        for (ReplicaDesc &d : desc) d.pl.pt_shift = -1;
    b->n_bonded_items = (int)bonded_items.size();
    b->n_items = (int)tile_items.size();
    GM_HIP(b->d_reps.ensure((size_t)R));
    GM_HIP(hipMemcpyAsync(b->d_reps.p, desc.data(), sizeof(ReplicaDesc) * (size_t)R, hipMemcpyHostToDevice, st));
    GM_HIP(b->d_items.ensure(items.size()));
    if (!items.empty())
        GM_HIP(hipMemcpyAsync(b->d_items.p, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice, st));
    GM_HIP(hipStreamSynchronize(st));

    drop_graph(b);
    const int grid = std::max(1, b->waves_per_cu * (b->ctx->n_cu > 0 ? b->ctx->n_cu : 256));
    // graph: prep -> fork -> {bonded terms on the side branch | partner lists or persistent tiles}
    // -> join -> finish
    // (the prep runs alone: sharing the machine with the bonded terms would
    // stretch it, and the tiles wait for it)
    GM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = launch_batch_prep(b->d_reps.p, b->atom_rep.p, b->atom_off.p, b->off.back(), b->ctr.p, st);
    if (e1 == hipSuccess) e1 = hipEventRecord(b->fork, st);
    if (e1 == hipSuccess) e1 = hipStreamWaitEvent(b->side, b->fork, 0);
    if (e1 == hipSuccess) e1 = launch_batch_bonded(b->d_reps.p, b->d_items.p, b->n_bonded_items, b->side);
    if (e1 == hipSuccess) e1 = hipEventRecord(b->join, b->side);
    if (e1 == hipSuccess)
        e1 = b->pl ? launch_batch_pl(b->d_reps.p, b->d_items.p + b->n_bonded_items, b->n_items, b->pair_table, st)
                   : launch_batch_tiles(b->d_reps.p, b->d_items.p + b->n_bonded_items, b->n_items, grid, b->ctr.p,
                                        st);
    if (e1 == hipSuccess) e1 = hipStreamWaitEvent(st, b->join, 0);
    if (e1 == hipSuccess)
        e1 = launch_batch_finish(b->d_reps.p, R, b->scl.p, b->n_scl_rec, b->n_scl - b->n_scl_rec, b->atom_off.p,
                                 b->E.p, b->fx.p, b->fy.p, b->fz.p, st);
    hipGraph_t g = nullptr;
    hipError_t e3 = hipStreamEndCapture(st, &g);
    if (e1 != hipSuccess) { if (g) (void)hipGraphDestroy(g); return from_hip(e1, "batched graph capture"); }
    GM_HIP(e3);
    b->graph = g;
    GM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return GM_OK;
}

// The replicas' layout after a build: partner lists when the batch's layout
// allows them and every replica's lists hold 16-bit codes (the batched kernel's
// form), else tiles for every replica.  Decided by the lists alone.
gm_status choose_layout(gm_batch *b) {
    bool all = b->layout != GM_LAYOUT_TILES;
    for (size_t r = 0; r < b->reps.size() && all; r++) {
        const gm_sys *s = b->reps[r];
        if (!pl_active(s) || s->pl_wide) {
            if (b->layout == GM_LAYOUT_PARTNERS)
                return set_error(GM_ERR_INVALID_ARG, "batched partner lists need 16-bit codes and LJ types; replica " +
                                                         std::to_string(r) + " has neither");
            all = false;
        }
    }
    if (!all)
        for (gm_sys *s : b->reps) {
            gm_status rc = set_pair_layout(s, GM_LAYOUT_TILES);
            if (rc != GM_OK) return rc;
        }
    b->pl = all;
    return GM_OK;
}

// The plan again when a replica's evaluation buffers were reallocated since
// it was made (a replica handle's pair-list export or statistics make a
// partner-list replica's tile list on demand, which can grow the LJ energy
// records the partner-list groups write).
gm_status plan_current(gm_batch *b) {
    for (size_t r = 0; r < b->reps.size(); r++)
        if (r >= b->plan_gen.size() || b->reps[r]->buf_gen != b->plan_gen[r]) return make_plan(b);
    return GM_OK;
}

gm_status apply_chunk(gm_batch *b) {
    for (gm_sys *s : b->reps) {
        s->chunk = b->chunk;
        if (s->built && !pl_active(s)) {
            gm_status rc = make_work_items(s);
            if (rc != GM_OK) return rc;
        }
    }
    return GM_OK;
}

}  // namespace

extern "C" {

gm_status gm_batch_create(gm_ctx *ctx, const gm_system_desc *descs, int32_t n_replicas, gm_batch **out) {
    GM_REQUIRE(ctx && descs && out && n_replicas > 0, "gm_batch_create: bad argument");
    for (int r = 0; r < n_replicas; r++)
        GM_REQUIRE(descs[r].box[0] == descs[r].box[1] && descs[r].box[1] == descs[r].box[2],
                   "gm_batch_create: batched replicas need cubic boxes");
    *out = nullptr;
    GM_HIP(hipSetDevice(ctx->device));
    gm_batch *b = new gm_batch();
    b->ctx = ctx;
    auto fail = [&](gm_status st) { gm_batch_destroy(b); return st; };
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(from_hip(e, "hipStreamCreate"));
    if ((e = hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking)) != hipSuccess)
        return fail(from_hip(e, "hipStreamCreate"));
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess || (e = hipEventCreate(&b->ev1)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming)) != hipSuccess)
        return fail(from_hip(e, "hipEventCreate"));
    b->off.push_back(0);
    for (int r = 0; r < n_replicas; r++) {
        gm_sys *s = nullptr;
        gm_status rc = system_create_impl(ctx, &descs[r], b->stream, &s);
        if (rc != GM_OK) return fail(rc);
        s->batch_owned = true;  // (the batch sets its layout at every build)
        b->reps.push_back(s);
        b->off.push_back(b->off.back() + descs[r].n_atoms);
    }
    const int64_t n = b->off.back();
    std::vector<int> rep((size_t)n);
    for (int r = 0; r < n_replicas; r++)
        for (int64_t a = b->off[(size_t)r]; a < b->off[(size_t)r + 1]; a++) rep[(size_t)a] = r;
    if ((e = b->atom_rep.ensure((size_t)std::max<int64_t>(n, 1))) != hipSuccess ||
        (e = b->atom_off.ensure(b->off.size())) != hipSuccess ||
        (e = b->E.ensure((size_t)n_replicas * 5)) != hipSuccess ||
        (e = b->bad_first.ensure((size_t)n_replicas)) != hipSuccess ||
        (e = b->ctr.ensure((size_t)kCtrStride * 8)) != hipSuccess)
        return fail(from_hip(e, "hipMalloc"));
    if ((e = launch_fill_i32(b->bad_first.p, (size_t)n_replicas, 0x7fffffff, b->stream)) != hipSuccess)
        return fail(from_hip(e, "fill"));
    for (DevBuf<double> *d : {&b->fx, &b->fy, &b->fz})
        if ((e = d->ensure((size_t)std::max<int64_t>(n, 1))) != hipSuccess) return fail(from_hip(e, "hipMalloc"));
    if (n > 0 && (e = hipMemcpy(b->atom_rep.p, rep.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(from_hip(e, "hipMemcpy"));
    if ((e = hipMemcpy(b->atom_off.p, b->off.data(), sizeof(int64_t) * b->off.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(from_hip(e, "hipMemcpy"));
    *out = b;
    return GM_OK;
}

gm_status gm_batch_destroy(gm_batch *b) {
    if (!b) return GM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (gm_sys *s : b->reps) gm_system_destroy(s);
    b->d_reps.release();
    b->d_items.release();
    b->atom_rep.release();
    b->atom_off.release();
    b->scl.release();
    b->fx.release();
    b->fy.release();
    b->fz.release();
    b->E.release();
    b->bad_first.release();
    b->ctr.release();
    for (hipEvent_t ev : {b->ev0, b->ev1, b->fork, b->join})
        if (ev) (void)hipEventDestroy(ev);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return GM_OK;
}

gm_status gm_batch_replica(gm_batch *b, int32_t r, gm_sys **out) {
    GM_REQUIRE(b && out && r >= 0 && r < (int32_t)b->reps.size(), "gm_batch_replica: bad argument");
    *out = b->reps[(size_t)r];
    return GM_OK;
}

gm_status gm_batch_build_neighbor_lists(gm_batch *b, double *x, double *y, double *z) {
    GM_REQUIRE(b, "NULL batch");
    const int64_t n = b->off.back();
    GM_REQUIRE(n == 0 || (x && y && z), "NULL positions");
    hipStream_t st = b->stream;
    b->built = false;
    for (size_t r = 0; r < b->reps.size(); r++) {
        gm_sys *s = b->reps[r];
        s->chunk = b->chunk;
        // every build starts from the batch's layout (a replica that needs
        // 32-bit codes turns to tiles in its own build, and choose_layout below
        // then turns the others)
        gm_status rl = set_pair_layout(s, b->layout);
        if (rl != GM_OK) return rl;
        const int64_t o = b->off[r];
        const size_t bytes = sizeof(double) * (size_t)s->n;
        if (s->n > 0) {
            GM_HIP(hipMemcpyAsync(s->x.p, x + o, bytes, hipMemcpyHostToDevice, st));
            GM_HIP(hipMemcpyAsync(s->y.p, y + o, bytes, hipMemcpyHostToDevice, st));
            GM_HIP(hipMemcpyAsync(s->z.p, z + o, bytes, hipMemcpyHostToDevice, st));
        }
        gm_status rc = build_on_device(s, s->periodic != 0);
        if (rc != GM_OK) return rc;
        if (s->periodic && s->n > 0) {
            GM_HIP(hipMemcpyAsync(x + o, s->x.p, bytes, hipMemcpyDeviceToHost, st));
            GM_HIP(hipMemcpyAsync(y + o, s->y.p, bytes, hipMemcpyDeviceToHost, st));
            GM_HIP(hipMemcpyAsync(z + o, s->z.p, bytes, hipMemcpyDeviceToHost, st));
        }
    }
    GM_HIP(hipStreamSynchronize(st));
    gm_status rc = choose_layout(b);
    if (rc != GM_OK) return rc;
    rc = make_plan(b);
    if (rc != GM_OK) return rc;
    b->built = true;
    return GM_OK;
}

gm_status gm_batch_set_positions(gm_batch *b, const double *x, const double *y, const double *z) {
    GM_REQUIRE(b, "NULL batch");
    GM_REQUIRE(b->off.back() == 0 || (x && y && z), "NULL positions");
    for (size_t r = 0; r < b->reps.size(); r++) {
        gm_sys *s = b->reps[r];
        const int64_t o = b->off[r];
        const size_t bytes = sizeof(double) * (size_t)s->n;
        if (s->n == 0) continue;
        GM_HIP(hipMemcpyAsync(s->x.p, x + o, bytes, hipMemcpyHostToDevice, b->stream));
        GM_HIP(hipMemcpyAsync(s->y.p, y + o, bytes, hipMemcpyHostToDevice, b->stream));
        GM_HIP(hipMemcpyAsync(s->z.p, z + o, bytes, hipMemcpyHostToDevice, b->stream));
    }
    GM_HIP(hipStreamSynchronize(b->stream));
    return GM_OK;
}

gm_status gm_batch_compute(gm_batch *b) {
    GM_REQUIRE(b, "NULL batch");
    if (!b->built) return set_error(GM_ERR_NOT_BUILT, "gm_batch_compute before gm_batch_build_neighbor_lists");
    gm_status rc = plan_current(b);
    if (rc != GM_OK) return rc;
    GM_HIP(hipEventRecord(b->ev0, b->stream));
    GM_HIP(hipGraphLaunch(b->exec, b->stream));
    GM_HIP(hipEventRecord(b->ev1, b->stream));
    return GM_OK;
}

gm_status gm_batch_sync(gm_batch *b, gm_energy_terms *e, float *ms) {
    GM_REQUIRE(b, "NULL batch");
    GM_HIP(hipStreamSynchronize(b->stream));
    if (ms) {
        float t = 0.f;
        GM_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
        *ms = t;
    }
    const size_t R = b->reps.size();
    std::vector<int> bad(R);
    GM_HIP(hipMemcpy(bad.data(), b->bad_first.p, sizeof(int) * R, hipMemcpyDeviceToHost));
    if (e) {
        std::vector<double> E(R * 5);
        GM_HIP(hipMemcpy(E.data(), b->E.p, sizeof(double) * R * 5, hipMemcpyDeviceToHost));
        for (size_t r = 0; r < R; r++) {
            const double *t = &E[r * 5];
            e[r].bond = t[0];
            e[r].angle = t[1];
            e[r].dihedral = t[2];
            e[r].lj14 = t[3];
            e[r].lj = t[4];
            double tot = 0.0;
            tot += t[0];
            tot += t[1];
            tot += t[2];
            tot += t[3];
            tot += t[4];
            e[r].total = tot;
        }
    }
    for (size_t r = 0; r < R; r++)
        if (bad[r] != 0x7fffffff)
            return set_error(GM_ERR_NUMERICAL, "force not finite in replica " + std::to_string(r) + " at atom " +
                                                   std::to_string(bad[r]) + " (forces written; those atoms hold NaN)");
    return GM_OK;
}

gm_status gm_batch_get_forces(gm_batch *b, double *fx, double *fy, double *fz) {
    GM_REQUIRE(b, "NULL batch");
    const int64_t n = b->off.back();
    GM_REQUIRE(n == 0 || (fx && fy && fz), "NULL array");
    GM_HIP(hipStreamSynchronize(b->stream));
    if (n > 0) {
        GM_HIP(hipMemcpy(fx, b->fx.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
        GM_HIP(hipMemcpy(fy, b->fy.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
        GM_HIP(hipMemcpy(fz, b->fz.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost));
    }
    return GM_OK;
}

gm_status gm_batch_set_shape(gm_batch *b, int32_t entries_per_item, int32_t waves_per_cu) {
    GM_REQUIRE(b && entries_per_item >= 0 && entries_per_item <= kMaxRun && waves_per_cu >= 1 &&
                   waves_per_cu <= 64, "gm_batch_set_shape: bad argument");
    b->chunk = entries_per_item;
    b->waves_per_cu = waves_per_cu;
    gm_status rc = apply_chunk(b);
    if (rc != GM_OK) return rc;
    if (b->built) return make_plan(b);
    return GM_OK;
}

gm_status gm_batch_set_pair_layout(gm_batch *b, int32_t layout) {
    GM_REQUIRE(b && (layout == GM_LAYOUT_TILES || layout == GM_LAYOUT_PARTNERS || layout == GM_LAYOUT_AUTO),
               "gm_batch_set_pair_layout: bad argument");
    b->layout = layout;
    if (!b->built) return GM_OK;
    b->built = false;  // (until the new plan is made)
    for (gm_sys *s : b->reps) {
        s->chunk = b->chunk;
        gm_status rc = set_pair_layout(s, layout);
        if (rc != GM_OK) return rc;
    }
    gm_status rc = choose_layout(b);
    if (rc != GM_OK) return rc;
    rc = make_plan(b);
    if (rc != GM_OK) return rc;
    b->built = true;
    return GM_OK;
}

gm_status gm_batch_get_pair_layout(gm_batch *b, int32_t *layout) {
    GM_REQUIRE(b && layout, "NULL argument");
    *layout = b->pl ? GM_LAYOUT_PARTNERS : GM_LAYOUT_TILES;
    return GM_OK;
}

gm_status gm_batch_autotune(gm_batch *b, int32_t *entries_per_item, int32_t *waves_per_cu, float *best_ms) {
    GM_REQUIRE(b, "NULL batch");
    if (!b->built) return set_error(GM_ERR_NOT_BUILT, "gm_batch_autotune before gm_batch_build_neighbor_lists");
    if (b->pl) {
        // partner lists have no tile shape: time the step, report shape 0 x 0
        gm_status rc = plan_current(b);
        if (rc != GM_OK) return rc;
        float total = 0.f;
        for (int it = 0; it < 8; it++) {
            GM_HIP(hipEventRecord(b->ev0, b->stream));
            GM_HIP(hipGraphLaunch(b->exec, b->stream));
            GM_HIP(hipEventRecord(b->ev1, b->stream));
            GM_HIP(hipEventSynchronize(b->ev1));
            float t = 0.f;
            GM_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
            if (it >= 2) total += t;
        }
        if (entries_per_item) *entries_per_item = 0;
        if (waves_per_cu) *waves_per_cu = 0;
        if (best_ms) *best_ms = total / 6.0f;
        return GM_OK;
    }
    static const int chunks[] = {8, 16, 24, 32};  // whole fixed-point blocks
    static const int waves[] = {8, 12, 16, 20};  // at most 5 per SIMD are resident
    float best = 1e30f;
    int bc = b->chunk, bw = b->waves_per_cu;
    for (int c : chunks) {
        for (int w : waves) {
            gm_status rc = gm_batch_set_shape(b, c, w);
            if (rc != GM_OK) return rc;
            float total = 0.f;
            for (int it = 0; it < 8; it++) {
                GM_HIP(hipEventRecord(b->ev0, b->stream));
                GM_HIP(hipGraphLaunch(b->exec, b->stream));
                GM_HIP(hipEventRecord(b->ev1, b->stream));
                GM_HIP(hipEventSynchronize(b->ev1));
                float t = 0.f;
                GM_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
                if (it >= 2) total += t;  // two warm-up replays
            }
            const float avg = total / 6.0f;
            if (avg < best) { best = avg; bc = c; bw = w; }
        }
    }
    gm_status rc = gm_batch_set_shape(b, bc, bw);
    if (rc != GM_OK) return rc;
    if (entries_per_item) *entries_per_item = bc;
    if (waves_per_cu) *waves_per_cu = bw;
    if (best_ms) *best_ms = best;
    return GM_OK;
}

}  // extern "C"
