// Internal definitions shared by the C-ABI implementation and the gfx950
// kernels.  Not part of the public interface (include/geodesic_hip.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "geodesic_hip.h"

namespace gm {

// Slot layout of the pair-tile list.  Atoms are reordered spatially into
// "slots": a supercluster is 64 consecutive slots (one wavefront of i-atoms),
// split into 8 clusters of 8 slots.  Padding slots hold atom -1.
constexpr int kClusterSize = 8;
constexpr int kClustersPerSc = 8;
constexpr int kScSlots = 64;
// Column sort capacity (atoms per xy column, sorted in LDS).
constexpr int kColumnCap = 8192;
// Longest run of list entries one nonbonded wavefront processes.
constexpr int kMaxRun = 32;
// Fixed-point block: a supercluster's list entries are cut into blocks of
// kBlock, counted from its first entry.  A block's i forces are summed in f32
// and converted to fixed point once, and its LJ energy is one f64 value, so a
// work item (a run of whole blocks) only decides which wavefront evaluates
// which blocks: results do not depend on the work shape.
constexpr int kBlock = 8;
static_assert(kMaxRun % kBlock == 0, "work items hold whole blocks");
// A fixed-point contribution of |v| * 2^32 at or above this (or not finite)
// marks its slot's force non-finite instead of being added.
constexpr float kFixedLimitF = 4.611686018427388e18f;  // 2^62
constexpr double kFixedLimit = 4.611686018427388e18;   // 2^62
// Work-item descriptor word 0: supercluster in the low 26 bits, entry count above.
constexpr int kWiEntryShift = 26;
constexpr int kWiScMask = (1 << kWiEntryShift) - 1;
static_assert(kMaxRun < (1 << (32 - kWiEntryShift)), "entry count must fit above the supercluster");
// list build: workgroups (of four wavefronts) of the pair-mask kernel, at most
constexpr int kMaskBlocks = 8192;
// list build: column histograms in LDS (32 KB) up to this many xy columns
constexpr int kLdsColumns = 8192;
// list build words in gm_sys::meta (ints): entries, entry-capacity overflow,
// entries per work item, work items, blocks
constexpr int kMetaNent = 16, kMetaEntOver = 17, kMetaChunk = 18, kMetaNwi = 19, kMetaNblk = 20;
// work-item costs (half tiles) lie in [8, 8 + kMaxRun * 17]: one sort bin each
constexpr int kCostBins = 8 + kMaxRun * 17 + 1;
// Partner lists (the default nonbonded path): every slot's listed partners,
// both directions, from the full pair search (entries (sc, cj) for every
// cluster cj near the supercluster).  build words: full entries, their
// capacity overflow, partner-list length (padding included), its overflow
constexpr int kMetaNfent = 21, kMetaFentOver = 22, kMetaNpl = 23, kMetaPlOver = 24;
// build word: some full entry's partners lie farther than 32767 slots from its
// group's first slot, so the lists hold 32-bit slot numbers (else 16-bit codes)
constexpr int kMetaPlWide = 25;
// Compact lists: a partner's code is its slot minus the group's first slot,
// plus kQposGuard (0 .. 65535).  Slot positions (qpos) sit inside an
// allocation with kQposGuard slots of padding on each side, so that any code
// decodes to an address inside it (codes read ahead past a list belong to
// other groups; their data is loaded and never used).
constexpr int kQposGuard = 32768;
constexpr int kMetaWords = 32;
// A partner-list wavefront evaluates kPlAtoms consecutive slots (a group),
// kPlLanes lanes each; lane l of a slot takes the list's entries l, l + 16, ...
constexpr int kPlAtoms = 4, kPlLanes = 16;
constexpr int kPlWaves = 4;  // wavefronts (groups) per workgroup of the partner-list kernels
constexpr int kPlPairTypes = 32;  // partner-list kernels: a type-pair table for up to this many types
// Each slot's list holds first the partners within r_cutoff + kPlInner (A)
// at the build, then the others; both parts are padded with the slot itself
// (r2 = 0: skipped) to the group's longest part rounded up to kPlPad (whole
// pairs of rounds: the force kernel takes two rounds per iteration).
constexpr int kPlPad = 2 * kPlLanes;
// The force kernel reads list codes up to three rounds past a part: the list
// buffer holds this many valid entries (slot 0) after the lists.
constexpr int kPlSlack = 4 * kPlLanes;
// The partner-list path reads LJ parameters from a table of the system's
// distinct (sigma/2, 2 sqrt eps) pairs in LDS, at most this many (else the
// system evaluates its tile list).
constexpr int kMaxLjTypes = 256;
constexpr double kPlInner = 0.1;
// Fixed-point force unit: 2^-32 kcal/mol/A per integer step.
constexpr double kFixedScale = 4294967296.0;
constexpr double kFixedInv = 1.0 / 4294967296.0;

// Per-system constants the nonbonded kernel reads.
struct NbParams {
    float sx, sy, sz;   // box_len / 2^32: A per fixed-point step, per axis
    float rc2;          // r_cutoff^2
    float rsw2;         // r_switch^2
    float rsw;          // r_switch
    float inv_w;        // 1 / (r_cutoff - r_switch)
    float ds_coef;      // -30 / (r_cutoff - r_switch)
    // the same constants with lengths in fixed-point steps (cubic boxes only)
    float inv_s;        // 2^32 / box_len
    float rc2_u, rsw2_u, rsw_u, inv_w_u, ds_coef_u;
    float fscale_u;     // force (kcal/mol per step) -> 2^-32 kcal/mol/A units
    // shell mode: the pair term is weighted by S(r) - S_in(r), S_in the same
    // quintic switch over [rin_sw, rin_sw + 1 / inv_w_in] (plain and step units)
    float rin_sw, inv_w_in, ds_coef_in;
    float rin_sw_u, inv_w_in_u, ds_coef_in_u;
    // -r_switch / (r_cutoff - r_switch) and the same for the inner switch:
    // u = r inv_w + nrw, dimensionless, so one value serves both length units
    float nrw, nrw_in;
};

// Per-system constants for the list build.
struct BuildParams {
    double box[3];
    double inv_q[3];    // 2^32 / box_len (quantisation)
    float sq[3];        // box_len / 2^32
    double r_skin2;     // r_skin^2 (f64, decision of neighbor.rs:38)
    double r_cut2;      // r_cutoff^2 (statistics only)
    float r_search;     // r_skin + margin, bounding-box search radius
    float r_skin2f, r_cut2f;   // f32 thresholds of the pair-mask pre-test
    float band_skin, band_cut; // |r2_f32 - threshold| at or below which f64 decides
    float wcol[2];      // column widths in x and y, A
    int kx, ky;         // column reach: columns more than k apart hold no pair within r_search
    float prune_rc;     // r_cutoff, A
    float prune_inv_step;  // 1 / width of one tile-distance byte step, 1/A
    float r_in2f;       // (r_cutoff + kPlInner)^2: a partner's list part
    float band_in;      // estimate at or below r_in2f + band_in: the first part
    int ncx, ncy, ncol;
    int periodic;
    int n_atoms;
    int max_sc;
};

// floor(x * 2^32 / L) mod 2^32, as a 32-bit pattern: the position as an
// unsigned fraction of the box.  t is an integer-valued double, and
// t - 2^32 floor(t / 2^32) is exact and lies in [0, 2^32) for |t| < 2^53.
__device__ inline int quantize_fixed(double x, double inv_q) {
    const double t = floor(x * inv_q);
    const double r = fma(-4294967296.0, floor(t * 2.3283064365386963e-10), t);
    return (int)(unsigned)r;
}

// Wrapping difference of two fixed-point coordinates as f32: the exact
// minimum-image displacement in fixed-point steps, rounded once.
__device__ inline float fixdiff(int a, int b) { return (float)(int)((unsigned)a - (unsigned)b); }

template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// v[l] + v[l ^ 16] (rows) or v[l] + v[l ^ 32], by the gfx950 permlane swaps.
__device__ inline double swap_self_f64(double v, bool rows) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
    const auto l = rows ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                        : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = rows ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                        : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __longlong_as_double(((long long)h[0] << 32) | l[0]) +
           __longlong_as_double(((long long)h[1] << 32) | l[1]);
}

// Sum over a full wavefront in a fixed order (lane bits 0, 1, 2, 3, 4, 5);
// every lane receives the total.
__device__ inline double wave_sum_f64(double v) {
    v += dpp_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);  // row_half_mirror: the other quad of the 8
    v += dpp_f64<0x128>(v);  // row_ror:8: the other 8 of the row
    v = swap_self_f64(v, true);
    return swap_self_f64(v, false);
}

// Tile pruning.  The list build stores, per tile, a byte b with
//   (closest listed pair of the tile at build) >= r_cutoff + b * step.
// If no atom has moved farther than d since the build, every pair of the tile
// is now at least r_cutoff + b * step - 2 d apart, so with b * step > 2 d plus
// a margin the whole tile lies beyond r_cutoff and contributes exactly zero.
// The margin (two steps and 1e-4 r_cutoff) covers the f32 rounding of the
// stored distance, of the displacement and of r2 in the force kernel.
constexpr int kPruneSteps = 254;  // bytes 0..254 span [r_cutoff, r_skin]
// The displacement bound is raised in kDispSlots words kDispStride apart (a
// block raises word blockIdx % kDispSlots) so that same-address atomics do not
// serialise; readers take the largest.
constexpr int kDispSlots = 16, kDispStride = 16;

// The device memory one force evaluation writes besides its outputs (its
// scratch): the int64 fixed-point accumulator (3 x slots), per-slot
// non-finite flags, the displacement words, bonded energies, per-block LJ
// energies and i-force records.  A system's own buffers, or a region of the
// library's static scratch array (gm_compute_forces_device_noalloc).
struct EvalBufs {
    unsigned long long *F;
    unsigned *bad;
    unsigned *disp;
    double *e_items;
    double *e_blk;
    float *fi_blk;  // null when the i forces go to F with atomics
    long long *brec;  // bonded force records, 3 x n_contrib (partner-list evaluations); null: bonded
                      // forces go to F with atomics
};
// Work counters of the batched tile kernel (one per XCD), this many words apart.
constexpr int kCtrStride = 32;

}  // namespace gm

// Device buffer with grow-only reallocation.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; n = 0; }
        hipError_t e = hipMalloc(&p, sizeof(T) * (count > 0 ? count : 1));
        if (e == hipSuccess) n = count;
        return e;
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

struct gm_ctx {
    int device = 0;
    int n_cu = 0;
    std::string arch;
    void *scratch = nullptr;   // this device's copy of the static scratch array (gm_noalloc.hip)
    size_t scratch_bytes = 0;
};

struct gm_sys {
    gm_ctx *ctx = nullptr;
    // bumped whenever an evaluation buffer that a batch's captured plan points
    // at (LJ energy records, i-force records) is reallocated: the batch then
    // makes its plan again before it launches
    uint64_t buf_gen = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = true;
    hipEvent_t ev_b0 = nullptr, ev_b1 = nullptr;

    // parameters
    int64_t n = 0;
    double box[3] = {0, 0, 0};
    int periodic = 0;
    // Span of the fixed-point coordinates per axis.  Periodic systems: the box.
    // Non-periodic systems: a cube fitted at each list build to the atoms'
    // extent (a power of two at least 2 (extent + 2 r_skin), when that fits in
    // the smallest box edge), so the 32-bit step stays fine in a large box.
    double qbox[3] = {0, 0, 0};
    bool fit_span = false;   // non-periodic: refit qbox at each list build
    int ncx_cap = 1;         // largest column count per axis the buffers hold (fitted spans)
    int box_ncx = 1, box_ncy = 1;  // column counts of a span equal to the box
    DevBuf<double> ext;      // atoms' bounding box (min xyz, max xyz) and a non-finite flag
    double rc = 0, rsw = 0, rskin = 0, scale14 = 0;
    int bonded_mi = 0;
    gm::NbParams nbp{};
    gm::BuildParams bp{};

    // topology (device, original atom order)
    int64_t nb = 0, na = 0, nd = 0, n14 = 0;
    DevBuf<double> sigma, eps;
    DevBuf<float2> atom_par;  // (sigma/2, sqrt(eps)) in f32
    DevBuf<uint32_t> bi, bj, ai, aj, ak, di, dj, dk, dl, dn, p14i, p14j;
    DevBuf<double> bk, br0, akth, ath0, dkphi;
    DevBuf<double2> dphase;  // per dihedral: cos(delta), sin(delta)
    DevBuf<int64_t> excl_off;
    DevBuf<uint32_t> excl_nbr;
    DevBuf<uint2> excl_rng;  // per atom (smallest, largest) exclusion partner above it; (1, 0) if none
    DevBuf<uint4> excl_bits; // per atom, mask of exclusion partners i+1..i+64 (bit k: partner i+1+k),
                             // then its largest partner if farther up (0 if none)
    DevBuf<unsigned long long> mask_parts;  // list statistics, 4 counts per pair-mask wavefront
    DevBuf<uint2> excl_rng_any;  // per atom (smallest, largest) exclusion partner, either side; (1, 0) if none

    // working positions (device, original order) and rebuild reference
    DevBuf<double> x, y, z, refx, refy, refz;
    DevBuf<double> fx, fy, fz;  // resident f64 forces (atom order)
    // x, y, z and fx, fy, fz are views of one allocation each (x at 0, y at n,
    // z at 2 n), so that a host array set moves in one transfer
    DevBuf<double> pos3, frc3;
    double *pinned_e = nullptr;  // 8 doubles of pinned host memory: the energy words read back
    DevBuf<double> sx, sy, sz;  // staging for needs_rebuild positions

    // spatial sort
    DevBuf<int> col_count, col_start, col_sc, col_fill, bucket, sorted_atoms;  // col_sc: ncol + 1 offsets;
                                                                               // col_count zero between builds
    DevBuf<int> sc_base, sc_m, sc_col;  // per supercluster: first sorted atom, atom count, column
    DevBuf<uint32_t> qatom;  // 3*n fixed-point coordinates, atom order
    DevBuf<int> meta;        // [0] nsc, [1] overflow, [2] max column, [5..7] rebuild test words;
                             // ints 8..15 hold u64 counters: list pairs, tiles, tiles with a pair
                             // within r_cutoff, pairs within r_cutoff (f64, build positions)
    int64_t max_sc = 0;

    // slots
    DevBuf<int> slot_atom, atom_slot;
    DevBuf<int4> sq;         // fixed-point coords per slot (x, y, z, 0) at the list build
    DevBuf<uint4> sxb;       // per slot, excl_bits of its atom; 0 for padding
    DevBuf<float2> sp;       // (sigma/2, sqrt eps) per slot
    DevBuf<int4> qpos;       // per evaluation: fixed-point coords per slot + bits of sigma/2
                             // (a view into qpos_mem, after its kQposGuard guard slots)
    DevBuf<int4> qpos_mem;
    DevBuf<float> qeps;      // 2 sqrt(eps) per slot
    DevBuf<unsigned> disp;   // largest displacement since the build (f32 bits, A), in
                             // kDispSlots words; raised by the prep kernel, cleared by the finish kernel
    bool prune = true;       // tile pruning requested (gm_set_tile_pruning)
    bool prune_geom = false; // box wide enough for the pruning bound (every edge >= 2 r_skin)
    DevBuf<int4> cl_ref;     // cluster bbox centre (fixed point) + valid flag
    DevBuf<float4> cl_half;  // cluster bbox half extents, A
    DevBuf<int4> sc_ref;
    DevBuf<float4> sc_half;

    // pair-tile list
    DevBuf<int> sc_count, sc_start;
    DevBuf<int> ent_sc, ent_cj;
    DevBuf<uint32_t> ent_imask;
    DevBuf<int2> ent_meta;                // per entry (j cluster, i-cluster mask)
    DevBuf<uint2> ent_dmin;               // per entry, per tile: closest listed pair byte
    DevBuf<uint8_t> ent_lb;               // per entry, per lane: mask bits over the 8 i clusters
    bool F_dirty = true;                  // accumulator not known to be zero
    DevBuf<int4> wi;          // work items (supercluster | entries << 26, first entry, first block,
                              // j-cluster of the first entry)
    DevBuf<float> fi_blk;                 // i forces per block (192 each, kernel force unit)
    bool fi_records = false;              // i forces as per-block records (small items) or atomics
    DevBuf<int> sc_blk_off;               // per supercluster, its first block (CSR offsets)
    int chunk = 0;  // entries per work item (0 = automatic; a multiple of kBlock)
    DevBuf<unsigned> bad;                 // per slot: force not finite this evaluation (cleared by the finish)

    // accumulators
    DevBuf<unsigned long long> F;  // 3 * n_pad, int64 fixed point, slot order
    DevBuf<double> e_items;        // per bonded term energies
    DevBuf<double> e_blk;          // per block LJ energies
    DevBuf<double> e_out;          // 5 reduced terms; the int at e_out + 6 is the first atom whose
                                   // force is not finite (INT_MAX if none), reset by the prep kernel

    // profiling: three events per timed evaluation (start, nonbonded end, end);
    // one evaluation in every prof_every is timed
    bool profiling = false;
    std::vector<hipEvent_t> prof_pool;
    int64_t prof_used = 0, prof_untimed = 0, prof_seen = 0;
    int prof_every = 1;
    int64_t last_timed = -1;  // pool index of the last profiled evaluation, -1 if unprofiled
    gm::NbParams nbp_shell{}; // shell-mode constants (set by a surrogate that owns this system)
    bool has_shell = false;   // nbp_shell in use: rebuilt from nbp whenever the span changes
    double shell_rsw = 0, shell_w = 0;  // inner switch start and width of the shell mode
    struct gm_dyn *dyn = nullptr;  // masses, constraints and dynamics buffers (gm_dynamics_set)
    // list build on the device: per-supercluster scans, work items, their sort
    DevBuf<int> item_off;                 // per supercluster, its first work item (CSR offsets)
    DevBuf<int4> wi_tmp;                  // work items in supercluster order (before the sort)
    DevBuf<int> wi_cost;                  // their estimated costs
    DevBuf<int> cost_bins;                // sort bins: counts, then offsets (heaviest first)
    int64_t ent_cap = 0;                  // entries the entry buffers hold (0: size them at the next build)
    int64_t item_cap = 0;                 // work items the work-item buffers hold
    // bonded force records, stored in atom order: per atom, the range of its
    // records (bc_off, n + 1); per term contribution (2 per bond, 3 per angle,
    // 4 per dihedral, 2 per scaled 1-4 pair, in term order) its record's
    // place (bc_pos); the records (3 x n_contrib, fixed point)
    DevBuf<int> bc_off, bc_pos;
    DevBuf<long long> brec;
    DevBuf<int2> srec;  // per slot: its atom's records [x, y) (made with the partner lists)
    int64_t n_contrib = 0;
    bool tiles_built = false;             // the tile list and work items of the last build exist
    // (a build with partner lists makes the tile list only when something
    // asks for it: pair-list export, statistics, shell-mode evaluations)
    // partner lists (built with the list when use_pl): the full pair search's
    // entries, per (full entry, lane) the listed and first-part masks over the
    // 8 j slots, per slot its two parts' counts, per group its offset and
    // padded part lengths
    bool use_pl = true;                   // nonbonded evaluation over partner lists (false: tiles)
    int layout = GM_LAYOUT_AUTO;          // the caller's choice (gm_set_pair_layout)
    bool auto_wide = false;               // automatic layout: this system's partner lists need 32-bit
                                          // slot numbers, so it evaluates tiles
    int n_types = 0;                      // distinct (sigma, eps) pairs, 0 if more than kMaxLjTypes
    DevBuf<int> atom_wtype;               // per atom: its type's byte offset in the table (8 B per type)
    DevBuf<float2> type_par;              // per type: sigma/2 (A), 2 sqrt(eps), f32
    bool pl_built = false;                // the last build made partner lists
    DevBuf<int> fsc_count, fsc_start;     // full entries per supercluster, their offsets (nsc + 1)
    DevBuf<int> fent_sc, fent_cj;
    DevBuf<uint16_t> fent_w;              // listed mask | first-part mask << 8
    DevBuf<int2> pl_cnt;                  // per slot: partners in the first part, in the second
    DevBuf<int> pl_goff;                  // per group: list offset
    DevBuf<int> pl_scoff;                 // per supercluster: list offset (nsc + 1)
    DevBuf<int> pl_gloc;                  // per group: list offset in its supercluster
    DevBuf<int2> pl_glen;                 // per group: padded first-part length, whole padded length
    DevBuf<int> pl_nbr;                   // partner slots
    int64_t fent_cap = 0, pl_cap = 0;     // capacities (0: size at the next build)
    int64_t nfent = 0, npl = 0;
    bool pl_wide = false;                 // the lists hold 32-bit slot numbers (else 16-bit codes)
    bool stats_pending = false;           // list statistics not yet summed (gm_get_stats sums them)
    bool cols_clean = false;              // column counters left at zero by the last build
    // host arrays in and out: a pinned staging buffer (3 n doubles each way) and
    // one event per array, so that the copy of one array into or out of the stage
    // overlaps the transfer of another
    double *stage = nullptr;
    size_t stage_n = 0;
    hipEvent_t stage_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool stage_in_pending = false;  // stage_ev[3] marks the end of transfers out of the stage
    hipGraph_t bgraph = nullptr;          // the captured sized build and what it was captured with
    hipGraphExec_t bexec = nullptr;
    std::vector<unsigned char> bkey;
    int bgraph_misses = 0;                // consecutive builds that needed a new capture
    bool bgraph_off = false;              // stream launches only (capture failed, or arguments keep changing)

    gm::EvalBufs ev{};  // the buffers the launchers use (own_bufs unless a caller gives others)
    bool built = false;
    int64_t nsc = 0, ncl = 0, nent = 0, n_wi = 0, n_blk = 0;
    int last_bad = 0x7fffffff;  // first atom with a non-finite force in the last collected evaluation
    bool batch_owned = false;  // a replica of a gm_batch: its list and work items belong to the batch
    gm_stats stats{};
};

// ---- error plumbing ----------------------------------------------------
namespace gm {
gm_status set_error(gm_status s, const std::string &msg);
gm_status from_hip(hipError_t e, const char *what);
}  // namespace gm

#define GM_HIP(call)                                                   \
    do {                                                               \
        hipError_t _e = (call);                                        \
        if (_e != hipSuccess) return gm::from_hip(_e, #call);          \
    } while (0)

struct gm_dyn;
void gm_dyn_free(gm_dyn *d);  // gm_dynamics.cpp

// ---- host helpers shared by the single-system and batch APIs (gm_api.cpp) --
namespace gm {
gm_status system_create_impl(gm_ctx *ctx, const gm_system_desc *d, hipStream_t ext_stream, gm_sys **out);
// positions from px, py, pz (the system's own x, y, z if null; wrapped there in
// place when wrap) into the system's x, y, z
gm_status build_on_device(gm_sys *s, bool wrap, double *px = nullptr, double *py = nullptr,
                          double *pz = nullptr);
// Forces at device positions into device arrays (f64, atom order), on the
// system's stream.  shell: LJ only, each pair weighted by S(r) - S_in(r)
// (nbp_shell), no bonded terms.
// bufs: the evaluation's scratch (null: the system's own); with bufs the
// caller has zeroed its accumulator, flags and displacement words, and no
// profiling events are recorded.
gm_status compute_on_device(gm_sys *s, const double *x, const double *y, const double *z, double *fx,
                            double *fy, double *fz, bool shell = false, const EvalBufs *bufs = nullptr);
// The system's own evaluation buffers.
EvalBufs own_bufs(gm_sys *s);
// gm_noalloc.hip: this device's address of the static scratch array into ctx
// (called by gm_init with the device current).
gm_status scratch_init(gm_ctx *ctx);
// Waits for the system's stream and reads the last evaluation's energies
// and first non-finite atom (last_bad).
gm_status collect(gm_sys *s, gm_energy_terms *e);
// GM_ERR_NUMERICAL naming the first atom with a non-finite force of the last
// collected evaluation, or GM_OK.
gm_status nonfinite_status(const gm_sys *s);
gm_status make_work_items(gm_sys *s);
// The tile list and work items of a system built with partner lists (made
// from the build's binning when first asked for).
gm_status ensure_tiles(gm_sys *s);
// The system's pair layout (GM_LAYOUT_*); lists of the other layout are made
// from the current build.
gm_status set_pair_layout(gm_sys *s, int32_t layout);
// Fixed-point span L (per axis) and xy column counts: every constant that
// depends on the length of one fixed-point step.
void set_quant_span(gm_sys *s, const double L[3], int ncx, int ncy);
// nbp_shell from nbp and the shell's inner switch (shell_rsw, shell_w).
void refresh_shell(gm_sys *s);
// cubic fixed-point span (the fast tile variant works in step units)
inline bool span_is_cubic(const gm_sys *s) { return s->qbox[0] == s->qbox[1] && s->qbox[1] == s->qbox[2]; }
// plain evaluations go over the partner lists (else the tile list)
inline bool pl_active(const gm_sys *s) { return s->use_pl && s->pl_built && s->n_types > 0; }
// LJ energy records of a plain evaluation: one per partner-list group, or one per fixed-point block
inline int64_t lj_records(const gm_sys *s) { return pl_active(s) ? s->nsc * (64 / kPlAtoms) : s->n_blk; }
// the first atom whose force was not finite in the last evaluation (INT_MAX if none)
inline int *bad_first_word(gm_sys *s) { return reinterpret_cast<int *>(s->e_out.p + 6); }
}  // namespace gm

// ---- kernel launchers (defined in the .hip files) --------------------------
namespace gm {
// gm_build.hip
// clear_cols: clear the column counters first (they are left at zero by the
// pair-mask kernel of a completed build)
hipError_t launch_build(gm_sys *s, bool wrap, double *px, double *py, double *pz, bool clear_cols);
// out[0..2] min x y z, out[3..5] max x y z, out[6] 1 if a coordinate is not finite
hipError_t launch_extent(gm_sys *s, const double *px, const double *py, const double *pz, double *out);
// entries per supercluster, then the per-supercluster scans (launch_sc_scan)
hipError_t launch_pair_count(gm_sys *s, int64_t ent_cap, int chunk_req);
// entry offsets, entry count (and capacity overflow), work shape (chunk_req
// entries per item, 0 automatic), work-item and block offsets: into meta
hipError_t launch_sc_scan(gm_sys *s, int64_t ent_cap, int chunk_req);
hipError_t launch_pair_fill_and_masks(gm_sys *s);
// partner lists: full pair search counts and their scan (full-entry count and
// overflow into meta), then the masks, per-slot counts, group offsets
// (length and overflow into meta), the lists and their padding
hipError_t launch_pl_count(gm_sys *s);
hipError_t launch_pl_fill(gm_sys *s);
// heaviest-first work items from the scanned list (after launch_sc_scan)
hipError_t launch_work_items(gm_sys *s);
// list statistics (pairs, tiles, within r_cutoff) into meta[8..15], counted on demand
hipError_t launch_list_stats(gm_sys *s);
// flag: where the answer goes (null: the system's own words, meta[5..7]):
// flag[0] rebuild needed, flag[1] the first atom whose position is not finite
// (INT_MAX if none), flag[2] the first atom whose force was not finite in the
// last evaluation (INT_MAX if none)
hipError_t launch_needs_rebuild(gm_sys *s, const double *x, const double *y, const double *z, int *flag = nullptr);
// gm_nonbonded.hip.  An evaluation is three launches:
//   prep: slot-order fixed-point positions and the displacement bound (pl, a
//     partner-list evaluation: the slot word w is the atom's type-table
//     offset instead of sigma/2);
//   forces: bonded terms + partner lists (each slot's LJ force stored to the
//     accumulator; bonded forces as records), or bonded terms + tiles into the
//     accumulator with atomics (shell: tiles only, with nbp_shell);
//   finish (gm_bonded.hip): energy reduction, f64 force output (accumulator
//     plus, after partner lists, the slot's bonded records), accumulator clear
//     (lj_only: the bonded energy segments are empty, as after a shell-mode
//     launch).
// ev_start / ev_stop (optional): stamped by the dispatch at the kernel's start and end
hipError_t launch_prep(gm_sys *s, const double *x, const double *y, const double *z, hipEvent_t ev_start = nullptr,
                       bool pl = false);
hipError_t launch_forces(gm_sys *s, const double *x, const double *y, const double *z, bool shell = false,
                         hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);
hipError_t launch_finish(gm_sys *s, double *fx, double *fy, double *fz, bool lj_only = false,
                         hipEvent_t ev_stop = nullptr);
}  // namespace gm
namespace gm {
struct ReplicaDesc;
struct TermArgs;
struct TileArgs;
TermArgs term_args(gm_sys *s, const double *x, const double *y, const double *z);
TileArgs tile_args(gm_sys *s);
struct PrepArgs;
PrepArgs prep_args(gm_sys *s, const double *x, const double *y, const double *z);
// ctr: kCtrStride * 8 words (the tile kernel's per-XCD counters, cleared by the prep)
hipError_t launch_batch_prep(const ReplicaDesc *reps, const int *atom_rep, const int64_t *atom_off,
                             int64_t n_atoms, unsigned *ctr, hipStream_t st);
hipError_t launch_batch_tiles(const ReplicaDesc *reps, const int2 *items, int n_items, int grid, unsigned *ctr,
                              hipStream_t st);
hipError_t launch_batch_bonded(const ReplicaDesc *reps, const int2 *items, int n_items, hipStream_t st);
// items: (replica, first group) per workgroup of kPlWaves partner-list groups
hipError_t launch_batch_pl(const ReplicaDesc *reps, const int2 *items, int n_items, bool pair_table,
                           hipStream_t st);
struct PlArgs;
PlArgs pl_args(gm_sys *s);
hipError_t launch_batch_finish(const ReplicaDesc *reps, int n_reps, const int2 *scl, int64_t n_rec, int64_t n_atm,
                               const int64_t *atom_off, double *eout, double *fx, double *fy, double *fz,
                               hipStream_t st);
hipError_t launch_fill_u64(unsigned long long *p, size_t n, hipStream_t st);
hipError_t launch_fill_i32(int *p, size_t n, int v, hipStream_t st);
}  // namespace gm
