// Device-resident dynamics, host side: masses and constraints attached to a
// system, the integrator steps on device arrays, and the reference's run loop
// (geodesic/src/lib.rs:255-330; paths relative to the reference repository
// root) with forces from the GPU force path and frames streamed to the host.
//
// SHAKE / RATTLE: when every connected component of the constraint graph
// fits a constraint group (GroupArgs), a solve is two launches and one read-
// back: each group iterates alone in LDS to its own first iteration below the
// tolerance, then every group continues to the largest of those counts, which
// is the reference's global stopping point unless some group is back above
// the tolerance there (not seen in practice; the per-iteration kernels below
// then take over from that state).
//
// The per-iteration kernels run without a host round trip per iteration:
// each iteration's kernel first compares the previous iteration's largest
// multiplier with the tolerance and raises a device flag, which makes the
// remaining launched iterations return at once (before a read-back, a
// one-thread kernel makes the same test for the last launched iteration).
// The host first reads the flag after as many iterations as the previous
// solve of the same kind needed (plus two), then every kCheckEvery
// iterations: about one round trip per solve.
// Either way the result is the reference's: same iterations, same stopping
// point, same arithmetic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gm_dynamics.hpp"

using namespace gm;
using namespace gm::dyn;

constexpr int kCtlWords = 11;  // control words: 4 per solver kind, then 3 rebuild test words

struct gm_dyn {
    int64_t n = 0, nc = 0;
    DevBuf<double> mass, dsq;
    DevBuf<uint32_t> ci, cj;
    DevBuf<int64_t> inc_off, inc_con;
    DevBuf<double> ref, scratch, lambda, force, ke;  // ref/scratch/force: 3n blocks
    DevBuf<unsigned long long> maxslot;
    DevBuf<int> flag;
    int max_iter_alloc = 0;
    int last_iters[2] = {0, 0};  // iterations the previous position / velocity solve took
    // constraint groups (none: a component exceeds a group's capacity)
    int n_groups = 0;
    int64_t n_free = 0;
    DevBuf<int> g_atom_off, g_con_off, g_inc_off, g_iters, g_conv, g_ctl;
    DevBuf<uint32_t> g_atoms, g_cons, g_con_loc, g_free;
    DevBuf<uint16_t> g_inc;
    bool groups_on = true;
    int group_cap = 0;      // > 0: phase 0 stops there (the per-iteration kernels continue)
    // Control words: g_ctl[4 kind + 2 buf + {0, 1}] (kind 0 SHAKE, 1 RATTLE;
    // two buffers per kind, alternating, each cleared by the other's phase 1),
    // g_ctl[8..10] the run loop's rebuild test words (rebuild, first atom with
    // a non-finite position, first atom with a non-finite force in the last
    // evaluation); h_ctl is their pinned copy.
    int *h_ctl = nullptr;
    int ctl_buf[2] = {0, 0};   // buffer the next solve of each kind uses
    int ctl_last[2] = {0, 0};  // buffer the last launched solve used
    bool forces_valid = false; // `force` holds the forces of the positions a run left (gm_md_run resume)
    gm_md_failure fail{};      // the last gm_md_run's NanInPos / NanInForce report
};

void gm_dyn_free(gm_dyn *d) {
    if (!d) return;
    for (DevBuf<double> *b : {&d->mass, &d->dsq, &d->ref, &d->scratch, &d->lambda, &d->force, &d->ke})
        b->release();
    d->ci.release();
    d->cj.release();
    d->inc_off.release();
    d->inc_con.release();
    d->maxslot.release();
    d->flag.release();
    for (DevBuf<int> *b : {&d->g_atom_off, &d->g_con_off, &d->g_inc_off, &d->g_iters, &d->g_conv, &d->g_ctl})
        b->release();
    for (DevBuf<uint32_t> *b : {&d->g_atoms, &d->g_cons, &d->g_con_loc, &d->g_free}) b->release();
    d->g_inc.release();
    if (d->h_ctl) (void)hipHostFree(d->h_ctl);
    delete d;
}

#define GM_REQUIRE(cond, msg)                                   \
    do {                                                        \
        if (!(cond)) return set_error(GM_ERR_INVALID_ARG, msg); \
    } while (0)

namespace {

constexpr int kCheckEvery = 4;

DynArgs dyn_args(const gm_dyn *d) {
    DynArgs D{};
    D.n = d->n;
    D.mass = d->mass.p;
    D.ci = d->ci.p;
    D.cj = d->cj.p;
    D.dsq = d->dsq.p;
    D.inc_off = d->inc_off.p;
    D.inc_con = d->inc_con.p;
    return D;
}

GroupArgs group_args(gm_dyn *d, bool velocity) {
    GroupArgs G{};
    G.n_groups = d->n_groups;
    G.atom_off = d->g_atom_off.p;
    G.atoms = d->g_atoms.p;
    G.con_off = d->g_con_off.p;
    G.cons = d->g_cons.p;
    G.con_loc = d->g_con_loc.p;
    G.inc_off = d->g_inc_off.p;
    G.inc = d->g_inc.p;
    G.free_atoms = d->g_free.p;
    G.n_free = d->n_free;
    G.iters = d->g_iters.p;
    G.conv = d->g_conv.p;
    const int k = velocity ? 1 : 0;
    G.ctl = d->g_ctl.p + 4 * k + 2 * d->ctl_buf[k];
    G.ctl_next = d->g_ctl.p + 4 * k + 2 * (d->ctl_buf[k] ^ 1);
    return G;
}

// Host-side layout of the constraint groups: connected components (union-
// find over constraint endpoints) in order of their lowest atom, packed whole
// and in that order into groups of at most kGroupAtoms atoms and kGroupCons
// constraints.  Returns false (no groups) when a component alone is larger.
struct GroupLayout {
    std::vector<int> atom_off{0}, con_off{0}, inc_off{0};
    std::vector<uint32_t> atoms, cons, con_loc, free_atoms;
    std::vector<uint16_t> inc;
};

bool make_groups(int64_t n, int64_t nc, const uint32_t *ci, const uint32_t *cj, const std::vector<int64_t> &off,
                 const std::vector<int64_t> &con, GroupLayout &L) {
    std::vector<int64_t> parent((size_t)n);
    for (int64_t a = 0; a < n; a++) parent[(size_t)a] = a;
    auto find = [&](int64_t a) {
        while (parent[(size_t)a] != a) a = parent[(size_t)a] = parent[(size_t)parent[(size_t)a]];
        return a;
    };
    for (int64_t c = 0; c < nc; c++) {
        const int64_t ra = find(ci[c]), rb = find(cj[c]);
        if (ra != rb) parent[(size_t)std::max(ra, rb)] = std::min(ra, rb);
    }
    std::vector<int64_t> comp((size_t)n, -1);  // root -> component
    std::vector<std::vector<uint32_t>> catoms, ccons;
    for (int64_t a = 0; a < n; a++) {
        if (off[(size_t)a + 1] == off[(size_t)a]) {
            L.free_atoms.push_back((uint32_t)a);
            continue;
        }
        const int64_t r = find(a);
        if (comp[(size_t)r] < 0) {
            comp[(size_t)r] = (int64_t)catoms.size();
            catoms.emplace_back();
            ccons.emplace_back();
        }
        catoms[(size_t)comp[(size_t)r]].push_back((uint32_t)a);
    }
    for (int64_t c = 0; c < nc; c++) ccons[(size_t)comp[(size_t)find(ci[c])]].push_back((uint32_t)c);
    std::vector<int> loc((size_t)n, -1), cloc((size_t)nc, -1);
    size_t k = 0;
    while (k < catoms.size()) {
        if (catoms[k].size() > (size_t)kGroupAtoms || ccons[k].size() > (size_t)kGroupCons) return false;
        size_t na = 0, ncg = 0, k1 = k;
        while (k1 < catoms.size() && na + catoms[k1].size() <= (size_t)kGroupAtoms &&
               ncg + ccons[k1].size() <= (size_t)kGroupCons) {
            na += catoms[k1].size();
            ncg += ccons[k1].size();
            k1++;
        }
        // each component's atoms and constraints are ascending, and no atom's
        // constraints span components: per atom, local order = constraint order
        const size_t a_base = L.atoms.size(), c_base = L.cons.size();
        for (size_t q = k; q < k1; q++) {
            for (uint32_t a : catoms[q]) {
                loc[a] = (int)(L.atoms.size() - a_base);
                L.atoms.push_back(a);
            }
            for (uint32_t c : ccons[q]) {
                cloc[c] = (int)(L.cons.size() - c_base);
                L.cons.push_back(c);
            }
        }
        for (size_t q = c_base; q < L.cons.size(); q++) {
            const uint32_t c = L.cons[q];
            L.con_loc.push_back((uint32_t)loc[ci[c]] | ((uint32_t)loc[cj[c]] << 16));
        }
        for (size_t q = a_base; q < L.atoms.size(); q++) {
            const uint32_t a = L.atoms[q];
            for (int64_t e = off[a]; e < off[(size_t)a + 1]; e++) {
                const int64_t c = con[(size_t)e];
                L.inc.push_back((uint16_t)((cloc[(size_t)c] << 1) | (cj[c] == a ? 1 : 0)));
            }
            L.inc_off.push_back((int)L.inc.size());
        }
        L.atom_off.push_back((int)L.atoms.size());
        L.con_off.push_back((int)L.cons.size());
        k = k1;
    }
    return true;
}

gm_status ensure_iter_buffers(gm_dyn *d, int max_iter) {
    if (d->max_iter_alloc < max_iter) {
        GM_HIP(d->maxslot.ensure((size_t)max_iter));
        d->max_iter_alloc = max_iter;
    }
    return GM_OK;
}

// The largest |multiplier| of the last iteration and its first constraint
// (constraint.rs:140-146, ties to the lowest index).
void worst_of(const std::vector<double> &lam, int64_t &idx, double &val) {
    idx = 0;
    val = 0.0;
    for (size_t c = 0; c < lam.size(); c++)
        if (std::fabs(lam[c]) > val) {
            val = std::fabs(lam[c]);
            idx = (int64_t)c;
        }
}

bool groups_active(const gm_dyn *d) { return d->nc > 0 && d->n_groups > 0 && d->groups_on; }

// Enqueue a group solve of one kind (fill of its control words, phase 0,
// phase 1), without a read-back: its control words end in
// g_ctl[2 velocity ..] on the device.
gm_status group_launch(gm_sys *s, bool velocity, Vec3 buf0, Vec3C other, int max_iter, double tol,
                       const int *guard) {
    gm_dyn *d = s->dyn;
    hipStream_t st = s->stream;
    const GroupArgs G = group_args(d, velocity);
    const DynArgs D = dyn_args(d);
    const int cap = d->group_cap > 0 ? std::min(d->group_cap, max_iter) : max_iter;
    GM_HIP(launch_group_solve(G, D, velocity, buf0, other, cap, tol, 0, d->lambda.p, st, guard));
    GM_HIP(launch_group_solve(G, D, velocity, buf0, other, cap, tol, 1, d->lambda.p, st, guard));
    const int k = velocity ? 1 : 0;
    d->ctl_last[k] = d->ctl_buf[k];
    d->ctl_buf[k] ^= 1;
    return GM_OK;
}

// The per-iteration kernels, as the reference's iterations start .. max_iter - 1
// from the state in buf0 (iteration start + j is their iteration j), to the
// first iteration below tol; GM_ERR_CONVERGENCE with the reference's report
// when there is none.  The scratch buffer is the second half of the double
// buffer; on return the result is in buf0.
gm_status iterate_from(gm_sys *s, bool velocity, Vec3 buf0, Vec3C other, int start, int max_iter, double tol,
                       int64_t step, int32_t *iters) {
    gm_dyn *d = s->dyn;
    hipStream_t st = s->stream;
    gm_status rc = ensure_iter_buffers(d, max_iter);
    if (rc != GM_OK) return rc;
    const int64_t n = d->n;
    const DynArgs D = dyn_args(d);
    Vec3 buf1{d->scratch.p, d->scratch.p + n, d->scratch.p + 2 * n};
    int flag[3] = {0, 0, 0};
    const int left = max_iter - start;
    if (left > 0) {
        GM_HIP(launch_fill_i32(d->flag.p, 3, 0, st));
        GM_HIP(launch_fill_u64(d->maxslot.p, (size_t)left, st));
        // the first read-back after as many iterations as the previous solve needed, plus two
        const int first = std::max(d->last_iters[velocity ? 1 : 0] - start, 0) + 2;
        for (int it = 0; it < left; it++) {
            const Vec3 in = (it & 1) ? buf1 : buf0, out = (it & 1) ? buf0 : buf1;
            if (velocity)
                GM_HIP(launch_rattle_iter(D, other, in, out, d->maxslot.p, it, tol, d->lambda.p, d->flag.p, st));
            else
                GM_HIP(launch_shake_iter(D, in, out, other, d->maxslot.p, it, tol, d->lambda.p, d->flag.p, st));
            if (it + 1 == first || (it + 1 > first && (it + 1 - first) % kCheckEvery == 0) || it + 1 == left) {
                // the test of this iteration (the next one's kernel would make it)
                GM_HIP(launch_check(d->maxslot.p, tol, it, d->flag.p, st));
                GM_HIP(hipMemcpyAsync(flag, d->flag.p, sizeof(flag), hipMemcpyDeviceToHost, st));
                GM_HIP(hipStreamSynchronize(st));
                if (flag[0]) break;
            }
        }
        if (flag[2] == 1) {
            const size_t b = sizeof(double) * (size_t)n;
            GM_HIP(hipMemcpyAsync(buf0.x, buf1.x, b, hipMemcpyDeviceToDevice, st));
            GM_HIP(hipMemcpyAsync(buf0.y, buf1.y, b, hipMemcpyDeviceToDevice, st));
            GM_HIP(hipMemcpyAsync(buf0.z, buf1.z, b, hipMemcpyDeviceToDevice, st));
        }
    }
    if (iters) *iters = start + flag[1];
    d->last_iters[velocity ? 1 : 0] = start + flag[1];
    if (!flag[0]) {
        std::vector<double> lam((size_t)d->nc);
        std::vector<uint32_t> hi(1), hj(1);
        GM_HIP(hipMemcpyAsync(lam.data(), d->lambda.p, sizeof(double) * lam.size(), hipMemcpyDeviceToHost, st));
        GM_HIP(hipStreamSynchronize(st));
        int64_t w = 0;
        double v = 0.0;
        worst_of(lam, w, v);
        GM_HIP(hipMemcpy(hi.data(), d->ci.p + w, sizeof(uint32_t), hipMemcpyDeviceToHost));
        GM_HIP(hipMemcpy(hj.data(), d->cj.p + w, sizeof(uint32_t), hipMemcpyDeviceToHost));
        return set_error(GM_ERR_CONVERGENCE,
                         std::string(velocity ? "velocity constraints" : "constraint solver") +
                             " did not converge at step " + std::to_string(step) + ": constraint " +
                             std::to_string(w) + " (atoms " + std::to_string(hi[0]) + ", " +
                             std::to_string(hj[0]) + "), residual " + std::to_string(v) + " after " +
                             std::to_string(max_iter) + " iterations");
    }
    return GM_OK;
}

// The end of a group solve of one kind whose control words are in h_ctl:
// the count every group reached is the reference's stopping point when no
// group is above tol there; otherwise the per-iteration kernels continue
// from that state.
gm_status group_finish(gm_sys *s, bool velocity, Vec3 buf0, Vec3C other, int max_iter, double tol, int64_t step,
                       int32_t *iters) {
    gm_dyn *d = s->dyn;
    const int k = velocity ? 1 : 0;
    const int *c = d->h_ctl + 4 * k + 2 * d->ctl_last[k];
    if (!c[1]) {
        if (iters) *iters = c[0];
        d->last_iters[velocity ? 1 : 0] = c[0];
        return GM_OK;
    }
    return iterate_from(s, velocity, buf0, other, c[0], max_iter, tol, step, iters);
}

// Every control word (and g_ctl[8..10], where the run loop's rebuild test writes) into h_ctl, one copy.
gm_status read_ctl(gm_sys *s) {
    gm_dyn *d = s->dyn;
    GM_HIP(hipMemcpyAsync(d->h_ctl, d->g_ctl.p, kCtlWords * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

// SHAKE (velocity = false) or RATTLE (velocity = true) on the 3n-block
// `buf0` (the caller's arrays): the group solve when there are groups, the
// per-iteration kernels otherwise.  On return the result is in buf0.
gm_status project(gm_sys *s, bool velocity, Vec3 buf0, Vec3C other, int max_iter, double tol, int64_t step,
                  int32_t *iters) {
    gm_dyn *d = s->dyn;
    if (d->nc == 0) {
        if (iters) *iters = 0;
        return GM_OK;
    }
    if (!groups_active(d)) return iterate_from(s, velocity, buf0, other, 0, max_iter, tol, step, iters);
    gm_status rc = group_launch(s, velocity, buf0, other, max_iter, tol, nullptr);
    if (rc != GM_OK || (rc = read_ctl(s)) != GM_OK) return rc;
    return group_finish(s, velocity, buf0, other, max_iter, tol, step, iters);
}

gm_status drift_and_constrain(gm_sys *s, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                              double dt, int max_iter, double tol, int64_t step, int32_t *iters) {
    gm_dyn *d = s->dyn;
    const int64_t n = d->n;
    double *rx = d->ref.p, *ry = d->ref.p + n, *rz = d->ref.p + 2 * n;
    GM_HIP(launch_drift(n, x, y, z, vx, vy, vz, rx, ry, rz, dt, s->stream));
    gm_status rc = project(s, false, Vec3{x, y, z}, Vec3C{rx, ry, rz}, max_iter, tol, step, iters);
    if (rc != GM_OK) return rc;
    GM_HIP(launch_resync(n, x, y, z, rx, ry, rz, vx, vy, vz, dt, s->stream));
    return GM_OK;
}

// 0.5 * sum_i m v^2 / 20.455^2 in atom order (geodesic/src/lib.rs:68-77).
gm_status kinetic_energy(gm_sys *s, const double *vx, const double *vy, const double *vz, double *out) {
    gm_dyn *d = s->dyn;
    GM_HIP(launch_kinetic_terms(d->n, vx, vy, vz, d->mass.p, d->ke.p, s->stream));
    std::vector<double> t((size_t)d->n);
    if (d->n > 0)
        GM_HIP(hipMemcpyAsync(t.data(), d->ke.p, sizeof(double) * t.size(), hipMemcpyDeviceToHost, s->stream));
    GM_HIP(hipStreamSynchronize(s->stream));
    double sum = 0.0;
    for (double v : t) sum += v;
    *out = 0.5 * sum * (1.0 / (20.455 * 20.455));
    return GM_OK;
}

gm_status frame_to_host(gm_sys *s, const double *x, const double *y, const double *z, float *dst) {
    const size_t n = (size_t)s->n;
    std::vector<double> h(3 * n);
    if (n > 0) {
        GM_HIP(hipMemcpyAsync(h.data(), x, 8 * n, hipMemcpyDeviceToHost, s->stream));
        GM_HIP(hipMemcpyAsync(h.data() + n, y, 8 * n, hipMemcpyDeviceToHost, s->stream));
        GM_HIP(hipMemcpyAsync(h.data() + 2 * n, z, 8 * n, hipMemcpyDeviceToHost, s->stream));
    }
    GM_HIP(hipStreamSynchronize(s->stream));
    for (size_t k = 0; k < 3 * n; k++) dst[k] = (float)h[k];
    return GM_OK;
}

}  // namespace

extern "C" {

gm_status gm_dynamics_set(gm_sys *s, const double *mass, int64_t n_constr, const uint32_t *ci, const uint32_t *cj,
                          const double *dsq) {
    GM_REQUIRE(s && (s->n == 0 || mass) && n_constr >= 0 && (n_constr == 0 || (ci && cj && dsq)),
               "gm_dynamics_set: bad argument");
    GM_REQUIRE(!s->batch_owned, "gm_dynamics_set: not available on a batch replica");
    const int64_t n = s->n;
    for (int64_t i = 0; i < n; i++)
        GM_REQUIRE(std::isfinite(mass[i]) && mass[i] > 0.0, "gm_dynamics_set: masses must be positive");
    for (int64_t c = 0; c < n_constr; c++)
        GM_REQUIRE(ci[c] < (uint64_t)n && cj[c] < (uint64_t)n && ci[c] != cj[c],
                   "gm_dynamics_set: constraint atom index out of range");
    gm_dyn_free(s->dyn);
    s->dyn = nullptr;
    gm_dyn *d = new gm_dyn();
    d->n = n;
    d->nc = n_constr;
    // per-atom incident constraints, ascending constraint index
    std::vector<int64_t> off((size_t)n + 1, 0), con((size_t)(2 * n_constr));
    for (int64_t c = 0; c < n_constr; c++) {
        off[(size_t)ci[c] + 1]++;
        off[(size_t)cj[c] + 1]++;
    }
    for (int64_t i = 0; i < n; i++) off[(size_t)i + 1] += off[(size_t)i];
    std::vector<int64_t> fill(off.begin(), off.end() - 1);
    for (int64_t c = 0; c < n_constr; c++) {
        con[(size_t)fill[ci[c]]++] = c;
        con[(size_t)fill[cj[c]]++] = c;
    }
    const size_t N = (size_t)std::max<int64_t>(n, 1), C = (size_t)std::max<int64_t>(n_constr, 1);
    auto fail = [&](hipError_t e) {
        gm_dyn_free(d);
        return from_hip(e, "gm_dynamics_set");
    };
    hipError_t e;
    if ((e = d->mass.ensure(N)) != hipSuccess || (e = d->dsq.ensure(C)) != hipSuccess ||
        (e = d->ci.ensure(C)) != hipSuccess || (e = d->cj.ensure(C)) != hipSuccess ||
        (e = d->inc_off.ensure(N + 1)) != hipSuccess || (e = d->inc_con.ensure(2 * C)) != hipSuccess ||
        (e = d->ref.ensure(3 * N)) != hipSuccess || (e = d->scratch.ensure(3 * N)) != hipSuccess ||
        (e = d->lambda.ensure(C)) != hipSuccess || (e = d->force.ensure(3 * N)) != hipSuccess ||
        (e = d->ke.ensure(N)) != hipSuccess || (e = d->flag.ensure(3)) != hipSuccess)
        return fail(e);
    if (n > 0 && (e = hipMemcpy(d->mass.p, mass, 8 * (size_t)n, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(d->inc_off.p, off.data(), 8 * off.size(), hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if (n_constr > 0) {
        if ((e = hipMemcpy(d->ci.p, ci, 4 * (size_t)n_constr, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(d->cj.p, cj, 4 * (size_t)n_constr, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(d->dsq.p, dsq, 8 * (size_t)n_constr, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(d->inc_con.p, con.data(), 8 * con.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e);
        GroupLayout L;
        if (make_groups(n, n_constr, ci, cj, off, con, L)) {
            const int G = (int)L.atom_off.size() - 1;
            auto up = [&](auto &buf, const auto &v) {
                hipError_t r = buf.ensure(std::max<size_t>(v.size(), 1));
                if (r == hipSuccess && !v.empty())
                    r = hipMemcpy(buf.p, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice);
                return r;
            };
            if ((e = up(d->g_atom_off, L.atom_off)) != hipSuccess || (e = up(d->g_con_off, L.con_off)) != hipSuccess ||
                (e = up(d->g_inc_off, L.inc_off)) != hipSuccess || (e = up(d->g_atoms, L.atoms)) != hipSuccess ||
                (e = up(d->g_cons, L.cons)) != hipSuccess || (e = up(d->g_con_loc, L.con_loc)) != hipSuccess ||
                (e = up(d->g_inc, L.inc)) != hipSuccess || (e = up(d->g_free, L.free_atoms)) != hipSuccess ||
                (e = d->g_iters.ensure((size_t)G)) != hipSuccess || (e = d->g_conv.ensure((size_t)G)) != hipSuccess ||
                (e = d->g_ctl.ensure(kCtlWords)) != hipSuccess ||
                (e = hipMemset(d->g_ctl.p, 0, kCtlWords * sizeof(int))) != hipSuccess ||  // every buffer starts clear
                (e = hipHostMalloc((void **)&d->h_ctl, kCtlWords * sizeof(int), hipHostMallocDefault)) != hipSuccess)
                return fail(e);
            d->n_groups = G;
            d->n_free = (int64_t)L.free_atoms.size();
        }
    }
    s->dyn = d;
    return GM_OK;
}

gm_status gm_set_constraint_groups(gm_sys *s, int32_t enable, int32_t iteration_cap, int32_t *n_groups) {
    GM_REQUIRE(s && s->dyn && (enable == 0 || enable == 1) && iteration_cap >= 0,
               "gm_set_constraint_groups: no dynamics attached or bad argument");
    s->dyn->groups_on = enable != 0;
    s->dyn->group_cap = iteration_cap;
    if (n_groups) *n_groups = s->dyn->n_groups;
    return GM_OK;
}

gm_status gm_half_kick_device(gm_sys *s, double *vx, double *vy, double *vz, const double *fx, const double *fy,
                              const double *fz, double dt_half) {
    GM_REQUIRE(s && s->dyn, "gm_half_kick_device: no dynamics attached (gm_dynamics_set)");
    GM_REQUIRE(s->n == 0 || (vx && vy && vz && fx && fy && fz), "gm_half_kick_device: NULL array");
    GM_HIP(launch_half_kick(s->n, vx, vy, vz, fx, fy, fz, s->dyn->mass.p, dt_half, s->stream));
    return GM_OK;
}

gm_status gm_drift_and_constrain_device(gm_sys *s, double *x, double *y, double *z, double *vx, double *vy,
                                        double *vz, double dt, int32_t max_iter, double tol, int64_t step,
                                        int32_t *iters) {
    GM_REQUIRE(s && s->dyn, "gm_drift_and_constrain_device: no dynamics attached (gm_dynamics_set)");
    GM_REQUIRE(s->n == 0 || (x && y && z && vx && vy && vz), "gm_drift_and_constrain_device: NULL array");
    GM_REQUIRE(max_iter > 0 && dt != 0.0, "gm_drift_and_constrain_device: need max_iter > 0, dt != 0");
    gm_status rc = drift_and_constrain(s, x, y, z, vx, vy, vz, dt, max_iter, tol, step, iters);
    if (rc != GM_OK) return rc;
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

gm_status gm_constrain_velocities_device(gm_sys *s, const double *x, const double *y, const double *z, double *vx,
                                         double *vy, double *vz, int32_t max_iter, double tol, int64_t step,
                                         int32_t *iters) {
    GM_REQUIRE(s && s->dyn, "gm_constrain_velocities_device: no dynamics attached (gm_dynamics_set)");
    GM_REQUIRE(s->n == 0 || (x && y && z && vx && vy && vz), "gm_constrain_velocities_device: NULL array");
    GM_REQUIRE(max_iter > 0, "gm_constrain_velocities_device: need max_iter > 0");
    gm_status rc = project(s, true, Vec3{vx, vy, vz}, Vec3C{x, y, z}, max_iter, tol, step, iters);
    if (rc != GM_OK) return rc;
    GM_HIP(hipStreamSynchronize(s->stream));
    return GM_OK;
}

gm_status gm_md_run(gm_sys *s, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                    const gm_md_desc *md, float *h_frames, double *h_epot, double *h_ekin, int64_t *n_frames) {
    GM_REQUIRE(s && s->dyn && md, "gm_md_run: no dynamics attached (gm_dynamics_set) or NULL description");
    GM_REQUIRE(s->n == 0 || (x && y && z && vx && vy && vz), "gm_md_run: NULL array");
    GM_REQUIRE(md->n_steps >= 0 && md->frame_interval > 0 && md->max_iter > 0 && md->dt != 0.0 &&
                   md->first_step >= 0 && (md->resume == 0 || md->resume == 1),
               "gm_md_run: need n_steps >= 0, frame_interval > 0, max_iter > 0, dt != 0, first_step >= 0");
    GM_REQUIRE(!md->resume || (s->built && s->dyn->forces_valid),
               "gm_md_run: resume needs a previous run on this system");
    const int64_t n = s->n;
    gm_dyn *d = s->dyn;
    d->fail = gm_md_failure{};
    d->forces_valid = false;
    double *fx = d->force.p, *fy = fx + n, *fz = fx + 2 * n;
    const double dt = md->dt;
    const int64_t s0 = md->first_step;
    int64_t frames = 0;
    if (n_frames) *n_frames = 0;
    auto done = [&](gm_status r) {
        if (n_frames) *n_frames = frames;
        return r;
    };
    auto record = [&](double epot) -> gm_status {
        double ke = 0.0;
        gm_status r = kinetic_energy(s, vx, vy, vz, &ke);
        if (r != GM_OK) return r;
        if (h_frames && (r = frame_to_host(s, x, y, z, h_frames + (size_t)frames * 3 * (size_t)n)) != GM_OK)
            return r;
        if (h_epot) h_epot[frames] = epot;
        if (h_ekin) h_ekin[frames] = ke;
        frames++;
        return GM_OK;
    };
    // NumericalError (error.rs) with the reference's message (lib.rs:79-107)
    auto nan_in_force = [&](int64_t step, int atom) {
        d->fail = gm_md_failure{2, 0, step, atom};  // a flagged atom's force is NaN in x, y and z
        return set_error(GM_ERR_NUMERICAL, "NaN in force_X at step " + std::to_string(step) + ", atom " +
                                               std::to_string(atom));
    };
    auto nan_in_pos = [&](int64_t step, int atom) -> gm_status {
        double p[3];
        const double *src[3] = {x, y, z};
        for (int c = 0; c < 3; c++) GM_HIP(hipMemcpy(&p[c], src[c] + atom, sizeof(double), hipMemcpyDeviceToHost));
        const int comp = !std::isfinite(p[0]) ? 0 : (!std::isfinite(p[1]) ? 1 : 2);
        d->fail = gm_md_failure{1, comp, step, atom};
        return set_error(GM_ERR_NUMERICAL, std::string("NaN in position_") + "XYZ"[comp] + " at step " +
                                               std::to_string(step) + ", atom " + std::to_string(atom));
    };
    constexpr int kNone = 0x7fffffff;
    gm_energy_terms e{};
    gm_status rc;
    if (!md->resume) {
        // lib.rs:261-283: list, forces, NaN check, velocity projection, frame 0
        if ((rc = gm_build_neighbor_list_device(s, x, y, z)) != GM_OK) return rc;
        if ((rc = compute_on_device(s, x, y, z, fx, fy, fz)) != GM_OK) return rc;
        if ((rc = collect(s, &e)) != GM_OK) return rc;
        if (s->last_bad != kNone) return done(nan_in_force(s0, s->last_bad));
        if ((rc = project(s, true, Vec3{vx, vy, vz}, Vec3C{x, y, z}, md->max_iter, md->tol, s0, nullptr)) != GM_OK)
            return done(rc);
        if ((rc = record(e.total)) != GM_OK) return done(rc);
    }
    // lib.rs:292-333: B, A (drift + SHAKE), NaN check of positions, rebuild if
    // needed, forces, NaN check of forces, B, RATTLE.
    // With constraint groups a step has one read-back: a step's RATTLE and
    // force check are read together with the next step's SHAKE, position
    // check and rebuild test.  The next step's kernels up to that point are
    // guarded by RATTLE's "continue" word, so when RATTLE must continue per
    // iteration none of them has changed anything: RATTLE continues, and the
    // step runs again.  Failures are reported in the reference's order: the
    // previous step's forces, its RATTLE, this step's SHAKE, its positions.
    hipStream_t st = s->stream;
    double *rx = d->ref.p, *ry = rx + n, *rz = rx + 2 * n;
    const bool deferred = groups_active(d);
    bool v_pending = false;  // the previous step's RATTLE is launched, not yet checked
    int words[3];            // rebuild test words of the per-iteration path
    auto rebuild_test = [&]() -> gm_status {
        GM_HIP(launch_needs_rebuild(s, x, y, z));
        GM_HIP(hipMemcpyAsync(words, s->meta.p + 5, sizeof(words), hipMemcpyDeviceToHost, st));
        GM_HIP(hipStreamSynchronize(st));
        return GM_OK;
    };
    for (int64_t k = 0; k < md->n_steps; k++) {
        const int64_t step = s0 + k;
        int32_t rebuild = 0;
        bool stepped = false;
        if (deferred) {
            const int *guard = v_pending ? d->g_ctl.p + 4 + 2 * d->ctl_last[1] + 1 : nullptr;
            GM_HIP(launch_half_kick(n, vx, vy, vz, fx, fy, fz, d->mass.p, dt / 2.0, st, guard));
            GM_HIP(launch_drift(n, x, y, z, vx, vy, vz, rx, ry, rz, dt, st, guard));
            if ((rc = group_launch(s, false, Vec3{x, y, z}, Vec3C{rx, ry, rz}, md->max_iter, md->tol, guard)) !=
                GM_OK)
                return done(rc);
            GM_HIP(launch_resync(n, x, y, z, rx, ry, rz, vx, vy, vz, dt, st, guard));
            GM_HIP(launch_needs_rebuild(s, x, y, z, d->g_ctl.p + 8));
            if ((rc = read_ctl(s)) != GM_OK) return done(rc);
            if (v_pending && d->h_ctl[10] != kNone) return done(nan_in_force(step - 1, d->h_ctl[10]));
            const bool replay = v_pending && d->h_ctl[4 + 2 * d->ctl_last[1] + 1];
            // the guarded SHAKE did nothing and cleared nothing: its buffer, still clear, is used again
            if (replay) d->ctl_buf[0] = d->ctl_last[0];
            if (v_pending) {
                rc = group_finish(s, true, Vec3{vx, vy, vz}, Vec3C{x, y, z}, md->max_iter, md->tol, step - 1, nullptr);
                if (rc != GM_OK) return done(rc);
                v_pending = false;
            }
            if (!replay) {
                // SHAKE continues per iteration: resync and test again after it
                if (d->h_ctl[2 * d->ctl_last[0] + 1]) {
                    rc = group_finish(s, false, Vec3{x, y, z}, Vec3C{rx, ry, rz}, md->max_iter, md->tol, step,
                                      nullptr);
                    if (rc != GM_OK) return done(rc);
                    GM_HIP(launch_resync(n, x, y, z, rx, ry, rz, vx, vy, vz, dt, st));
                    if ((rc = rebuild_test()) != GM_OK) return done(rc);
                    if (words[1] != kNone) return done(nan_in_pos(step, words[1]));
                    rebuild = words[0];
                } else {
                    if (d->h_ctl[9] != kNone) return done(nan_in_pos(step, d->h_ctl[9]));
                    rebuild = d->h_ctl[8] ? 1 : 0;
                }
                stepped = true;
            }
        }
        if (!stepped) {
            GM_HIP(launch_half_kick(n, vx, vy, vz, fx, fy, fz, d->mass.p, dt / 2.0, st));
            if ((rc = drift_and_constrain(s, x, y, z, vx, vy, vz, dt, md->max_iter, md->tol, step, nullptr)) != GM_OK)
                return done(rc);
            if ((rc = rebuild_test()) != GM_OK) return done(rc);
            if (words[1] != kNone) return done(nan_in_pos(step, words[1]));
            rebuild = words[0];
        }
        if (rebuild && (rc = gm_build_neighbor_list_device(s, x, y, z)) != GM_OK) return done(rc);
        if ((rc = compute_on_device(s, x, y, z, fx, fy, fz)) != GM_OK) return done(rc);
        const bool frame = (step + 1) % md->frame_interval == 0, last = k + 1 == md->n_steps;
        // forces checked before the velocity projection can fail: at once where
        // the host waits anyway, otherwise with the next step's read-back
        if (frame || last || !deferred) {
            if ((rc = collect(s, &e)) != GM_OK) return done(rc);
            if (s->last_bad != kNone) return done(nan_in_force(step, s->last_bad));
        }
        GM_HIP(launch_half_kick(n, vx, vy, vz, fx, fy, fz, d->mass.p, dt / 2.0, st));
        if (deferred) {
            rc = group_launch(s, true, Vec3{vx, vy, vz}, Vec3C{x, y, z}, md->max_iter, md->tol, nullptr);
            if (rc != GM_OK) return done(rc);
            v_pending = true;
            if (frame || last) {  // checked before a frame and at the end
                if ((rc = read_ctl(s)) != GM_OK) return done(rc);
                rc = group_finish(s, true, Vec3{vx, vy, vz}, Vec3C{x, y, z}, md->max_iter, md->tol, step, nullptr);
                if (rc != GM_OK) return done(rc);
                v_pending = false;
            }
        } else if ((rc = project(s, true, Vec3{vx, vy, vz}, Vec3C{x, y, z}, md->max_iter, md->tol, step,
                                 nullptr)) != GM_OK) {
            return done(rc);
        }
        if (frame && (rc = record(e.total)) != GM_OK) return done(rc);
    }
    GM_HIP(hipStreamSynchronize(s->stream));
    d->forces_valid = true;
    return done(GM_OK);
}

gm_status gm_md_last_failure(gm_sys *s, gm_md_failure *out) {
    GM_REQUIRE(s && s->dyn && out, "gm_md_last_failure: no dynamics attached or NULL argument");
    *out = s->dyn->fail;
    return GM_OK;
}

gm_status gm_jacobi_factor(gm_sys *s, double total_energy, double *out) {
    GM_REQUIRE(s && out, "gm_jacobi_factor: NULL argument");
    if (!s->built) return set_error(GM_ERR_NOT_BUILT, "gm_jacobi_factor before any force evaluation");
    gm_energy_terms e{};
    gm_status rc = collect(s, &e);
    if (rc != GM_OK) return rc;
    *out = 2.0 * (total_energy - e.total);
    return GM_OK;
}

}  // extern "C"
