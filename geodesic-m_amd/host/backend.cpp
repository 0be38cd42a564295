// GpuBackend: the ComputeBackend trait (geodesic-core/src/backend.rs:11-23;
// paths relative to the reference repository root) over libgeodesic_hip.so,
// replacing geodesic-gpu/src/gpu_backend.rs:11-100.
//
// try_new copies AtomData / BondedTopology into a gm_system_desc (the library
// copies every array), attaches masses and constraints for the device
// dynamics, and allocates six device arrays (positions, velocities) that
// geodesic_drift and run_md work on.  Evaluation is synchronous with host
// SoA buffers, as the trait is; run_md keeps everything on the device.
// The simulation parameters are the ones given to try_new: the reference's
// callers pass the same SimParams to every build_neighbor_list.
#include <cstring>
#include <sstream>

#include "geodesic.hpp"

namespace geodesic {
namespace {

void check(gm_status st, const char *what) {
    if (st == GM_OK) return;
    std::string msg = std::string(what) + ": " + gm_last_error();
    if (st == GM_ERR_CONVERGENCE) throw ConvergenceError(msg);
    throw BackendError(st, msg);
}

template <typename T>
const T *ptr_or_null(const std::vector<T> &v) {
    return v.empty() ? nullptr : v.data();
}

}  // namespace

std::unique_ptr<GpuBackend> GpuBackend::try_new(AtomData atoms, BondedTopology topology, const SimParams &params,
                                                int device) {
    const size_t n = atoms.sigma.size();
    if (atoms.epsilon.size() != n || atoms.mass.size() != n)
        throw BackendError(GM_ERR_INVALID_ARG, "AtomData arrays differ in length");
    std::unique_ptr<GpuBackend> b(new GpuBackend());
    b->atoms_ = std::move(atoms);
    b->topology_ = std::move(topology);
    b->params_ = params;
    const AtomData &a = b->atoms_;
    const BondedTopology &t = b->topology_;
    check(gm_init(device, &b->ctx_), "gm_init");

    gm_system_desc d;
    std::memset(&d, 0, sizeof(d));
    d.n_atoms = (int64_t)n;
    d.sigma = ptr_or_null(a.sigma);
    d.epsilon = ptr_or_null(a.epsilon);
    for (int k = 0; k < 3; k++) d.box[k] = params.box_size[(size_t)k];
    d.periodic = params.periodic ? 1 : 0;
    d.r_cutoff = params.r_cutoff;
    d.r_switch = params.r_switch;
    d.r_skin = params.r_skin;
    d.scale_14 = t.pair14_i.empty() ? 0.0 : t.scale_14;
    d.bonded_min_image = params.bonded_min_image ? 1 : 0;
    d.n_bonds = (int64_t)t.bond_i.size();
    d.bond_i = ptr_or_null(t.bond_i);
    d.bond_j = ptr_or_null(t.bond_j);
    d.bond_k = ptr_or_null(t.bond_k);
    d.bond_r0 = ptr_or_null(t.bond_r0);
    d.n_angles = (int64_t)t.angle_i.size();
    d.angle_i = ptr_or_null(t.angle_i);
    d.angle_j = ptr_or_null(t.angle_j);
    d.angle_k = ptr_or_null(t.angle_k);
    d.angle_kth = ptr_or_null(t.angle_kth);
    d.angle_th0 = ptr_or_null(t.angle_th0);
    d.n_dihedrals = (int64_t)t.dihed_i.size();
    d.dihed_i = ptr_or_null(t.dihed_i);
    d.dihed_j = ptr_or_null(t.dihed_j);
    d.dihed_k = ptr_or_null(t.dihed_k);
    d.dihed_l = ptr_or_null(t.dihed_l);
    d.dihed_kphi = ptr_or_null(t.dihed_kphi);
    d.dihed_n = ptr_or_null(t.dihed_n);
    d.dihed_delta = ptr_or_null(t.dihed_delta);
    d.n_excl = (int64_t)t.excl_i.size();
    d.excl_i = ptr_or_null(t.excl_i);
    d.excl_j = ptr_or_null(t.excl_j);
    d.n_pairs14 = (int64_t)t.pair14_i.size();
    d.pair14_i = ptr_or_null(t.pair14_i);
    d.pair14_j = ptr_or_null(t.pair14_j);
    check(gm_system_create(b->ctx_, &d, &b->sys_), "gm_system_create");
    check(gm_dynamics_set(b->sys_, ptr_or_null(a.mass), (int64_t)t.constr_i.size(), ptr_or_null(t.constr_i),
                          ptr_or_null(t.constr_j), ptr_or_null(t.constr_dsq)),
          "gm_dynamics_set");
    for (double *&p : b->d_) {
        void *q = nullptr;
        check(gm_device_alloc(b->ctx_, sizeof(double) * (n > 0 ? n : 1), &q), "gm_device_alloc");
        p = static_cast<double *>(q);
    }
    b->forces_.fx.assign(n, 0.0);
    b->forces_.fy.assign(n, 0.0);
    b->forces_.fz.assign(n, 0.0);
    return b;
}

GpuBackend::~GpuBackend() {
    for (double *p : d_)
        if (p) (void)gm_device_free(ctx_, p);
    if (sys_) (void)gm_system_destroy(sys_);
    if (ctx_) (void)gm_ctx_destroy(ctx_);
}

void GpuBackend::build_neighbor_list(SimState &state, const SimParams &) {
    check(gm_build_neighbor_list(sys_, state.pos_x.data(), state.pos_y.data(), state.pos_z.data()),
          "gm_build_neighbor_list");
}

// Infallible as the trait's compute_forces is: a non-finite force
// (GM_ERR_NUMERICAL) is returned as NaN in the buffer, for the caller's
// finiteness check (lib.rs:79-92) to report.
const ForceBuffer &GpuBackend::compute_forces(const SimState &state) {
    const gm_status st = gm_compute_forces(sys_, state.pos_x.data(), state.pos_y.data(), state.pos_z.data(),
                                           forces_.fx.data(), forces_.fy.data(), forces_.fz.data(), &terms_);
    if (st != GM_ERR_NUMERICAL) check(st, "gm_compute_forces");
    return forces_;
}

bool GpuBackend::needs_rebuild(const SimState &state) const {
    int32_t out = 0;
    check(gm_needs_rebuild(sys_, state.pos_x.data(), state.pos_y.data(), state.pos_z.data(), &out),
          "gm_needs_rebuild");
    return out != 0;
}

void GpuBackend::upload_state(const SimState &s) {
    const size_t bytes = sizeof(double) * s.pos_x.size();
    const std::vector<double> *src[6] = {&s.pos_x, &s.pos_y, &s.pos_z, &s.vel_x, &s.vel_y, &s.vel_z};
    for (int k = 0; k < 6; k++) check(gm_copy_to_device(ctx_, d_[k], src[k]->data(), bytes), "gm_copy_to_device");
}

void GpuBackend::download_state(SimState &s) {
    const size_t bytes = sizeof(double) * s.pos_x.size();
    std::vector<double> *dst[6] = {&s.pos_x, &s.pos_y, &s.pos_z, &s.vel_x, &s.vel_y, &s.vel_z};
    for (int k = 0; k < 6; k++) check(gm_copy_to_host(ctx_, dst[k]->data(), d_[k], bytes), "gm_copy_to_host");
}

// integrator.rs:40-71 (drift_and_constrain) with the backend's
// max_constr_iter / constr_tol, as CpuBackend::geodesic_drift does.
void GpuBackend::geodesic_drift(SimState &state, double dt) {
    upload_state(state);
    int32_t iters = 0;
    const gm_status st = gm_drift_and_constrain_device(sys_, d_[0], d_[1], d_[2], d_[3], d_[4], d_[5], dt,
                                                       (int32_t)params_.max_constr_iter, params_.constr_tol,
                                                       (int64_t)state.step, &iters);
    if (st == GM_ERR_CONVERGENCE) {
        std::ostringstream m;
        m << "constraint solver failed to converge at step " << state.step << " after "
          << params_.max_constr_iter << " iterations: " << gm_last_error();
        throw ConvergenceError(m.str());
    }
    check(st, "gm_drift_and_constrain_device");
    download_state(state);
}

MdRecord GpuBackend::run_md(SimState &state, uint64_t first_step, uint64_t n_steps, double dt,
                            uint32_t frame_interval, bool resume) {
    const size_t n = state.pos_x.size();
    if (frame_interval == 0) throw ConfigError("invalid value for key 'run.frame_interval': '0' (must be >= 1)");
    MdRecord r;
    const uint64_t cap = n_steps / frame_interval + 2;
    r.frames.resize((size_t)cap * 3 * n);
    r.epot.resize((size_t)cap);
    r.ekin.resize((size_t)cap);
    upload_state(state);
    gm_md_desc md;
    std::memset(&md, 0, sizeof(md));
    md.n_steps = (int64_t)n_steps;
    md.dt = dt;
    md.frame_interval = (int32_t)frame_interval;
    md.max_iter = (int32_t)params_.max_constr_iter;
    md.tol = params_.constr_tol;
    md.first_step = (int64_t)first_step;
    md.resume = resume ? 1 : 0;
    const gm_status st = gm_md_run(sys_, d_[0], d_[1], d_[2], d_[3], d_[4], d_[5], &md, r.frames.data(),
                                   r.epot.data(), r.ekin.data(), &r.n_frames);
    if (st == GM_ERR_NUMERICAL || st == GM_ERR_CONVERGENCE) {
        r.status = st;
        r.error = gm_last_error();
    } else {
        check(st, "gm_md_run");
    }
    download_state(state);
    state.step = first_step + n_steps;
    if (r.n_frames > 0) {
        state.potential_energy = r.epot[(size_t)r.n_frames - 1];
        state.kinetic_energy = r.ekin[(size_t)r.n_frames - 1];
    }
    return r;
}

}  // namespace geodesic
