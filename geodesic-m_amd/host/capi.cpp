// extern "C" layer of libgeodesic_host.so (include/geodesic_host.h) over the
// C++ host API (geodesic.hpp).  Exceptions become status codes with the
// message kept for gmh_last_error().
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>

#include "geodesic.hpp"
#include "geodesic_host.h"

using namespace geodesic;

struct gmh_topology {
    AtomData atoms;
    BondedTopology topo;
    std::vector<uint8_t> element;
    std::vector<uint32_t> residue;
};

struct gmh_backend {
    std::unique_ptr<GpuBackend> backend;
    SimState state;
    SimParams params;
};

namespace {

thread_local std::string g_error;

gm_status fail(gm_status st, const std::string &msg) {
    g_error = msg;
    return st;
}

template <typename F>
gm_status guarded(F &&f) {
    try {
        f();
        g_error.clear();
        return GM_OK;
    } catch (const BackendError &e) {
        return fail(e.status, e.what());
    } catch (const ConvergenceError &e) {
        return fail(GM_ERR_CONVERGENCE, e.what());
    } catch (const IoError &e) {
        return fail(GMH_ERR_IO, e.what());
    } catch (const ConfigError &e) {
        return fail(GMH_ERR_CONFIG, e.what());
    } catch (const NumericalError &e) {
        return fail(GMH_ERR_NUMERICAL, e.what());
    } catch (const std::bad_alloc &) {
        return fail(GM_ERR_OOM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(GM_ERR_INVALID_ARG, e.what());
    }
}

#define GMH_REQUIRE(cond, msg)                                   \
    do {                                                         \
        if (!(cond)) return fail(GM_ERR_INVALID_ARG, (msg));     \
    } while (0)

void blocks_in(const double *src, size_t n, std::vector<double> &x, std::vector<double> &y, std::vector<double> &z) {
    x.assign(src, src + n);
    y.assign(src + n, src + 2 * n);
    z.assign(src + 2 * n, src + 3 * n);
}

void blocks_out(double *dst, const std::vector<double> &x, const std::vector<double> &y, const std::vector<double> &z) {
    const size_t n = x.size();
    std::memcpy(dst, x.data(), n * sizeof(double));
    std::memcpy(dst + n, y.data(), n * sizeof(double));
    std::memcpy(dst + 2 * n, z.data(), n * sizeof(double));
}

template <typename T>
void expose(const std::vector<T> &v, const void **data, int64_t *count) {
    *data = v.empty() ? nullptr : v.data();
    *count = (int64_t)v.size();
}

}  // namespace

extern "C" {

const char *gmh_last_error(void) { return g_error.c_str(); }

gm_status gmh_prmtop_read(const char *path, int32_t fidelity, gmh_topology **out) {
    GMH_REQUIRE(path && out, "gmh_prmtop_read: NULL argument");
    *out = nullptr;
    return guarded([&] {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw IoError(path, "cannot open file");
        std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        PrmtopOptions opt;
        opt.fidelity = fidelity != 0;
        auto parsed = prmtop::parse(text, opt);
        auto *t = new gmh_topology();
        t->atoms = std::move(parsed.first);
        t->topo = std::move(parsed.second);
        for (const AtomMeta &m : t->atoms.meta) {
            t->element.push_back((uint8_t)m.element);
            t->residue.push_back(m.residue_id);
        }
        *out = t;
    });
}

gm_status gmh_topology_free(gmh_topology *t) {
    delete t;
    return GM_OK;
}

gm_status gmh_topology_field(const gmh_topology *t, int32_t field, const void **data, int64_t *count) {
    GMH_REQUIRE(t && data && count, "gmh_topology_field: NULL argument");
    const AtomData &a = t->atoms;
    const BondedTopology &b = t->topo;
    switch (field) {
        case GMH_F_SIGMA: expose(a.sigma, data, count); break;
        case GMH_F_EPSILON: expose(a.epsilon, data, count); break;
        case GMH_F_MASS: expose(a.mass, data, count); break;
        case GMH_F_CHARGE: expose(a.charge, data, count); break;
        case GMH_F_ELEMENT: expose(t->element, data, count); break;
        case GMH_F_RESIDUE: expose(t->residue, data, count); break;
        case GMH_F_BOND_I: expose(b.bond_i, data, count); break;
        case GMH_F_BOND_J: expose(b.bond_j, data, count); break;
        case GMH_F_BOND_K: expose(b.bond_k, data, count); break;
        case GMH_F_BOND_R0: expose(b.bond_r0, data, count); break;
        case GMH_F_ANGLE_I: expose(b.angle_i, data, count); break;
        case GMH_F_ANGLE_J: expose(b.angle_j, data, count); break;
        case GMH_F_ANGLE_K: expose(b.angle_k, data, count); break;
        case GMH_F_ANGLE_KTH: expose(b.angle_kth, data, count); break;
        case GMH_F_ANGLE_TH0: expose(b.angle_th0, data, count); break;
        case GMH_F_DIHED_I: expose(b.dihed_i, data, count); break;
        case GMH_F_DIHED_J: expose(b.dihed_j, data, count); break;
        case GMH_F_DIHED_K: expose(b.dihed_k, data, count); break;
        case GMH_F_DIHED_L: expose(b.dihed_l, data, count); break;
        case GMH_F_DIHED_KPHI: expose(b.dihed_kphi, data, count); break;
        case GMH_F_DIHED_N: expose(b.dihed_n, data, count); break;
        case GMH_F_DIHED_DELTA: expose(b.dihed_delta, data, count); break;
        case GMH_F_CONSTR_I: expose(b.constr_i, data, count); break;
        case GMH_F_CONSTR_J: expose(b.constr_j, data, count); break;
        case GMH_F_CONSTR_DSQ: expose(b.constr_dsq, data, count); break;
        case GMH_F_EXCL_I: expose(b.excl_i, data, count); break;
        case GMH_F_EXCL_J: expose(b.excl_j, data, count); break;
        case GMH_F_PAIR14_I: expose(b.pair14_i, data, count); break;
        case GMH_F_PAIR14_J: expose(b.pair14_j, data, count); break;
        default: return fail(GM_ERR_INVALID_ARG, "gmh_topology_field: unknown field");
    }
    return GM_OK;
}

gm_status gmh_topology_scales(const gmh_topology *t, double *scale_14, double *scee) {
    GMH_REQUIRE(t, "gmh_topology_scales: NULL topology");
    if (scale_14) *scale_14 = t->topo.pair14_i.empty() ? 0.0 : t->topo.scale_14;
    if (scee) *scee = t->topo.scee;
    return GM_OK;
}

gm_status gmh_topology_promote_hydrogen_bonds(gmh_topology *t) {
    GMH_REQUIRE(t, "gmh_topology_promote_hydrogen_bonds: NULL topology");
    return guarded([&] { promote_hydrogen_bonds(t->topo, t->atoms); });
}

gm_status gmh_inpcrd_read(const char *path, int64_t n_atoms, int32_t has_box, double *pos, double *vel) {
    GMH_REQUIRE(path && n_atoms >= 0 && (n_atoms == 0 || (pos && vel)), "gmh_inpcrd_read: bad argument");
    return guarded([&] {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw IoError(path, "cannot open file");
        std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        const SimState s = inpcrd::parse(text, (size_t)n_atoms, has_box != 0);
        if (n_atoms > 0) {
            blocks_out(pos, s.pos_x, s.pos_y, s.pos_z);
            blocks_out(vel, s.vel_x, s.vel_y, s.vel_z);
        }
    });
}

gm_status gmh_dcd_write(const char *path, int64_t natom, int32_t frame_interval, double dt_ps, int64_t n_frames,
                        const float *frames) {
    GMH_REQUIRE(path && natom >= 0 && n_frames >= 0 && frame_interval >= 0 && (n_frames == 0 || frames),
                "gmh_dcd_write: bad argument");
    return guarded([&] {
        DcdWriter w(path, (size_t)natom, (uint32_t)frame_interval, dt_ps);
        for (int64_t k = 0; k < n_frames; k++) {
            const float *fr = frames + (size_t)k * 3 * (size_t)natom;
            w.write_frame(fr, fr + natom, fr + 2 * natom, (size_t)natom);
        }
        w.close();
    });
}

gm_status gmh_energy_log_write(const char *path, int64_t n_atoms, int64_t n_rows, const uint64_t *step,
                               const double *time_ps, const double *potential_kcal, const double *kinetic_kcal) {
    GMH_REQUIRE(path && n_atoms >= 0 && n_rows >= 0 && (n_rows == 0 || (step && time_ps && potential_kcal && kinetic_kcal)),
                "gmh_energy_log_write: bad argument");
    return guarded([&] {
        EnergyLogWriter w(path, (size_t)n_atoms);
        for (int64_t k = 0; k < n_rows; k++) w.write_row(step[k], time_ps[k], potential_kcal[k], kinetic_kcal[k]);
    });
}

gm_status gmh_energy_from_files(const char *prmtop_path, const char *inpcrd_path, int32_t fidelity, int32_t device,
                                gmh_energy_report *out) {
    GMH_REQUIRE(prmtop_path && inpcrd_path && out, "gmh_energy_from_files: NULL argument");
    return guarded([&] {
        PrmtopOptions opt;
        opt.fidelity = fidelity != 0;
        const EnergyReport r = energy_from_files(prmtop_path, inpcrd_path, opt, device);
        out->bond = r.bond;
        out->angle = r.angle;
        out->dihedral = r.dihedral;
        out->nonbonded = r.nonbonded;
        out->total = r.total;
        out->lj14 = r.lj14;
        out->r_switch = r.r_switch;
        out->r_cutoff = r.r_cutoff;
    });
}

gm_status gmh_run(const gmh_run_config *cfg, gmh_run_summary *out) {
    GMH_REQUIRE(cfg && out && cfg->prmtop && cfg->inpcrd && cfg->trajectory && cfg->energy_log && cfg->backend,
                "gmh_run: NULL argument");
    GMH_REQUIRE(cfg->n_steps >= 0 && cfg->frame_interval >= 1 && cfg->max_constr_iter >= 1, "gmh_run: bad argument");
    return guarded([&] {
        RunConfig c;
        c.n_steps = (uint64_t)cfg->n_steps;
        c.frame_interval = (uint32_t)cfg->frame_interval;
        c.backend = cfg->backend;
        c.n_threads = cfg->n_threads > 0 ? (size_t)cfg->n_threads : 1;
        c.prmtop = cfg->prmtop;
        c.inpcrd = cfg->inpcrd;
        for (int k = 0; k < 3; k++) c.box_size[(size_t)k] = cfg->box_size[k];
        c.periodic = cfg->periodic != 0;
        c.dt = cfg->dt;
        c.total_energy = cfg->total_energy;
        c.r_cutoff = cfg->r_cutoff;
        c.r_skin = cfg->r_skin;
        c.r_switch = cfg->r_switch;
        c.max_constr_iter = (uint32_t)cfg->max_constr_iter;
        c.constr_tol = cfg->constr_tol;
        c.constrain_hydrogens = cfg->constrain_hydrogens != 0;
        c.trajectory = cfg->trajectory;
        c.energy_log = cfg->energy_log;
        c.monitoring = cfg->monitoring != 0;
        c.energy_drift_threshold_kcal = cfg->energy_drift_threshold_kcal;
        c.energy_drift_stop = cfg->energy_drift_stop != 0;
        c.prmtop_options.fidelity = cfg->fidelity != 0;
        c.device = cfg->device;
        const RunSummary s = run(c);
        out->n_steps = (int64_t)s.n_steps;
        out->n_frames = (int64_t)s.n_frames;
        out->final_potential = s.final_potential;
        out->final_kinetic = s.final_kinetic;
    });
}

gm_status gmh_backend_create(const gmh_topology *t, const double box[3], int32_t periodic, double r_cutoff,
                             double r_skin, double r_switch, int32_t max_constr_iter, double constr_tol,
                             int32_t bonded_min_image, int32_t device, gmh_backend **out) {
    GMH_REQUIRE(t && box && out && max_constr_iter >= 1, "gmh_backend_create: bad argument");
    *out = nullptr;
    return guarded([&] {
        auto *b = new gmh_backend();
        const size_t n = t->atoms.mass.size();
        b->params.n_atoms = n;
        b->params.box_size = {{box[0], box[1], box[2]}};
        b->params.periodic = periodic != 0;
        b->params.r_cutoff = r_cutoff;
        b->params.r_skin = r_skin;
        b->params.r_switch = r_switch;
        b->params.max_constr_iter = (uint32_t)max_constr_iter;
        b->params.constr_tol = constr_tol;
        b->params.bonded_min_image = bonded_min_image != 0;
        b->state = SimState(n);
        try {
            b->backend = GpuBackend::try_new(t->atoms, t->topo, b->params, device);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

gm_status gmh_backend_free(gmh_backend *b) {
    delete b;
    return GM_OK;
}

gm_status gmh_backend_set_state(gmh_backend *b, const double *pos, const double *vel, uint64_t step) {
    GMH_REQUIRE(b && pos && vel, "gmh_backend_set_state: NULL argument");
    const size_t n = b->state.pos_x.size();
    blocks_in(pos, n, b->state.pos_x, b->state.pos_y, b->state.pos_z);
    blocks_in(vel, n, b->state.vel_x, b->state.vel_y, b->state.vel_z);
    b->state.step = step;
    return GM_OK;
}

gm_status gmh_backend_get_state(const gmh_backend *b, double *pos, double *vel) {
    GMH_REQUIRE(b && pos && vel, "gmh_backend_get_state: NULL argument");
    blocks_out(pos, b->state.pos_x, b->state.pos_y, b->state.pos_z);
    blocks_out(vel, b->state.vel_x, b->state.vel_y, b->state.vel_z);
    return GM_OK;
}

gm_status gmh_backend_build_neighbor_list(gmh_backend *b) {
    GMH_REQUIRE(b, "NULL backend");
    return guarded([&] { b->backend->build_neighbor_list(b->state, b->params); });
}

gm_status gmh_backend_compute_forces(gmh_backend *b, double *forces, gm_energy_terms *e) {
    GMH_REQUIRE(b, "NULL backend");
    return guarded([&] {
        const ForceBuffer &f = b->backend->compute_forces(b->state);
        b->state.force_x = f.fx;
        b->state.force_y = f.fy;
        b->state.force_z = f.fz;
        b->state.potential_energy = b->backend->potential_energy();
        if (forces) blocks_out(forces, f.fx, f.fy, f.fz);
        if (e) *e = b->backend->energy_terms();
    });
}

gm_status gmh_backend_needs_rebuild(const gmh_backend *b, int32_t *out) {
    GMH_REQUIRE(b && out, "NULL argument");
    return guarded([&] { *out = b->backend->needs_rebuild(b->state) ? 1 : 0; });
}

gm_status gmh_backend_geodesic_drift(gmh_backend *b, double dt) {
    GMH_REQUIRE(b, "NULL backend");
    return guarded([&] { b->backend->geodesic_drift(b->state, dt); });
}

gm_status gmh_backend_n_threads(const gmh_backend *b, int64_t *out) {
    GMH_REQUIRE(b && out, "NULL argument");
    *out = (int64_t)b->backend->n_threads();
    return GM_OK;
}

}  // extern "C"
