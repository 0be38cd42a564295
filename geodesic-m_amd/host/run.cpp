// The `energy` and `run` entry points of geodesic/src/lib.rs (paths relative
// to the reference repository root) over the GPU backend.
//
// energy_from_files (lib.rs:113-171): prmtop + inpcrd at the standard
// cutoffs (r_switch 10, r_cutoff 12, r_skin 14 A) in a 1000 A non-periodic
// box, no constraint promotion; per-term energies from one GPU evaluation
// instead of the reference's CPU-only routines.
//
// run (lib.rs:187-330): parse, promote hydrogen bonds to constraints when
// configured, then the BAB + RATTLE loop (initial list, forces and velocity
// projection; per step half-kick, drift + SHAKE, rebuild when needs_rebuild
// fires, forces, half-kick, RATTLE) device-resident in one gm_md_run call,
// with f32 frames and the potential / kinetic energy of every frame-interval
// step returned to the host.  Frames go to the DCD and rows to the CSV log in
// step order; energy-drift monitoring (lib.rs:365-386) is applied per frame
// in the same order, so a Stop leaves the same files the reference would.
// Non-finite positions and forces are checked on the device at every step
// (lib.rs:79-107) and stop the run there, after the frames before it are
// written.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "geodesic.hpp"

namespace geodesic {
namespace {

std::string read_file(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw IoError(path, "cannot open file");
    std::ostringstream s;
    s << f.rdbuf();
    if (!f.good() && !f.eof()) throw IoError(path, "read failed");
    return s.str();
}

// lib.rs:30-33
constexpr double kEnergyRSwitch = 10.0, kEnergyRCutoff = 12.0, kEnergyRSkin = 14.0, kEnergyBox = 1000.0;

}  // namespace

EnergyReport energy_from_files(const std::string &prmtop_path, const std::string &inpcrd_path,
                               const PrmtopOptions &opt, int device) {
    auto parsed = prmtop::parse(read_file(prmtop_path), opt);
    const size_t n = parsed.first.mass.size();
    SimState state = inpcrd::parse(read_file(inpcrd_path), n, false);
    SimParams p;
    p.n_atoms = n;
    p.box_size = {{kEnergyBox, kEnergyBox, kEnergyBox}};
    p.periodic = false;
    p.r_cutoff = kEnergyRCutoff;
    p.r_skin = kEnergyRSkin;
    p.r_switch = kEnergyRSwitch;
    p.max_constr_iter = 100;
    p.constr_tol = 1e-10;
    auto b = GpuBackend::try_new(std::move(parsed.first), std::move(parsed.second), p, device);
    b->build_neighbor_list(state, p);
    b->compute_forces(state);
    const gm_energy_terms &e = b->energy_terms();
    EnergyReport r;
    r.bond = e.bond;
    r.angle = e.angle;
    r.dihedral = e.dihedral;
    r.nonbonded = e.lj;
    r.lj14 = e.lj14;
    r.total = e.total;  // bond + angle + dihedral + lj14 + nonbonded, in that order
    r.r_switch = kEnergyRSwitch;
    r.r_cutoff = kEnergyRCutoff;
    return r;
}

RunSummary run(const RunConfig &c) {
    if (c.backend != "gpu" && c.backend != "hybrid")
        throw ConfigError("invalid value for key 'run.backend': '" + c.backend +
                          "' (this library provides the GPU backend: \"gpu\" or \"hybrid\")");
    if (c.frame_interval == 0) throw ConfigError("invalid value for key 'run.frame_interval': '0' (must be >= 1)");
    auto parsed = prmtop::parse(read_file(c.prmtop), c.prmtop_options);
    AtomData &atoms = parsed.first;
    BondedTopology &topology = parsed.second;
    const size_t n = atoms.mass.size();
    SimState state = inpcrd::parse(read_file(c.inpcrd), n, false);
    if (c.constrain_hydrogens) promote_hydrogen_bonds(topology, atoms);

    SimParams p;
    p.n_atoms = n;
    p.n_steps = c.n_steps;
    p.dt = c.dt;
    p.box_size = c.box_size;
    p.periodic = c.periodic;
    p.r_cutoff = c.r_cutoff;
    p.r_skin = c.r_skin;
    p.r_switch = c.r_switch;
    p.max_constr_iter = c.max_constr_iter;
    p.constr_tol = c.constr_tol;
    p.frame_interval = c.frame_interval;
    p.n_threads = c.n_threads;
    p.total_energy = c.total_energy;
    // prmtop fidelity: periodic systems take minimum-image bonded displacements
    p.bonded_min_image = c.prmtop_options.fidelity && c.periodic;

    auto backend = GpuBackend::try_new(std::move(atoms), std::move(topology), p, c.device);
    DcdWriter dcd(c.trajectory, n, c.frame_interval, c.dt);
    EnergyLogWriter csv(c.energy_log, n);

    // The device loop runs in segments of whole frame intervals; each
    // segment's frames are written before the next starts, so host memory is
    // bounded and a failure leaves every frame taken before it on disk.
    const uint64_t frame_bytes = 12 * (uint64_t)(n > 0 ? n : 1);
    const uint64_t seg_frames = std::max<uint64_t>(1, std::min<uint64_t>(4096, (256ull << 20) / frame_bytes));
    const uint64_t seg_steps = seg_frames * c.frame_interval;
    uint64_t done = 0, frame_no = 0;
    double e0 = 0.0;
    bool first = true;
    while (first || done < c.n_steps) {
        const uint64_t ns = std::min<uint64_t>(seg_steps, c.n_steps - done);
        MdRecord rec = backend->run_md(state, done, ns, c.dt, c.frame_interval, !first);
        for (int64_t k = 0; k < rec.n_frames; k++, frame_no++) {
            const float *fr = rec.frames.data() + (size_t)k * 3 * n;
            const uint64_t step = frame_no * c.frame_interval;
            dcd.write_frame(fr, fr + n, fr + 2 * n, n);
            csv.write_row(step, (double)step * c.dt, rec.epot[(size_t)k], rec.ekin[(size_t)k]);
            if (frame_no == 0) {
                e0 = rec.epot[(size_t)k] + rec.ekin[(size_t)k];
            } else if (c.monitoring) {
                const double drift = std::fabs(rec.epot[(size_t)k] + rec.ekin[(size_t)k] - e0);
                if (drift > c.energy_drift_threshold_kcal) {
                    std::ostringstream m;
                    if (c.energy_drift_stop) {
                        dcd.close();
                        m << "energy drift " << drift << " kcal/mol exceeds threshold "
                          << c.energy_drift_threshold_kcal << " kcal/mol at step " << step;
                        throw NumericalError(m.str());
                    }
                    std::fprintf(stderr, "warning: energy drift %.4f kcal/mol exceeds threshold %.4f at step %llu\n",
                                 drift, c.energy_drift_threshold_kcal, (unsigned long long)step);
                }
            }
        }
        if (rec.status != GM_OK) {
            dcd.close();
            if (rec.status == GM_ERR_CONVERGENCE) throw ConvergenceError(rec.error);
            throw NumericalError(rec.error);
        }
        done += ns;
        first = false;
    }
    dcd.close();
    RunSummary s;
    s.n_steps = c.n_steps;
    s.n_frames = (size_t)frame_no;
    s.final_potential = state.potential_energy;
    s.final_kinetic = state.kinetic_energy;
    s.trajectory = c.trajectory;
    s.energy_log = c.energy_log;
    return s;
}

}  // namespace geodesic
