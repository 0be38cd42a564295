// Host side of the MI355X force engine in C++: the reference's core types
// and `ComputeBackend` trait, a `GpuBackend` over the C-ABI of
// libgeodesic_hip.so, the AMBER prmtop/inpcrd readers, the DCD and CSV
// writers, and the `energy` and `run` entry points.
//
// Mirrors, with the same names, argument meaning and error behaviour
// (paths relative to the reference repository root):
//   geodesic-core/src/{atoms,topology,state,params,buffers,backend,error}.rs
//   geodesic-gpu/src/gpu_backend.rs:11-100          (GpuBackend)
//   geodesic-io/src/prmtop.rs:182-344, inpcrd.rs:12-97
//   geodesic-io/src/dcd.rs, export.rs:20-55          (DcdWriter, EnergyLogWriter)
//   geodesic-engine/src/constraint.rs:10-40          (promote_hydrogen_bonds)
//   geodesic/src/lib.rs:113-171, 187-330             (energy_from_files, run)
// Rust's Result<T, E> becomes T-or-throw with one exception type per error
// enum; the C shim (capi.cpp) turns them into status codes.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "geodesic_hip.h"

namespace geodesic {

// ---- errors (error.rs) -----------------------------------------------------
struct IoError : std::runtime_error {
    std::string path;
    IoError(const std::string &p, const std::string &what) : std::runtime_error(p + ": " + what), path(p) {}
};
struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NumericalError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ConvergenceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// BackendError (error.rs:82-91); `status` is the gm_status of the failure.
struct BackendError : std::runtime_error {
    int status;
    BackendError(int st, const std::string &what) : std::runtime_error(what), status(st) {}
};

// ---- core types ------------------------------------------------------------
enum class Element : uint8_t { H, C, N, O, S, Unknown };

struct AtomMeta {
    Element element = Element::Unknown;
    uint32_t residue_id = 0;
    std::array<char, 4> atom_name{{' ', ' ', ' ', ' '}};
    uint8_t chain_id = 0;
};

// atoms.rs:3-12
struct AtomData {
    std::vector<double> epsilon;  // kcal/mol
    std::vector<double> sigma;    // A
    std::vector<double> mass;     // amu
    std::vector<double> charge;   // elementary charge (not used by the force field)
    std::vector<AtomMeta> meta;
};

// topology.rs:4-37, plus the scaled 1-4 pairs read in fidelity mode [ext]
struct BondedTopology {
    std::vector<uint32_t> bond_i, bond_j;
    std::vector<double> bond_k, bond_r0;
    std::vector<uint32_t> angle_i, angle_j, angle_k;
    std::vector<double> angle_kth, angle_th0;
    std::vector<uint32_t> dihed_i, dihed_j, dihed_k, dihed_l;
    std::vector<double> dihed_kphi;
    std::vector<uint32_t> dihed_n;
    std::vector<double> dihed_delta;
    std::vector<uint32_t> constr_i, constr_j;
    std::vector<double> constr_dsq;
    std::vector<uint32_t> excl_i, excl_j;
    std::vector<uint32_t> pair14_i, pair14_j;  // [ext]
    double scale_14 = 0.0;                     // [ext] LJ scale of pair14_*, 1 / SCNB
    double scee = 0.0;                         // [ext] SCEE as read (no electrostatics here)
};

// state.rs:3-23
struct SimState {
    std::vector<double> pos_x, pos_y, pos_z;
    std::vector<double> vel_x, vel_y, vel_z;
    std::vector<double> force_x, force_y, force_z;
    double potential_energy = 0.0;
    double kinetic_energy = 0.0;
    uint64_t step = 0;
    explicit SimState(size_t n_atoms = 0)
        : pos_x(n_atoms), pos_y(n_atoms), pos_z(n_atoms), vel_x(n_atoms), vel_y(n_atoms), vel_z(n_atoms),
          force_x(n_atoms), force_y(n_atoms), force_z(n_atoms) {}
};

// params.rs:4-23, plus [ext] fields of the C-ABI descriptor
struct SimParams {
    size_t n_atoms = 0;
    uint64_t n_steps = 0;
    double dt = 0.0;
    std::array<double, 3> box_size{{0.0, 0.0, 0.0}};
    bool periodic = false;
    double r_cutoff = 0.0, r_skin = 0.0, r_switch = 0.0;
    uint32_t max_constr_iter = 100;
    double constr_tol = 1e-10;
    uint32_t frame_interval = 1;
    size_t n_threads = 1;
    double total_energy = 0.0;
    bool bonded_min_image = false;  // [ext]
};

// buffers.rs:5-9
struct ForceBuffer {
    std::vector<double> fx, fy, fz;
};

// backend.rs:11-23
class ComputeBackend {
public:
    virtual ~ComputeBackend() = default;
    virtual void build_neighbor_list(SimState &state, const SimParams &params) = 0;
    virtual const ForceBuffer &compute_forces(const SimState &state) = 0;
    virtual void geodesic_drift(SimState &state, double dt) = 0;  // throws ConvergenceError
    virtual ForceBuffer reduce_forces() const = 0;
    virtual double potential_energy() const = 0;
    virtual const AtomData &atoms() const = 0;
    virtual const BondedTopology &topology() const = 0;
    virtual bool needs_rebuild(const SimState &state) const = 0;
    virtual size_t n_threads() const = 0;
};

// Frames and energies of a device-resident run segment (GpuBackend::run_md),
// and how the segment ended: status GM_OK, or GM_ERR_NUMERICAL /
// GM_ERR_CONVERGENCE with the reference's message (the frames taken before
// the failure are kept).
struct MdRecord {
    int64_t n_frames = 0;
    std::vector<float> frames;  // n_frames x (x block, y block, z block) x n_atoms
    std::vector<double> epot, ekin;
    gm_status status = GM_OK;
    std::string error;
};

// gpu_backend.rs:11-100 over libgeodesic_hip.so.  Owns one device context,
// one gm_sys and the device arrays the dynamics kernels work on.
class GpuBackend : public ComputeBackend {
public:
    // GpuBackend::try_new (gpu_backend.rs:24-28): takes atoms and topology by
    // value; throws BackendError (GM_ERR_NO_DEVICE, GM_ERR_OOM, ...).
    static std::unique_ptr<GpuBackend> try_new(AtomData atoms, BondedTopology topology, const SimParams &params,
                                               int device = 0);
    ~GpuBackend() override;
    GpuBackend(const GpuBackend &) = delete;
    GpuBackend &operator=(const GpuBackend &) = delete;

    void build_neighbor_list(SimState &state, const SimParams &params) override;
    const ForceBuffer &compute_forces(const SimState &state) override;
    // drift + SHAKE (integrator.rs:40-71) on the device, state round-tripped
    void geodesic_drift(SimState &state, double dt) override;
    ForceBuffer reduce_forces() const override { return forces_; }
    double potential_energy() const override { return terms_.total; }
    const AtomData &atoms() const override { return atoms_; }
    const BondedTopology &topology() const override { return topology_; }
    bool needs_rebuild(const SimState &state) const override;
    size_t n_threads() const override { return 1; }

    // Inherent: per-term energies of the last evaluation (EnergyReport fields).
    const gm_energy_terms &energy_terms() const { return terms_; }
    // Inherent: the run loop of lib.rs:255-330 on the device (gm_md_run),
    // steps [first_step, first_step + n_steps); resume continues the previous
    // segment (which must end on a frame step); positions and velocities
    // advance in `state`.  A NaN or convergence failure ends the segment
    // early and is returned in the record (the caller writes its frames, then
    // raises it); other failures throw.
    MdRecord run_md(SimState &state, uint64_t first_step, uint64_t n_steps, double dt, uint32_t frame_interval,
                    bool resume);
    gm_sys *raw() const { return sys_; }

private:
    GpuBackend() = default;
    void upload_state(const SimState &state);
    void download_state(SimState &state);

    AtomData atoms_;
    BondedTopology topology_;
    SimParams params_;
    ForceBuffer forces_;
    gm_energy_terms terms_{};
    gm_ctx *ctx_ = nullptr;
    gm_sys *sys_ = nullptr;
    double *d_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // x y z vx vy vz
};

// ---- geodesic-io -----------------------------------------------------------
struct PrmtopOptions {
    // false: the reference's reading (prmtop.rs:132-180: 1-4 and improper
    // flags dropped, 1-4 pairs only excluded).  true [ext]: proper dihedrals
    // whose 1-4 flag is clear contribute a scaled 1-4 LJ pair, scale 1/SCNB
    // (SCNB_SCALE_FACTOR, default 2.0 when the section is absent).
    bool fidelity = false;
};
namespace prmtop {
std::pair<AtomData, BondedTopology> parse(const std::string &text, const PrmtopOptions &opt = {});
}
namespace inpcrd {
// inpcrd.rs:12-84; has_box disambiguates a trailing box line from velocities.
SimState parse(const std::string &text, size_t expected_n_atoms, bool has_box);
}
// constraint.rs:10-40: bonds touching a hydrogen become constraints.
void promote_hydrogen_bonds(BondedTopology &topology, const AtomData &atoms);

// dcd.rs: CHARMM/NAMD DCD, f32 x/y/z, no unit cell; NSET patched on close.
class DcdWriter {
public:
    DcdWriter(const std::string &path, size_t natom, uint32_t frame_interval, double dt_ps);
    ~DcdWriter();
    void write_frame(const float *x, const float *y, const float *z, size_t n);
    void close();

private:
    std::string path_;
    void *file_ = nullptr;
    size_t natom_ = 0;
    uint32_t n_frames_ = 0;
    bool closed_ = false;
};

// export.rs:20-55: one CSV row per call, flushed.
class EnergyLogWriter {
public:
    EnergyLogWriter(const std::string &path, size_t n_atoms);
    ~EnergyLogWriter();
    void write_row(uint64_t step, double time_ps, double potential_kcal, double kinetic_kcal);

private:
    std::string path_;
    void *file_ = nullptr;
    size_t n_atoms_ = 0;
};

// ---- geodesic (lib.rs) -----------------------------------------------------
struct EnergyReport {
    double bond = 0, angle = 0, dihedral = 0, nonbonded = 0, total = 0;
    double lj14 = 0;  // [ext] scaled 1-4 LJ (fidelity mode; 0 otherwise)
    double r_switch = 0, r_cutoff = 0;
};
// lib.rs:113-171 through the GPU backend: standard cutoffs 12 / 14 / 10 A,
// 1000 A non-periodic box, no constraint promotion.
EnergyReport energy_from_files(const std::string &prmtop_path, const std::string &inpcrd_path,
                               const PrmtopOptions &opt = {}, int device = 0);

// config.rs:102-128 (file paths already resolved by the caller).
struct RunConfig {
    uint64_t n_steps = 0;
    uint32_t frame_interval = 1;
    std::string backend = "gpu";  // "gpu" or "hybrid" (the same backend, as in lib.rs:240)
    size_t n_threads = 1;
    std::string prmtop, inpcrd;
    std::array<double, 3> box_size{{1000.0, 1000.0, 1000.0}};
    bool periodic = false;
    double dt = 0.001, total_energy = 0.0;
    double r_cutoff = 12.0, r_skin = 14.0, r_switch = 10.0;
    uint32_t max_constr_iter = 100;
    double constr_tol = 1e-10;
    bool constrain_hydrogens = true;
    std::string trajectory, energy_log;
    bool monitoring = false;           // MonitoringConfig present
    double energy_drift_threshold_kcal = 0.0;
    bool energy_drift_stop = false;    // DriftAction::Stop (else Warn)
    PrmtopOptions prmtop_options;      // [ext]
    int device = 0;
};
struct RunSummary {
    uint64_t n_steps = 0;
    size_t n_frames = 0;
    double final_potential = 0.0, final_kinetic = 0.0;
    std::string trajectory, energy_log;
};
// lib.rs:187-330: the BAB + RATTLE loop on the GPU (device resident), frames
// to the DCD, energies to the CSV log, energy-drift monitoring per frame.
RunSummary run(const RunConfig &config);

}  // namespace geodesic
