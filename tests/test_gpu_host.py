"""The C++ host layer over the GPU backend (libgeodesic_host.so).

Paths relative to the reference repository root:
  - `geodesic energy` (geodesic/src/lib.rs:113-171) through the backend:
    per-term energies of ala_dipeptide within 1e-6 of the f64 oracle, in the
    reference's reading of the prmtop and with the 1-4 flags kept (fidelity);
  - `geodesic run` (lib.rs:187-330) with the golden config of
    geodesic/tests/golden_reference.rs:27-94: 100 frames, the DCD's header and
    layout byte-identical to the committed golden file and every coordinate
    within 1e-4 A of it (FP32 pair forces), and the CSV rows equal to the run's
    own energies at 3 decimals; energy-drift monitoring with action Stop ends
    the run after writing the frame that exceeded the threshold;
  - the ComputeBackend trait (backend.rs:11-23) of the C++ GpuBackend: forces
    byte-identical to the engine's host entry point, geodesic_drift byte-
    identical to the oracle's drift_and_constrain, needs_rebuild, n_threads.
"""
import os
import sys

import numpy as np
import pytest

import geodesic_hip as gh
from geodesic_hip import host
from oracle import amber
from oracle import oracle as orc

from _common import FIX

pytestmark = pytest.mark.gpu

PRM = os.path.join(FIX, "ala_dipeptide.prmtop")
CRD = os.path.join(FIX, "ala_dipeptide.inpcrd")
BIG = (1000.0, 1000.0, 1000.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_terms(scale14):
    with open(PRM) as f:
        atoms, topo = amber.parse_prmtop(f.read())
    with open(CRD) as f:
        x, y, z, *_ = amber.parse_inpcrd(f.read(), len(atoms.mass))
    if scale14 == 0.0:
        topo.pair14_i = np.zeros(0, np.uint32)
        topo.pair14_j = np.zeros(0, np.uint32)
    pi, pj = orc.neighbor_build(x, y, z, BIG, 14.0, topo.excl_i, topo.excl_j)
    t, *_ = orc.cpu_backend_compute(x, y, z, atoms.sigma, atoms.epsilon, topo, pi, pj, BIG, 12.0, 10.0,
                                    scale14=scale14, periodic=False)
    return t


@pytest.mark.parametrize("fidelity", [False, True])
def test_energy_from_files_matches_oracle(fidelity):
    say(f"[gpu] geodesic energy, fidelity={fidelity}")
    r = host.energy_from_files(PRM, CRD, fidelity=fidelity)
    t = oracle_terms(0.5 if fidelity else 0.0)
    for a, b in (("bond", "bond"), ("angle", "angle"), ("dihedral", "dihedral"), ("nonbonded", "lj"),
                 ("lj14", "lj14"), ("total", "total")):
        ref = getattr(t, b)
        if ref == 0.0:
            assert r[a] == 0.0, a
        else:
            assert abs(r[a] - ref) <= 1e-6 * abs(ref), (a, r[a], ref)
    assert r["r_switch"] == 10.0 and r["r_cutoff"] == 12.0
    assert (r["lj14"] != 0.0) == fidelity


def golden_run(tmp_path, **kw):
    dcd, csv = str(tmp_path / "t.dcd"), str(tmp_path / "e.csv")
    s = host.run(PRM, CRD, dcd, csv, n_steps=99, frame_interval=1, dt=0.001, box_size=BIG, periodic=False,
                 r_cutoff=12.0, r_skin=14.0, r_switch=10.0, max_constr_iter=100, constr_tol=1e-8,
                 constrain_hydrogens=True, total_energy=5.12, **kw)
    return s, dcd, csv


def test_run_reproduces_golden_trajectory(tmp_path):
    say("[gpu] geodesic run, golden config")
    s, dcd, csv = golden_run(tmp_path)
    assert s["n_frames"] == 100 and s["n_steps"] == 99
    with open(os.path.join(FIX, "ala_dipeptide_ref.dcd"), "rb") as f:
        golden = f.read()
    with open(dcd, "rb") as f:
        out = f.read()
    assert len(out) == len(golden)
    head = 4 + 4 + 80 + 4 + 4 + 4 + 80 + 4 + 12
    assert out[:head] == golden[:head]
    _, gf = amber.read_dcd(golden)
    _, of = amber.read_dcd(out)
    dev = float(np.abs(of.astype(np.float64) - gf).max())
    say(f"[gpu] max |frame - golden| = {dev:.3g} A")
    assert of[0].tobytes() == gf[0].tobytes() and dev <= 1e-4
    rows = open(csv).read().splitlines()
    assert rows[0] == "step,time_ps,potential_kcal,kinetic_kcal,total_kcal,temperature_K"
    assert len(rows) == 101
    # the same loop through the engine's own entry point gives the energies the log formats
    with open(PRM) as f:
        atoms, topo = amber.parse_prmtop(f.read())
    amber.promote_hydrogen_bonds(topo, atoms)
    with open(CRD) as f:
        x, y, z, *_ = amber.parse_inpcrd(f.read(), len(atoms.mass))
    g = gh.System(atoms.sigma, atoms.epsilon, topo,
                  gh.Params(box=BIG, periodic=False, r_cutoff=12.0, r_switch=10.0, r_skin=14.0))
    g.dynamics_set(atoms.mass, topo.constr_i, topo.constr_j, topo.constr_dsq)
    d = [gh.DeviceArray.from_host(np.ascontiguousarray(a, np.float64)) for a in (x, y, z, *([np.zeros(22)] * 3))]
    frames, epot, ekin = g.md_run(*d, n_steps=99, dt=0.001, frame_interval=1, max_iter=100, tol=1e-8)
    assert frames.tobytes() == of.tobytes()
    kb = 0.0019872041
    for k in (0, 1, 50, 99):
        f = rows[k + 1].split(",")
        assert f[0] == str(k) and f[1] == f"{k * 0.001:.3f}"
        assert f[2] == f"{epot[k]:.3f}" and f[3] == f"{ekin[k]:.3f}" and f[4] == f"{epot[k] + ekin[k]:.3f}"
        assert f[5] == f"{2.0 * ekin[k] / (3.0 * 22 * kb):.2f}"


def test_run_energy_drift_stop(tmp_path):
    with pytest.raises(gh.GmError) as ei:
        golden_run(tmp_path, monitoring=(1e-12, "stop"))
    assert ei.value.status == host.GMH_ERR_NUMERICAL and "energy drift" in str(ei.value)
    hdr, frames = amber.read_dcd(open(tmp_path / "t.dcd", "rb").read())
    rows = open(tmp_path / "e.csv").read().splitlines()
    # the reference writes the frame and the row, then checks (lib.rs:322-328)
    assert hdr["nset"] == len(frames) == 2 and len(rows) == 3


def test_run_rejects_a_cpu_backend(tmp_path):
    with pytest.raises(gh.GmError) as ei:
        golden_run(tmp_path, backend="cpu")
    assert ei.value.status == host.GMH_ERR_CONFIG and "run.backend" in str(ei.value)


def test_backend_trait_over_gpu_backend():
    say("[gpu] ComputeBackend trait of the C++ GpuBackend")
    t = host.Topology(PRM)
    t.promote_hydrogen_bonds()
    n = 22
    pos, vel = host.read_inpcrd(CRD, n)
    rng = np.random.default_rng(3)
    vel = rng.normal(0, 2.0, (3, n))
    b = host.Backend(t, BIG, False, 12.0, 14.0, 10.0, max_constr_iter=100, constr_tol=1e-8)
    assert b.n_threads() == 1
    b.set_state(pos, vel, step=0)
    b.build_neighbor_list()
    f, e = b.compute_forces()
    # the engine's host entry point on the same system
    with open(PRM) as fh:
        atoms, topo = amber.parse_prmtop(fh.read())
    amber.promote_hydrogen_bonds(topo, atoms)
    g = gh.System(atoms.sigma, atoms.epsilon, topo,
                  gh.Params(box=BIG, periodic=False, r_cutoff=12.0, r_switch=10.0, r_skin=14.0))
    x, y, z = pos.copy()
    g.build_neighbor_list(x, y, z)
    e2, fx, fy, fz = g.compute_forces(x, y, z)
    assert np.stack([fx, fy, fz]).tobytes() == f.tobytes() and e.total == e2.total
    assert not b.needs_rebuild()
    # geodesic_drift == integrator::drift_and_constrain (oracle) byte for byte
    b.geodesic_drift(0.001)
    p2, v2 = b.state()
    ox, oy, oz = pos.copy()
    ovx, ovy, ovz = vel.copy()
    orc.drift_and_constrain(ox, oy, oz, ovx, ovy, ovz, topo, atoms.mass, 0.001, 100, 1e-8)
    assert np.stack([ox, oy, oz]).tobytes() == p2.tobytes()
    assert np.stack([ovx, ovy, ovz]).tobytes() == v2.tobytes()
    # a non-converging drift is ConvergenceError
    b2 = host.Backend(t, BIG, False, 12.0, 14.0, 10.0, max_constr_iter=1, constr_tol=0.0)
    b2.set_state(pos, np.full((3, n), 5.0))
    with pytest.raises(gh.GmError) as ei:
        b2.geodesic_drift(0.01)
    assert ei.value.status == 9 and "converge" in str(ei.value)


def test_python_backend_mirror_geodesic_drift():
    """The ctypes GpuBackend mirror's geodesic_drift, in place on a SimState-
    like object, equals the oracle's drift_and_constrain byte for byte."""
    with open(PRM) as f:
        atoms, topo = amber.parse_prmtop(f.read())
    amber.promote_hydrogen_bonds(topo, atoms)
    with open(CRD) as f:
        x, y, z, *_ = amber.parse_inpcrd(f.read(), len(atoms.mass))

    class State:
        pass

    st = State()
    rng = np.random.default_rng(9)
    st.pos_x, st.pos_y, st.pos_z = x.copy(), y.copy(), z.copy()
    st.vel_x, st.vel_y, st.vel_z = (rng.normal(0, 2.0, 22) for _ in range(3))
    st.step = 0
    ref = [a.copy() for a in (st.pos_x, st.pos_y, st.pos_z, st.vel_x, st.vel_y, st.vel_z)]
    b = gh.GpuBackend.try_new(atoms, topo, gh.Params(box=BIG, periodic=False, r_cutoff=12.0, r_switch=10.0,
                                                     r_skin=14.0), max_constr_iter=100, constr_tol=1e-8)
    b.geodesic_drift(st, 0.001)
    orc.drift_and_constrain(*ref, topo, atoms.mass, 0.001, 100, 1e-8)
    for a, r in zip((st.pos_x, st.pos_y, st.pos_z, st.vel_x, st.vel_y, st.vel_z), ref):
        assert a.tobytes() == r.tobytes()
