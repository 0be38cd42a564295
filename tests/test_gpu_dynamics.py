"""Device-resident dynamics on the GPU against the f64 oracle and the
reference's golden trajectory.

integrator.rs:14-71 and constraint.rs:62-259 (paths relative to the reference
repository root): half-kick, drift + SHAKE and RATTLE must be byte-identical
to the oracle's restatement (itself pinned by the golden DCD) given the same
inputs.  The whole run loop (geodesic/src/lib.rs:255-330) with GPU forces is
compared with geodesic/tests/golden/ala_dipeptide_ref.dcd."""
import sys

import numpy as np
import pytest

import geodesic_hip as gh
from geodesic_hip import systems
from oracle import amber
from oracle import oracle as orc

from _common import FIX, load

pytestmark = pytest.mark.gpu

BIG = (1000.0, 1000.0, 1000.0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def dev(a):
    return gh.DeviceArray.from_host(np.ascontiguousarray(a, np.float64))


def host(*arrs):
    return [a.to_host() for a in arrs]


# Constraint solver modes (gm_set_constraint_groups): groups of whole
# components iterated in LDS; one launch per iteration over all atoms; and
# groups handing over to per-iteration launches after two iterations.  All
# three must reproduce the oracle's bytes and iteration counts.
SOLVERS = {"groups": (True, 0), "iterations": (False, 0), "handover": (True, 2)}


def use_solver(g, name):
    enable, cap = SOLVERS[name]
    n_groups = g.set_constraint_groups(enable, cap)
    return n_groups


def ala_system():
    (x, y, z), atoms, topo = load("ala_dipeptide")
    amber.promote_hydrogen_bonds(topo, atoms)
    params = gh.Params(box=BIG, periodic=False, r_cutoff=12.0, r_switch=10.0, r_skin=14.0)
    g = gh.System(atoms.sigma, atoms.epsilon, topo, params)
    g.dynamics_set(atoms.mass, topo.constr_i, topo.constr_j, topo.constr_dsq)
    return (x, y, z), atoms, topo, params, g


@pytest.mark.parametrize("solver", list(SOLVERS))
def test_half_kick_drift_rattle_match_oracle_bytes(solver):
    say(f"[gpu] dynamics primitives, ala_dipeptide, {solver}")
    (x, y, z), atoms, topo, params, g = ala_system()
    assert use_solver(g, solver) == 1  # the 12 X-H constraints fit one group
    rng = np.random.default_rng(3)
    vx, vy, vz = (rng.normal(0, 2.0, 22) for _ in range(3))
    pi, pj = orc.neighbor_build(x, y, z, BIG, 14.0, topo.excl_i, topo.excl_j)
    t, fx, fy, fz = orc.cpu_backend_compute(x, y, z, atoms.sigma, atoms.epsilon, topo, pi, pj, BIG, 12.0, 10.0,
                                            n_threads=1, periodic=False)
    d = [dev(a) for a in (x, y, z, vx, vy, vz, fx, fy, fz)]
    # B
    g.half_kick_device(*d[3:6], *d[6:9], 0.0005)
    ox, oy, oz, ovx, ovy, ovz = (a.copy() for a in (x, y, z, vx, vy, vz))
    orc.half_kick(ovx, ovy, ovz, fx, fy, fz, atoms.mass, 0.0005)
    g.sync()
    assert np.stack(host(*d[3:6])).tobytes() == np.stack([ovx, ovy, ovz]).tobytes()
    # A: drift + SHAKE + velocity resync
    it = g.drift_and_constrain_device(*d[0:6], 0.001, 100, 1e-8, step=0)
    oit = orc.drift_and_constrain(ox, oy, oz, ovx, ovy, ovz, topo, atoms.mass, 0.001, 100, 1e-8)
    assert it == oit and it > 1
    assert np.stack(host(*d[0:3])).tobytes() == np.stack([ox, oy, oz]).tobytes()
    assert np.stack(host(*d[3:6])).tobytes() == np.stack([ovx, ovy, ovz]).tobytes()
    # RATTLE
    it = g.constrain_velocities_device(*d[0:6], 100, 1e-8)
    oit = orc.constrain_velocities(ox, oy, oz, ovx, ovy, ovz, topo, atoms.mass, 100, 1e-8)
    assert it == oit
    assert np.stack(host(*d[3:6])).tobytes() == np.stack([ovx, ovy, ovz]).tobytes()


@pytest.mark.parametrize("solver", list(SOLVERS))
def test_shake_rattle_many_constraints_match_oracle_bytes(solver):
    """3,000 atoms; every triplet is a water-like unit (heavy centre, two
    light arms held at their current lengths), so the centre atoms take two
    constraints each: the per-atom Jacobi sums still match byte for byte.
    The 875 triplets pack into groups of 85 (255 atoms)."""
    say(f"[gpu] dynamics primitives, constrained triplets, {solver}")
    n = 3000
    s = systems.protein_solvent(n, (n / 0.1) ** (1 / 3), n // 8, seed=51)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    tri = np.arange(n // 8, n - 2, 3)
    ci = np.concatenate([tri, tri]).astype(np.uint32)
    cj = np.concatenate([tri + 1, tri + 2]).astype(np.uint32)
    dsq = (x[ci] - x[cj]) ** 2 + (y[ci] - y[cj]) ** 2 + (z[ci] - z[cj]) ** 2
    z32, zf = np.zeros(0, np.uint32), np.zeros(0)
    topo = amber.Topology(z32, z32, zf, zf, z32, z32, z32, zf, zf, z32, z32, z32, z32, zf, z32, zf, z32, z32,
                          constr_i=ci, constr_j=cj, constr_dsq=dsq)
    mass = np.full(n, 12.0)
    mass[tri], mass[tri + 1], mass[tri + 2] = 16.0, 1.008, 1.008
    g = gh.System(s.sigma, s.epsilon, s.topo, s.params)
    g.dynamics_set(mass, ci, cj, dsq)
    assert use_solver(g, solver) == -(-len(tri) // 85)
    rng = np.random.default_rng(5)
    vx, vy, vz = (rng.normal(0, 3.0, n) for _ in range(3))
    d = [dev(a) for a in (x, y, z, vx, vy, vz)]
    it = g.drift_and_constrain_device(*d, 0.002, 200, 1e-8)
    oit = orc.drift_and_constrain(x, y, z, vx, vy, vz, topo, mass, 0.002, 200, 1e-8)
    assert it == oit and it > 2
    assert np.stack(host(*d[0:3])).tobytes() == np.stack([x, y, z]).tobytes()
    assert np.stack(host(*d[3:6])).tobytes() == np.stack([vx, vy, vz]).tobytes()
    it = g.constrain_velocities_device(*d, 200, 1e-8)
    oit = orc.constrain_velocities(x, y, z, vx, vy, vz, topo, mass, 200, 1e-8)
    assert it == oit
    assert np.stack(host(*d[3:6])).tobytes() == np.stack([vx, vy, vz]).tobytes()


def test_constraint_solver_reports_non_convergence():
    (x, y, z), atoms, topo, params, g = ala_system()
    msgs = []
    for solver in SOLVERS:
        use_solver(g, solver)
        d = [dev(a) for a in (x, y, z, np.full(22, 5.0), np.zeros(22), np.zeros(22))]
        with pytest.raises(gh.GmError) as ei:
            g.drift_and_constrain_device(*d, 0.01, 3, 0.0)
        assert ei.value.status == 9 and "did not converge" in str(ei.value)
        msgs.append(str(ei.value))
    # the same worst constraint, residual and count whichever solver ran
    assert msgs[0] == msgs[1] == msgs[2]


@pytest.mark.parametrize("chain,groups", [(300, 0), (256, 1), (257, 0)])
def test_constrained_chain_at_and_past_group_capacity(chain, groups):
    """A chain with every bond constrained is one component of `chain` atoms:
    256 atoms (255 constraints) still fit one group, 257 or 300 do not and
    take the per-iteration launches.  The oracle's bytes either way."""
    say(f"[gpu] dynamics primitives, {chain}-atom constrained chain")
    n = 320
    s = systems.protein_solvent(n, (n / 0.1) ** (1 / 3), 300, seed=53)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    ci = np.arange(0, chain - 1, dtype=np.uint32)
    cj = ci + 1
    dsq = (x[ci] - x[cj]) ** 2 + (y[ci] - y[cj]) ** 2 + (z[ci] - z[cj]) ** 2
    z32, zf = np.zeros(0, np.uint32), np.zeros(0)
    topo = amber.Topology(z32, z32, zf, zf, z32, z32, z32, zf, zf, z32, z32, z32, z32, zf, z32, zf, z32, z32,
                          constr_i=ci, constr_j=cj, constr_dsq=dsq)
    mass = np.where(np.arange(n) % 2 == 0, 12.0, 14.0)
    g = gh.System(s.sigma, s.epsilon, s.topo, s.params)
    g.dynamics_set(mass, ci, cj, dsq)
    assert g.set_constraint_groups(True, 0) == groups
    rng = np.random.default_rng(7)
    vx, vy, vz = (rng.normal(0, 1.0, n) for _ in range(3))
    d = [dev(a) for a in (x, y, z, vx, vy, vz)]
    it = g.drift_and_constrain_device(*d, 0.001, 500, 1e-8)
    oit = orc.drift_and_constrain(x, y, z, vx, vy, vz, topo, mass, 0.001, 500, 1e-8)
    assert it == oit and it > 2
    assert np.stack(host(*d[0:3])).tobytes() == np.stack([x, y, z]).tobytes()
    it = g.constrain_velocities_device(*d, 500, 1e-8)
    oit = orc.constrain_velocities(x, y, z, vx, vy, vz, topo, mass, 500, 1e-8)
    assert it == oit
    assert np.stack(host(*d[3:6])).tobytes() == np.stack([vx, vy, vz]).tobytes()


def test_md_run_reproduces_golden_trajectory():
    """geodesic/tests/golden_reference.rs:27-94 on the device: 99 steps of
    1 fs, one frame per step.  Forces are FP32 pair arithmetic here, so frames
    agree with the committed f64 trajectory to a small tolerance, not bytes."""
    say("[gpu] md run vs golden DCD")
    with open(f"{FIX}/ala_dipeptide_ref.dcd", "rb") as f:
        hdr, golden = amber.read_dcd(f.read())
    (x, y, z), atoms, topo, params, g = ala_system()
    d = [dev(a) for a in (x, y, z, np.zeros(22), np.zeros(22), np.zeros(22))]
    frames, epot, ekin = g.md_run(*d, n_steps=99, dt=0.001, frame_interval=1, max_iter=100, tol=1e-8)
    assert frames.shape == golden.shape == (100, 3, 22)
    dev_max = float(np.abs(frames.astype(np.float64) - golden).max())
    say(f"[gpu] max |frame - golden| = {dev_max:.3g} A; E0 = {epot[0] + ekin[0]:.6f}, "
        f"E99 = {epot[-1] + ekin[-1]:.6f}")
    assert frames[0].tobytes() == golden[0].tobytes()
    assert dev_max <= 1e-4
    # potential of frame 0 against the oracle; the reference records 5.12
    pi, pj = orc.neighbor_build(x, y, z, BIG, 14.0, topo.excl_i, topo.excl_j)
    t, *_ = orc.cpu_backend_compute(x, y, z, atoms.sigma, atoms.epsilon, topo, pi, pj, BIG, 12.0, 10.0,
                                    periodic=False)
    assert abs(epot[0] - t.total) <= 1e-6 * abs(t.total)
    assert abs((epot[-1] + ekin[-1]) - (epot[0] + ekin[0])) <= 1e-2


def test_md_run_same_bytes_with_every_solver():
    """The run loop checks a step's RATTLE with the next step's SHAKE (one
    read-back a step) and replays the step when RATTLE must continue per
    iteration; caps of 1..6 group iterations send some or all solves of both
    kinds down that path.  Every way, the run's bytes are those of the
    per-iteration solver."""
    say("[gpu] md run, solver modes")
    (x, y, z), atoms, topo, params, g = ala_system()
    rng = np.random.default_rng(11)
    v0 = [rng.normal(0, 3.0, 22) for _ in range(3)]
    runs = {}
    for name, (enable, cap) in [("iterations", (False, 0)), ("groups", (True, 0))] + \
            [(f"cap{c}", (True, c)) for c in range(1, 7)]:
        g.set_constraint_groups(enable, cap)
        d = [dev(a) for a in (x, y, z, *v0)]
        frames, epot, ekin = g.md_run(*d, n_steps=30, dt=0.001, frame_interval=7, max_iter=100, tol=1e-8)
        runs[name] = (frames.tobytes(), epot.tobytes(), ekin.tobytes(), np.stack(host(*d)).tobytes())
    for name, r in runs.items():
        assert r == runs["iterations"], name


def test_jacobi_factor():
    (x, y, z), atoms, topo, params, g = ala_system()
    g.build_neighbor_list(x, y, z)
    terms, fx, fy, fz = g.compute_forces(x, y, z)
    assert g.jacobi_factor(5.12) == 2.0 * (5.12 - terms.total)
