"""The run loop's failure contract and the force path's non-finite reporting.

Paths relative to the reference repository root.  The reference checks
positions after every drift and forces after every evaluation
(geodesic/src/lib.rs:79-107, called at :269, :296 and :307) and stops with
NumericalError::NanInPos / NanInForce {step, atom, component}
(geodesic-core/src/error.rs:35-41).  The engine reproduces that order and
report; the expected failure of each case comes from the f64 oracle running
the same loop with the same checks.

Forces are accumulated in int64 fixed point (2^-32 kcal/mol/A): a pair or
bonded contribution of 2^30 kcal/mol/A or more does not fit, and is reported
as a non-finite force (NaN in x, y and z of that atom) instead of wrapping.
The f64 reference would carry such a force on; the tests state that range
(two atoms 0.3 A apart) and check that forces just inside it still match the
oracle (1.0 A apart).
"""
import os
import sys

import numpy as np
import pytest

import geodesic_hip as gh
from geodesic_hip import host, systems
from oracle import amber
from oracle import oracle as orc

from _common import FIX, force_rms_rel, load

pytestmark = pytest.mark.gpu

BIG = (1000.0, 1000.0, 1000.0)
FIXED_RANGE = 2.0 ** 30  # kcal/mol/A: the largest force contribution int64 fixed point holds


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def no_topology():
    z32, zf = np.zeros(0, np.uint32), np.zeros(0)
    return systems.Topology(z32, z32, zf, zf, z32, z32, z32, zf, zf, z32, z32, z32, z32, zf, z32, zf, z32, z32)


def two_atoms(r):
    x, y, z = np.array([100.0, 100.0 + r]), np.array([100.0, 100.0]), np.array([100.0, 100.0])
    sig, eps = np.array([3.8, 3.8]), np.array([0.25, 0.25])
    p = gh.Params(box=BIG, periodic=False, r_cutoff=12.0, r_switch=10.0, r_skin=14.0)
    return (x, y, z), sig, eps, p


def test_contact_inside_the_fixed_point_range_matches_oracle():
    (x, y, z), sig, eps, p = two_atoms(1.0)
    g = gh.System(sig, eps, no_topology(), p)
    g.build_neighbor_list(x, y, z)
    e, fx, fy, fz = g.compute_forces(x, y, z)
    t, ox, oy, oz = orc.cpu_backend_compute(x, y, z, sig, eps, no_topology(), np.array([0], np.uint32),
                                            np.array([1], np.uint32), BIG, 12.0, 10.0, periodic=False)
    assert 1e7 < abs(ox[0]) < FIXED_RANGE
    assert abs(e.lj - t.lj) <= 1e-6 * abs(t.lj)
    assert force_rms_rel(fx, fy, fz, ox, oy, oz) <= 1e-5


def test_contact_beyond_the_fixed_point_range_is_reported_not_wrapped():
    """0.3 A with sigma 3.8, eps 0.25: |F| ~ 7e14 kcal/mol/A, past 2^30."""
    (x, y, z), sig, eps, p = two_atoms(0.3)
    g = gh.System(sig, eps, no_topology(), p)
    g.build_neighbor_list(x, y, z)
    _, ox, _, _ = orc.cpu_backend_compute(x, y, z, sig, eps, no_topology(), np.array([0], np.uint32),
                                          np.array([1], np.uint32), BIG, 12.0, 10.0, periodic=False)
    assert abs(ox[0]) > FIXED_RANGE and np.isfinite(ox[0])
    with pytest.raises(gh.GmError) as ei:
        g.compute_forces(x, y, z)
    assert ei.value.status == gh.GM_ERR_NUMERICAL and "atom 0" in str(ei.value)
    e, fx, fy, fz = ei.value.result
    assert np.isnan(fx).all() and np.isnan(fy).all() and np.isnan(fz).all()
    # the flag belongs to that evaluation: a contact back in range is clean again
    x2 = np.array([100.0, 101.0])
    e2, fx2, _, _ = g.compute_forces(x2, y, z)
    assert np.isfinite(fx2).all()


@pytest.fixture(scope="module")
def sys3k():
    s = systems.protein_solvent(3000, (3000 / 0.1) ** (1 / 3), 375, seed=7)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    orc.wrap_positions(x, y, z, s.params.box)
    pi, pj = orc.neighbor_build(x, y, z, s.params.box, s.params.r_skin, s.topo.excl_i, s.topo.excl_j,
                                method="brute")
    return s, (x, y, z), (pi, pj)


def test_nan_position_marks_the_atom_and_every_listed_and_bonded_partner(sys3k):
    """compute_forces with one coordinate NaN, on a list built before it: the
    f64 path gives NaN forces on that atom and on every atom sharing a listed
    pair or a bonded term with it, a NaN energy, and finite forces elsewhere."""
    s, (x, y, z), (pi, pj) = sys3k
    p = s.params
    g = gh.System(s.sigma, s.epsilon, s.topo, p)
    gx, gy, gz = x.copy(), y.copy(), z.copy()
    g.build_neighbor_list(gx, gy, gz)
    for k in (17, 2999):  # a chain atom (bonded partners) and a solvent atom
        say(f"[gpu] NaN position at atom {k}")
        bx = gx.copy()
        bx[k] = np.nan
        t, ox, oy, oz = orc.cpu_backend_compute(bx, gy, gz, s.sigma, s.epsilon, s.topo, pi, pj, p.box,
                                                p.r_cutoff, p.r_switch, n_threads=8, scale14=p.scale_14,
                                                bonded_min_image=p.bonded_min_image, periodic=p.periodic)
        bad_ref = ~(np.isfinite(ox) & np.isfinite(oy) & np.isfinite(oz))
        with pytest.raises(gh.GmError) as ei:
            g.compute_forces(bx, gy, gz)
        assert ei.value.status == gh.GM_ERR_NUMERICAL
        e, fx, fy, fz = ei.value.result
        bad = np.isnan(fx) & np.isnan(fy) & np.isnan(fz)
        assert np.array_equal(bad, bad_ref), (np.flatnonzero(bad), np.flatnonzero(bad_ref))
        assert f"atom {int(np.flatnonzero(bad_ref)[0])}" in str(ei.value)
        assert np.isnan(t.total) and np.isnan(e.total)
        ok = ~bad_ref
        assert np.isfinite(fx[ok]).all()
        assert force_rms_rel(fx[ok], fy[ok], fz[ok], ox[ok], oy[ok], oz[ok]) <= 1e-5
    e0, *f0 = g.compute_forces(gx, gy, gz)  # finite positions again: clean
    assert np.isfinite(e0.total)


def test_nan_position_is_never_listed(sys3k):
    """neighbor.rs:25-43: r2 <= r_skin^2 is false for a NaN distance."""
    s, (x, y, z), _ = sys3k
    bx = x.copy()
    bx[1234] = np.nan
    pi, pj = orc.neighbor_build(bx, y, z, s.params.box, s.params.r_skin, s.topo.excl_i, s.topo.excl_j,
                                method="brute")
    assert not np.isin(1234, pi) and not np.isin(1234, pj)
    g = gh.System(s.sigma, s.epsilon, s.topo, s.params)
    gx, gy, gz = bx.copy(), y.copy(), z.copy()
    g.build_neighbor_list(gx, gy, gz)
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)


# ---- the run loop -----------------------------------------------------------

def oracle_run(x, y, z, vx, vy, vz, atoms, topo, n_steps, dt, box=BIG, rc=12.0, rs=14.0, rsw=10.0,
               mi=100, tol=1e-8, fi=1):
    """lib.rs:255-333 over the oracle, with the reference's checks (and the
    fixed-point range as a force check).  Returns (kind, step, atom,
    component, frames) of the first failure, kind 0 if none."""
    x, y, z, vx, vy, vz = (a.copy() for a in (x, y, z, vx, vy, vz))

    def build():
        pi, pj = orc.neighbor_build(x, y, z, box, rs, topo.excl_i, topo.excl_j, method="brute")
        return pi, pj, x.copy(), y.copy(), z.copy()

    def forces(pi, pj):
        return orc.cpu_backend_compute(x, y, z, atoms.sigma, atoms.epsilon, topo, pi, pj, box, rc, rsw,
                                       n_threads=1, periodic=False)[1:]

    def first_bad(arrs, limit=None):
        bad = np.zeros(len(x), bool)
        for a in arrs:
            bad |= ~np.isfinite(a) if limit is None else ~(np.abs(a) < limit)
        if not bad.any():
            return None
        k = int(np.flatnonzero(bad)[0])
        comp = next(c for c, a in enumerate(arrs) if not (np.isfinite(a[k]) if limit is None else abs(a[k]) < limit))
        return k, comp

    pi, pj, rx, ry, rz = build()
    f = forces(pi, pj)
    b = first_bad(f, FIXED_RANGE)
    if b:
        return 2, 0, b[0], 0, 0
    assert orc.constrain_velocities(x, y, z, vx, vy, vz, topo, atoms.mass, mi, tol) >= 0
    frames = 1
    for step in range(n_steps):
        orc.half_kick(vx, vy, vz, *f, atoms.mass, dt / 2)
        assert orc.drift_and_constrain(x, y, z, vx, vy, vz, topo, atoms.mass, dt, mi, tol) >= 0
        b = first_bad((x, y, z))
        if b:
            return 1, step, b[0], b[1], frames
        if orc.needs_rebuild(x, y, z, rx, ry, rz, rs, rc):
            pi, pj, rx, ry, rz = build()
        f = forces(pi, pj)
        b = first_bad(f, FIXED_RANGE)
        if b:
            return 2, step, b[0], 0, frames  # a flagged atom's force is NaN in x, y and z
        orc.half_kick(vx, vy, vz, *f, atoms.mass, dt / 2)
        assert orc.constrain_velocities(x, y, z, vx, vy, vz, topo, atoms.mass, mi, tol) >= 0
        if (step + 1) % fi == 0:
            frames += 1
    return 0, n_steps, -1, 0, frames


def ala():
    (x, y, z), atoms, topo = load("ala_dipeptide")
    amber.promote_hydrogen_bonds(topo, atoms)
    return (x, y, z), atoms, topo


def gpu_run(x, y, z, vx, vy, vz, atoms, topo, n_steps, dt=0.001, groups=True, p=None, **kw):
    p = p or gh.Params(box=BIG, periodic=False, r_cutoff=12.0, r_switch=10.0, r_skin=14.0)
    g = gh.System(atoms.sigma, atoms.epsilon, topo, p)
    g.dynamics_set(atoms.mass, topo.constr_i, topo.constr_j, topo.constr_dsq)
    g.set_constraint_groups(groups, 0)
    d = [gh.DeviceArray.from_host(np.ascontiguousarray(a, np.float64)) for a in (x, y, z, vx, vy, vz)]
    try:
        frames, *_ = g.md_run(*d, n_steps=n_steps, dt=dt, **kw)
        return g, None, len(frames)
    except gh.GmError as err:
        return g, err, len(err.result[0])


@pytest.mark.parametrize("groups", [True, False])
@pytest.mark.parametrize("which", ["free", "constrained"])
def test_md_run_stops_at_a_nan_position(groups, which):
    """A NaN velocity: on an atom in no constraint the next drift makes its
    position NaN (NanInPos at step 0); on a hydrogen, the velocity projection
    first spreads it to its constraint partner (a NaN multiplier does not
    keep the solver from converging: constraint.rs:127-131), so the first
    atom of the pair is reported."""
    (x, y, z), atoms, topo = ala()
    cons = set(topo.constr_i.tolist()) | set(topo.constr_j.tolist())
    k = next(a for a in range(22) if (a not in cons) == (which == "free"))
    vx, vy, vz = np.zeros(22), np.zeros(22), np.zeros(22)
    vy[k] = np.nan
    exp = oracle_run(x, y, z, vx, vy, vz, atoms, topo, 5, 0.001)
    say(f"[gpu] md NaN velocity at atom {k} ({which}, groups {groups}): oracle {exp}")
    assert exp[0] == 1
    g, err, nf = gpu_run(x, y, z, vx, vy, vz, atoms, topo, 5, groups=groups)
    assert err is not None and err.status == gh.GM_ERR_NUMERICAL
    f = g.md_last_failure()
    assert (f["kind"], f["step"], f["atom"], f["component"], nf) == exp
    axis = "XYZ"[exp[3]]
    assert f"NaN in position_{axis} at step {exp[1]}, atom {exp[2]}" in str(err)


@pytest.mark.parametrize("groups", [True, False])
def test_md_run_stops_at_a_force_out_of_range(groups):
    """Two atoms closing at 2400 A/ps from 2.7 A are 0.3 A apart after the
    first drift: the force of step 0 is past the fixed-point range and is
    reported as NanInForce at step 0, after the one frame taken at the start."""
    (x, y, z), sig, eps, p = two_atoms(2.7)
    atoms = type("A", (), {})()
    atoms.sigma, atoms.epsilon, atoms.mass = sig, eps, np.array([12.0, 12.0])
    topo = no_topology()
    vx, vy, vz = np.array([1200.0, -1200.0]), np.zeros(2), np.zeros(2)
    exp = oracle_run(x, y, z, vx, vy, vz, atoms, topo, 3, 0.001)
    say(f"[gpu] md force out of range: oracle {exp}")
    assert exp[:3] == (2, 0, 0)
    g, err, nf = gpu_run(x, y, z, vx, vy, vz, atoms, topo, 3, groups=groups, p=p)
    assert err is not None and err.status == gh.GM_ERR_NUMERICAL
    f = g.md_last_failure()
    assert (f["kind"], f["step"], f["atom"], f["component"], nf) == exp
    assert "NaN in force_X at step 0, atom 0" in str(err)


def test_md_run_in_resumed_segments_equals_one_run():
    """gm_md_run cut into calls that end on frame steps (resume) gives the
    bytes of one call: frames, energies, final positions and velocities."""
    (x, y, z), atoms, topo = ala()
    rng = np.random.default_rng(5)
    v0 = [rng.normal(0, 3.0, 22) for _ in range(3)]
    p = gh.Params(box=BIG, periodic=False, r_cutoff=12.0, r_switch=10.0, r_skin=14.0)
    outs = []
    for cuts in ([30], [10, 10, 10], [5, 20, 5]):
        g = gh.System(atoms.sigma, atoms.epsilon, topo, p)
        g.dynamics_set(atoms.mass, topo.constr_i, topo.constr_j, topo.constr_dsq)
        d = [gh.DeviceArray.from_host(np.ascontiguousarray(a, np.float64)) for a in (x, y, z, *v0)]
        fr, ep, ek, done = [], [], [], 0
        for c in cuts:
            f, e, k = g.md_run(*d, n_steps=c, dt=0.001, frame_interval=5, first_step=done, resume=done > 0)
            fr.append(f), ep.append(e), ek.append(k)
            done += c
        outs.append((np.concatenate(fr).tobytes(), np.concatenate(ep).tobytes(), np.concatenate(ek).tobytes(),
                     np.stack([a.to_host() for a in d]).tobytes()))
    assert outs[0] == outs[1] == outs[2]


def test_run_failure_keeps_the_frames_before_it(tmp_path):
    """`geodesic run` with a NaN velocity in the inpcrd's place: the frames
    before the failing step are on disk, the error is NanInPos."""
    (x, y, z), atoms, topo = ala()
    crd = tmp_path / "nan.inpcrd"
    # an inpcrd of the fixture's positions; the run then starts at rest, so
    # force the NaN through a huge time step instead: dt 1e300 ps overflows
    # the first drift to a non-finite position
    with open(os.path.join(FIX, "ala_dipeptide.inpcrd")) as f:
        crd.write_text(f.read())
    dcd, csv = str(tmp_path / "t.dcd"), str(tmp_path / "e.csv")
    with pytest.raises(gh.GmError) as ei:
        host.run(os.path.join(FIX, "ala_dipeptide.prmtop"), str(crd), dcd, csv, n_steps=5, frame_interval=1,
                 dt=1e300, box_size=BIG, periodic=False, r_cutoff=12.0, r_skin=14.0, r_switch=10.0,
                 max_constr_iter=100, constr_tol=1e-8, constrain_hydrogens=False, total_energy=0.0)
    assert ei.value.status == host.GMH_ERR_NUMERICAL and "NaN in position_" in str(ei.value)
    hdr, frames = amber.read_dcd(open(dcd, "rb").read())
    assert hdr["nset"] == len(frames) == 1 and len(open(csv).read().splitlines()) == 2


def test_hybrid_backend_is_an_alias_of_gpu(tmp_path):
    """geodesic/tests/gpu_hybrid_alias.rs: the DCD bytes of backend "hybrid"
    and "gpu" are identical."""
    out = {}
    for b in ("gpu", "hybrid"):
        dcd, csv = str(tmp_path / f"{b}.dcd"), str(tmp_path / f"{b}.csv")
        host.run(os.path.join(FIX, "ala_dipeptide.prmtop"), os.path.join(FIX, "ala_dipeptide.inpcrd"), dcd, csv,
                 n_steps=20, frame_interval=2, dt=0.001, box_size=BIG, periodic=False, r_cutoff=12.0,
                 r_skin=14.0, r_switch=10.0, max_constr_iter=100, constr_tol=1e-8, constrain_hydrogens=True,
                 total_energy=0.0, backend=b)
        out[b] = (open(dcd, "rb").read(), open(csv).read())
    assert out["gpu"] == out["hybrid"] and len(out["gpu"][0]) > 0


def test_jacobi_factor_against_the_oracle():
    """params.rs:21-22: 2 (E - V(x)) in f64, with V the f64 oracle's
    potential (the GPU potential is within 1e-6 of it)."""
    (x, y, z), atoms, topo = load("ala_dipeptide")
    p = gh.Params(box=BIG, periodic=False, r_cutoff=12.0, r_switch=10.0, r_skin=14.0)
    g = gh.System(atoms.sigma, atoms.epsilon, topo, p)
    g.build_neighbor_list(x, y, z)
    g.compute_forces(x, y, z)
    pi, pj = orc.neighbor_build(x, y, z, BIG, 14.0, topo.excl_i, topo.excl_j)
    t, *_ = orc.cpu_backend_compute(x, y, z, atoms.sigma, atoms.epsilon, topo, pi, pj, BIG, 12.0, 10.0,
                                    periodic=False)
    for E in (5.12, 100.0, -3.0):
        jf = g.jacobi_factor(E)
        ref = 2.0 * (E - t.total)
        assert abs(jf - ref) <= 2e-6 * abs(t.total) + 1e-12 * abs(ref)
