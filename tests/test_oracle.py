"""Pins the f64 CPU oracle against what the reference itself fixes.

Known answers and invariants come from the reference's own tests (cited per
test, paths relative to the reference repository root), its golden DCD, and
closed forms.  None of these re-derives a value from the oracle's formula.
"""
import numpy as np
import pytest

from oracle import amber
from oracle import oracle as orc

from _common import FIX, load

BIG = (1000.0, 1000.0, 1000.0)


def empty_topology():
    z32, zf = np.zeros(0, np.uint32), np.zeros(0, np.float64)
    return amber.Topology(z32, z32, zf, zf, z32, z32, z32, zf, zf, z32, z32, z32, z32, zf, z32, zf,
                          z32, z32)


# ---- golden trajectory: geodesic/tests/golden_reference.rs:27-94 -------------

def test_ala_dipeptide_100_frame_trajectory_matches_golden():
    """The oracle's bonded + LJ + BAB/RATTLE restatement reproduces the
    reference's committed DCD byte for byte (n_threads=1, dt 1 fs, non-periodic,
    box 1000, r_c/r_skin/r_sw 12/14/10, H-bonds constrained, tol 1e-8)."""
    with open(f"{FIX}/ala_dipeptide_ref.dcd", "rb") as f:
        hdr, golden = amber.read_dcd(f.read())
    assert hdr["nset"] == 100 and hdr["natom"] == 22
    (x, y, z), atoms, topo = load("ala_dipeptide")
    vx, vy, vz = np.zeros(22), np.zeros(22), np.zeros(22)
    amber.promote_hydrogen_bonds(topo, atoms)
    rc, rs, rsw, dt, mi, tol = 12.0, 14.0, 10.0, 0.001, 100, 1e-8

    def build():
        pi, pj = orc.neighbor_build(x, y, z, BIG, rs, topo.excl_i, topo.excl_j)
        return pi, pj, x.copy(), y.copy(), z.copy()

    def forces(pi, pj):
        t, fx, fy, fz = orc.cpu_backend_compute(x, y, z, atoms.sigma, atoms.epsilon, topo, pi, pj,
                                                BIG, rc, rsw, n_threads=1, periodic=False)
        return t, fx, fy, fz

    pi, pj, rx, ry, rz = build()
    t, fx, fy, fz = forces(pi, pj)
    assert orc.constrain_velocities(x, y, z, vx, vy, vz, topo, atoms.mass, mi, tol) >= 0
    frames = [np.stack([x, y, z]).astype(np.float32)]
    for _ in range(99):
        orc.half_kick(vx, vy, vz, fx, fy, fz, atoms.mass, dt / 2)
        assert orc.drift_and_constrain(x, y, z, vx, vy, vz, topo, atoms.mass, dt, mi, tol) >= 0
        if orc.needs_rebuild(x, y, z, rx, ry, rz, rs, rc):
            pi, pj, rx, ry, rz = build()
        t, fx, fy, fz = forces(pi, pj)
        orc.half_kick(vx, vy, vz, fx, fy, fz, atoms.mass, dt / 2)
        assert orc.constrain_velocities(x, y, z, vx, vy, vz, topo, atoms.mass, mi, tol) >= 0
        frames.append(np.stack([x, y, z]).astype(np.float32))
    frames = np.array(frames)
    assert frames.tobytes() == golden.tobytes()


def test_ala_dipeptide_energy_matches_golden_total():
    """geodesic/tests/golden_reference.rs:57 records total_energy = 5.12 for
    this fixture (the number `geodesic energy` prints, lib.rs:113-171)."""
    (x, y, z), atoms, topo = load("ala_dipeptide")
    pi, pj = orc.neighbor_build(x, y, z, BIG, 14.0, topo.excl_i, topo.excl_j)
    assert len(pi) == 231 - 98
    t, *_ = orc.cpu_backend_compute(x, y, z, atoms.sigma, atoms.epsilon, topo, pi, pj, BIG, 12.0,
                                    10.0, periodic=False)
    assert round(t.total, 2) == 5.12


# ---- prmtop known answers: geodesic-io/tests/prmtop_roundtrip.rs:4-29 ---------

def test_lj_pair_atom_data():
    (x, y, z), atoms, topo = load("lj_pair")
    assert list(atoms.mass) == [1.0, 2.0]
    assert abs(atoms.sigma[0] - 2.0) < 1e-8 and abs(atoms.epsilon[0] - 1.0) < 1e-8
    assert abs(atoms.sigma[1] - 2.0) < 1e-8 and abs(atoms.epsilon[1] - 1.0) < 1e-8
    assert atoms.names == ["AR1 ", "AR2 "] and atoms.element[0] == amber.ELEM_UNKNOWN
    assert len(topo.bond_i) == len(topo.angle_i) == len(topo.dihed_i) == len(topo.excl_i) == 0


def test_tip3p_oxygen_lj_parameters():
    """TIP3P oxygen (water_box_4): sigma 3.15061 A, epsilon 0.1521 kcal/mol."""
    _, atoms, _ = load("water_box_4")
    o = [i for i, e in enumerate(atoms.element) if e == amber.ELEM_O]
    assert len(o) == 4
    assert abs(atoms.sigma[o[0]] - 3.15061) < 1e-4
    assert abs(atoms.epsilon[o[0]] - 0.1521) < 1e-4


def test_ala_dipeptide_topology_counts():
    _, atoms, topo = load("ala_dipeptide")
    assert len(atoms.mass) == 22
    assert (len(topo.bond_i), len(topo.angle_i), len(topo.dihed_i)) == (21, 36, 52)
    assert len(topo.excl_i) == 98
    # 41 unique proper 1-4 pairs, all of them inside the reference's exclusions
    assert len(topo.pair14_i) == 41
    ex = set(zip(np.minimum(topo.excl_i, topo.excl_j).tolist(), np.maximum(topo.excl_i, topo.excl_j).tolist()))
    assert all((a, b) in ex for a, b in zip(topo.pair14_i.tolist(), topo.pair14_j.tolist()))


# ---- LJ closed forms and invariants: nonbonded_gradient.rs ------------------

def two_atom(r):
    x = np.array([0.0, r]); y = np.zeros(2); z = np.zeros(2)
    return x, y, z, np.array([2.0, 2.0]), np.array([1.0, 1.0])


def potential_at(r, rc, rsw):
    x, y, z, s, e = two_atom(r)
    return orc.pair_forces(x, y, z, s, e, [0], [1], rc, rsw, BIG)[0]


def force_x1(r, rc, rsw):
    x, y, z, s, e = two_atom(r)
    return orc.pair_forces(x, y, z, s, e, [0], [1], rc, rsw, BIG)[1][1]


@pytest.mark.parametrize("r,v,f", [
    (3.0, -3.203365942785745e-1, -5.790144155230776e-1),
    (2.5, -7.736980930560002e-1, -1.1971684466688e0),
    (10.5, -1.712482547001078e-4, -1.985857393975714e-4),
    (11.5, -1.145636429712952e-5, -6.433967872816221e-5),
])
def test_switched_lj_closed_form(r, v, f):
    """sigma=2, eps=1, r_c=12, r_sw=10: V and the radial force F on atom 1."""
    assert abs(potential_at(r, 12.0, 10.0) - v) <= 1e-12 * abs(v)
    assert abs(force_x1(r, 12.0, 10.0) - f) <= 1e-12 * abs(f)


@pytest.mark.parametrize("r", [2.5, 3.0, 10.5, 11.5])
def test_gradient_matches_finite_difference(r):
    eps = 1e-6
    num = -(potential_at(r + eps, 12.0, 10.0) - potential_at(r - eps, 12.0, 10.0)) / (2 * eps)
    ana = force_x1(r, 12.0, 10.0)
    assert abs(ana - num) / (abs(num) + 1.0) < 1e-5


def test_switching_function_is_continuous_at_boundaries():
    assert abs(potential_at(10.0 - 1e-7, 12.0, 10.0) - potential_at(10.0 + 1e-7, 12.0, 10.0)) < 1e-4
    assert abs(potential_at(12.0 - 1e-6, 12.0, 10.0)) < 1e-3
    assert potential_at(12.0 + 1e-9, 12.0, 10.0) == 0.0


def test_repulsive_force_pushes_atoms_apart():
    assert force_x1(1.5, 12.0, 10.0) > 0.0


def test_newtons_third_law_holds():
    x, y, z, s, e = two_atom(3.0)
    _, fx, _, _ = orc.pair_forces(x, y, z, s, e, [0], [1], 12.0, 10.0, BIG)
    assert abs(fx[0] + fx[1]) < 1e-12


def test_coincident_atoms_are_skipped():
    """nonbonded.rs:129: r2 == 0 contributes nothing (no NaN)."""
    x, y, z, s, e = two_atom(0.0)
    en, fx, fy, fz = orc.pair_forces(x, y, z, s, e, [0], [1], 12.0, 10.0, BIG)
    assert en == 0.0 and not np.any(fx) and np.all(np.isfinite(fx))


# ---- bonded gradients: bonded_gradient.rs:32-186 ----------------------------

def fd_check(x, y, z, fn, tol=1e-4):
    _, fx, fy, fz = fn(x, y, z)
    eps = 1e-6
    for i in range(len(x)):
        for arr, ana in ((x, fx[i]), (y, fy[i]), (z, fz[i])):
            keep = arr[i]
            arr[i] = keep + eps
            vp = fn(x, y, z)[0]
            arr[i] = keep - eps
            vm = fn(x, y, z)[0]
            arr[i] = keep
            num = -(vp - vm) / (2 * eps)
            assert abs(ana - num) / (abs(num) + 1.0) < tol, (i, ana, num)


def test_bond_gradient_matches_finite_difference():
    t = empty_topology()
    t.bond_i, t.bond_j = np.array([0], np.uint32), np.array([1], np.uint32)
    t.bond_k, t.bond_r0 = np.array([300.0]), np.array([1.5])
    fd_check(np.array([0.0, 1.7]), np.zeros(2), np.zeros(2), lambda x, y, z: orc.bond_forces(x, y, z, t))


def test_angle_gradient_matches_finite_difference():
    t = empty_topology()
    t.angle_i, t.angle_j, t.angle_k = (np.array([v], np.uint32) for v in (0, 1, 2))
    t.angle_kth, t.angle_th0 = np.array([50.0]), np.array([1.9106])
    fd_check(np.array([1.0, 0.0, 0.3]), np.array([0.0, 0.0, 1.0]), np.zeros(3),
             lambda x, y, z: orc.angle_forces(x, y, z, t))


def dihedral_topo(k, n, d):
    t = empty_topology()
    t.dihed_i, t.dihed_j, t.dihed_k, t.dihed_l = (np.array([v], np.uint32) for v in (0, 1, 2, 3))
    t.dihed_kphi, t.dihed_n, t.dihed_delta = np.array([k]), np.array([n], np.uint32), np.array([d])
    return t


@pytest.mark.parametrize("geom,par", [
    (([0.0, 0.0, 1.0, 1.5], [0.0, 1.0, 1.0, 2.0], [0.0, 0.0, 0.0, 0.7]), (2.0, 2, 0.0)),
    (([0.3, -0.4, 0.9, 2.1], [-0.6, 0.5, 1.3, 0.8], [1.1, 0.2, -0.3, 1.4]), (3.5, 3, 0.7)),
])
def test_dihedral_gradient_matches_finite_difference(geom, par):
    t = dihedral_topo(*par)
    x, y, z = (np.array(v) for v in geom)
    fd_check(x, y, z, lambda x, y, z: orc.dihedral_forces(x, y, z, t))


def test_dihedral_total_force_sums_to_zero():
    t = dihedral_topo(2.0, 2, 0.0)
    _, fx, fy, fz = orc.dihedral_forces(np.array([0.0, 0.0, 1.0, 1.5]), np.array([0.0, 1.0, 1.0, 2.0]),
                                        np.array([0.0, 0.0, 0.0, 0.7]), t)
    assert abs(fx.sum()) < 1e-10 and abs(fy.sum()) < 1e-10 and abs(fz.sum()) < 1e-10


# ---- fixtures: fixture_gradient_check.rs, newton_third_law.rs ---------------

FIXTURES = ["lj_pair", "harmonic_dimer", "water_box_4", "ala_dipeptide"]


def total_potential_and_force(x, y, z, atoms, topo, box=(100.0,) * 3):
    e1, fx, fy, fz = orc.bond_forces(x, y, z, topo)
    e2, *_ = orc.angle_forces(x, y, z, topo, fx=fx, fy=fy, fz=fz)
    e3, *_ = orc.dihedral_forces(x, y, z, topo, fx=fx, fy=fy, fz=fz)
    wx, wy, wz = x.copy(), y.copy(), z.copy()
    orc.wrap_positions(wx, wy, wz, box)
    pi, pj = orc.neighbor_build(wx, wy, wz, box, 14.0, topo.excl_i, topo.excl_j)
    e4, *_ = orc.pair_forces(x, y, z, atoms.sigma, atoms.epsilon, pi, pj, 12.0, 10.0, box, fx, fy, fz)
    return e1 + e2 + e3 + e4, fx, fy, fz


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_gradient_matches_finite_difference(name):
    (x, y, z), atoms, topo = load(name)
    fd_check(x, y, z, lambda x, y, z: total_potential_and_force(x, y, z, atoms, topo))


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_net_force_vanishes(name):
    (x, y, z), atoms, topo = load(name)
    _, fx, fy, fz = total_potential_and_force(x, y, z, atoms, topo)
    n_eps = len(x) * 1e-10
    assert abs(fx.sum()) < n_eps and abs(fy.sum()) < n_eps and abs(fz.sum()) < n_eps


# ---- neighbour list: neighbor_list.rs:48-109 ---------------------------------

def test_finds_pairs_within_skin_and_excludes_far_pairs():
    pi, pj = orc.neighbor_build([0.0, 3.0, 20.0], [0.0] * 3, [0.0] * 3, BIG, 6.0, [], [])
    assert list(pi) == [0] and list(pj) == [1]


def test_bonded_exclusions_are_filtered_out():
    pi, _ = orc.neighbor_build([0.0, 1.0], [0.0] * 2, [0.0] * 2, BIG, 6.0, [0], [1])
    assert len(pi) == 0


def test_minimum_image_finds_pair_across_periodic_boundary():
    pi, pj = orc.neighbor_build([0.5, 9.5], [0.0] * 2, [0.0] * 2, (10.0,) * 3, 6.0, [], [])
    assert list(pi) == [0] and list(pj) == [1]


def test_positions_are_wrapped_into_box():
    x, y, z = np.array([-1.0]), np.array([11.0]), np.array([0.0])
    orc.wrap_positions(x, y, z, (10.0,) * 3)
    assert abs(x[0] - 9.0) < 1e-12 and abs(y[0] - 1.0) < 1e-12


def test_rebuild_triggered_past_half_skin_displacement():
    z1 = np.zeros(1)
    assert not orc.needs_rebuild([0.4], z1, z1, z1, z1, z1, 6.0, 5.0)
    assert orc.needs_rebuild([0.6], z1, z1, z1, z1, z1, 6.0, 5.0)


@pytest.mark.parametrize("n,box,seed", [(900, 17.0, 1), (3000, 31.0, 2), (2000, 40.0, 3), (500, 9.0, 4)])
def test_cell_build_equals_brute_force(n, box, seed):
    """The cell-grid oracle gives the brute-force list exactly (order and set),
    including boxes with fewer than 3 cells per side."""
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(0, box, n) for _ in range(3))
    ei = rng.integers(0, n, 200).astype(np.uint32)
    ej = rng.integers(0, n, 200).astype(np.uint32)
    a = orc.neighbor_build(x, y, z, (box,) * 3, 10.5, ei, ej, method="brute")
    b = orc.neighbor_build(x, y, z, (box,) * 3, 10.5, ei, ej, method="cells")
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---- CPU backend determinism: geodesic-engine/tests/cpu_backend.rs ------------

def lj_cluster():
    c = [(0, 0, 0), (3, 0, 0), (0, 3, 0), (3, 3, 0), (0, 0, 3), (3, 0, 3), (0, 3, 3), (3, 3, 3),
         (1.5, 1.5, 1.5), (4.5, 1.5, 1.5), (1.5, 4.5, 1.5), (1.5, 1.5, 4.5)]
    a = np.array(c, dtype=np.float64)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy(), np.full(12, 1.5), np.full(12, 0.2)


def cluster_backend(T):
    x, y, z, s, e = lj_cluster()
    box = (100.0,) * 3
    orc.wrap_positions(x, y, z, box)
    pi, pj = orc.neighbor_build(x, y, z, box, 10.0, [], [])
    return orc.cpu_backend_compute(x, y, z, s, e, empty_topology(), pi, pj, box, 8.0, 6.0, n_threads=T), \
        (x, y, z, s, e, pi, pj)


def test_compute_forces_matches_direct_nonbonded_call_single_thread():
    (t, fx, fy, fz), (x, y, z, s, e, pi, pj) = cluster_backend(1)
    en, gx, gy, gz = orc.pair_forces(x, y, z, s, e, pi, pj, 8.0, 6.0, (100.0,) * 3)
    assert np.all(np.abs(fx - gx) < 1e-12) and np.all(np.abs(fy - gy) < 1e-12)
    assert np.all(np.abs(fz - gz) < 1e-12) and t.lj == en


@pytest.mark.parametrize("T", [1, 4])
def test_nonbonded_reduction_is_repeatable_for_a_fixed_thread_count(T):
    (_, a, b, c), _ = cluster_backend(T)
    (_, d, e, f), _ = cluster_backend(T)
    assert a.tobytes() == d.tobytes() and b.tobytes() == e.tobytes() and c.tobytes() == f.tobytes()


def test_nonbonded_reduction_agrees_across_thread_counts_within_tolerance():
    (_, a, b, c), _ = cluster_backend(1)
    (_, d, e, f), _ = cluster_backend(4)
    assert np.all(np.abs(a - d) < 1e-10) and np.all(np.abs(b - e) < 1e-10) and np.all(np.abs(c - f) < 1e-10)


# ---- extensions with no reference counterpart (closed forms / invariants) ----

def test_lj14_is_half_the_unswitched_pair_energy():
    """1-4 term = 0.5 * V_LJ(r) and 0.5 * F_LJ(r), pinned against the
    golden-pinned pair routine in its unswitched range."""
    for r in (2.2, 3.1, 3.9):
        x, y, z, s, e = two_atom(r)
        en, fx, _, _ = orc.lj14_forces(x, y, z, s, e, [0], [1], 0.5, BIG)
        ep, gx, _, _ = orc.pair_forces(x, y, z, s, e, [0], [1], 12.0, 10.0, BIG)
        assert abs(en - 0.5 * ep) <= 1e-15 * abs(ep) and abs(fx[1] - 0.5 * gx[1]) <= 1e-15 * abs(gx[1])


def test_lj14_gradient_matches_finite_difference():
    (x, y, z), atoms, topo = load("ala_dipeptide")
    fd_check(x, y, z, lambda x, y, z: orc.lj14_forces(x, y, z, atoms.sigma, atoms.epsilon,
                                                      topo.pair14_i, topo.pair14_j, 0.5, BIG))


def test_min_image_bonded_terms_equal_raw_terms_for_a_whole_molecule():
    (x, y, z), atoms, topo = load("ala_dipeptide")
    box = (40.0,) * 3
    for fn in (orc.bond_forces, orc.angle_forces, orc.dihedral_forces):
        a = fn(x, y, z, topo)
        b = fn(x, y, z, topo, min_image=True, box=box)
        assert a[0] == b[0] and np.array_equal(a[1], b[1])


def test_min_image_bonded_terms_are_invariant_across_the_boundary():
    """Shift the molecule so it straddles the box face and wrap every atom:
    min-image bonded energies and forces match the unwrapped molecule."""
    (x, y, z), atoms, topo = load("ala_dipeptide")
    box = (40.0,) * 3
    sx, sy, sz = x - 3.5, y + 37.0, z - 1.0
    orc.wrap_positions(sx, sy, sz, box)
    for fn in (orc.bond_forces, orc.angle_forces, orc.dihedral_forces):
        a = fn(x, y, z, topo)
        b = fn(sx, sy, sz, topo, min_image=True, box=box)
        assert abs(a[0] - b[0]) <= 1e-9 * max(1.0, abs(a[0]))
        assert np.max(np.abs(a[1] - b[1])) < 1e-7 and np.max(np.abs(a[3] - b[3])) < 1e-7
