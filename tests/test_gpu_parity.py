"""GPU parity: the gfx950 engine, called through the C-ABI, against the f64 oracle.

Mirrors the reference's GPU acceptance tests (geodesic-gpu/tests/
forces_gpu_vs_cpu.rs:7-24, gpu_determinism.rs:7-20; paths relative to the
reference repository root) with tighter bars:
  - Verlet list: bit-exact pair set vs neighbor.rs:25-43;
  - per-term energies: relative error <= 1e-6 (terms that are exactly zero
    in the oracle must be exactly zero);
  - forces: RMS error relative to the RMS force <= 1e-5, and every component
    within 1e-4 * max(|F|, 1) (the reference's own bar);
  - two evaluations, and two independent builds, byte-identical.
"""
import sys

import numpy as np
import pytest

import geodesic_hip as gh
from geodesic_hip import systems
from oracle import oracle as orc

from _common import force_rms_rel, load

pytestmark = pytest.mark.gpu

E_REL = 1e-6
F_RMS = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_reference(x, y, z, sigma, epsilon, topo, p: gh.Params, method="auto", threads=8):
    """Wrap (when periodic) + list + CpuBackend::compute_forces in f64."""
    x, y, z = x.copy(), y.copy(), z.copy()
    if p.periodic:
        orc.wrap_positions(x, y, z, p.box)
    pi, pj = orc.neighbor_build(x, y, z, p.box, p.r_skin, topo.excl_i, topo.excl_j, method=method)
    t, fx, fy, fz = orc.cpu_backend_compute(x, y, z, sigma, epsilon, topo, pi, pj, p.box, p.r_cutoff,
                                            p.r_switch, n_threads=threads, scale14=p.scale_14,
                                            bonded_min_image=p.bonded_min_image, periodic=p.periodic)
    return (x, y, z), (pi, pj), t, (fx, fy, fz)


def check_terms(e, t, tag):
    for name in ("bond", "angle", "dihedral", "lj", "lj14", "total"):
        g, r = getattr(e, name), getattr(t, name)
        if r == 0.0:
            assert g == 0.0, f"{tag}: {name} gpu={g} oracle=0"
        else:
            assert abs(g - r) <= E_REL * abs(r), f"{tag}: {name} gpu={g!r} oracle={r!r} rel={abs(g - r) / abs(r):.3e}"


def check_forces(f, g, tag, synthetic=False):
    """Force RMS bar, then every component within the reference's
    1e-4 * max(|F|, 1) (forces_gpu_vs_cpu.rs:17-20).  On the dense synthetic
    systems single pair forces reach ~1e4 kcal/mol/A while an atom's net force
    can cancel to ~1, so FP32 pair math leaves ~1e-3 absolute error there; those
    systems get an explicit allowance of 1e-6 x the RMS force on top."""
    fx, fy, fz = f
    gx, gy, gz = g
    rms = force_rms_rel(fx, fy, fz, gx, gy, gz)
    assert rms <= F_RMS, f"{tag}: force RMS rel {rms:.3e}"
    f_rms = float(np.sqrt(np.mean(gx ** 2 + gy ** 2 + gz ** 2)))
    extra = 1e-6 * f_rms if synthetic else 0.0
    worst = 0.0
    for a, b in ((fx, gx), (fy, gy), (fz, gz)):
        bound = 1e-4 * np.maximum(np.abs(b), 1.0) + extra
        worst = max(worst, float(np.max(np.abs(a - b) / bound)))
        assert np.all(np.abs(a - b) <= bound), f"{tag}: component bound, max {np.max(np.abs(a - b) / bound):.3e}"
    say(f"[gpu] {tag}: force rms rel {rms:.3e}, worst component / bound {worst:.3f}")
    return rms


def run_gpu(x, y, z, sigma, epsilon, topo, p, layout=None):
    s = gh.System(sigma, epsilon, topo, p)
    if layout is not None:
        s.set_pair_layout(layout)
    x, y, z = x.copy(), y.copy(), z.copy()
    s.build_neighbor_list(x, y, z)
    e, fx, fy, fz = s.compute_forces(x, y, z)
    return s, (x, y, z), e, (fx, fy, fz)


def fixture_params(n):
    """geodesic-gpu/tests/common/mod.rs:16-32."""
    return gh.Params(box=(100.0,) * 3, periodic=True, r_cutoff=12.0, r_switch=10.0, r_skin=14.0)


@pytest.mark.parametrize("name", ["lj_pair", "harmonic_dimer", "water_box_4", "ala_dipeptide"])
def test_fixture_full_backend_matches_cpu(name):
    say(f"[gpu] fixture {name}")
    (x, y, z), atoms, topo = load(name)
    p = fixture_params(len(x))
    (wx, wy, wz), (pi, pj), t, f = oracle_reference(x, y, z, atoms.sigma, atoms.epsilon, topo, p)
    s, (gx, gy, gz), e, g = run_gpu(x, y, z, atoms.sigma, atoms.epsilon, topo, p)
    assert gx.tobytes() == wx.tobytes() and gy.tobytes() == wy.tobytes() and gz.tobytes() == wz.tobytes()
    qi, qj = s.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    check_terms(e, t, name)
    check_forces(g, f, name)


def ala_dipeptide_cfg0():
    """22-atom alanine dipeptide, vacuum, coordinates + U[-0.05, 0.05] (seed 0),
    1-4 LJ scaled by 1/2, r_c/r_skin/r_sw 12/14/10 (lib.rs:31-34), box 1000."""
    (x, y, z), atoms, topo = load("ala_dipeptide")
    r = systems.Rng(0)
    x, y, z = (v + r.uniform(len(v), -0.05, 0.05) for v in (x, y, z))
    p = gh.Params(box=(1000.0,) * 3, periodic=False, r_cutoff=12.0, r_switch=10.0, r_skin=14.0,
                  scale_14=0.5)
    return (x, y, z), atoms, topo, p


def test_ala_dipeptide_vacuum_with_scaled_14_matches_cpu():
    say("[gpu] ala_dipeptide vacuum + 1-4")
    (x, y, z), atoms, topo, p = ala_dipeptide_cfg0()
    (wx, wy, wz), (pi, pj), t, f = oracle_reference(x, y, z, atoms.sigma, atoms.epsilon, topo, p)
    assert len(pi) == 133 and t.lj14 != 0.0
    s, (gx, gy, gz), e, g = run_gpu(x, y, z, atoms.sigma, atoms.epsilon, topo, p)
    assert gx.tobytes() == x.tobytes()  # non-periodic: not wrapped
    qi, qj = s.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    check_terms(e, t, "ala_dipeptide")
    check_forces(g, f, "ala_dipeptide")


@pytest.fixture(scope="module")
def sys24k():
    s = systems.system_24k()
    say("[gpu] 24k oracle reference")
    ref = oracle_reference(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params, method="cells")
    return s, ref


def test_24k_list_energies_forces_match_oracle(sys24k):
    s, ((wx, wy, wz), (pi, pj), t, f) = sys24k
    say("[gpu] 24k build + compute")
    g, (gx, gy, gz), e, fg = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params)
    assert gx.tobytes() == wx.tobytes() and gy.tobytes() == wy.tobytes() and gz.tobytes() == wz.tobytes()
    qi, qj = g.pair_list()
    assert len(qi) == len(pi)
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    st = g.stats()
    assert st["n_list_pairs"] == len(pi)
    # pair interactions counted by the build: min-image r2 <= r_c^2 in f64
    d = np.stack([wx[pj] - wx[pi], wy[pj] - wy[pi], wz[pj] - wz[pi]])
    d -= 62.0 * np.round(d / 62.0)
    assert st["n_pairs_cutoff"] == int(np.count_nonzero(np.einsum("ij,ij->j", d, d) <= 81.0))
    check_terms(e, t, "24k")
    rms = check_forces(fg, f, "24k", synthetic=True)
    say(f"[gpu] 24k force rms rel {rms:.3e} stats {st}")


def test_24k_reruns_are_byte_identical(sys24k):
    """gpu_determinism.rs:7-20, at 24k atoms, across evaluations and builds."""
    s, _ = sys24k
    say("[gpu] 24k determinism")
    g1, (x, y, z), e1, f1 = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params)
    e2, *f2 = g1.compute_forces(x, y, z)
    g2, _, e3, f3 = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params)
    for a, b in ((f1, f2), (f1, f3)):
        for u, v in zip(a, b):
            assert u.tobytes() == v.tobytes()
    for name in ("bond", "angle", "dihedral", "lj", "lj14", "total"):
        assert getattr(e1, name) == getattr(e2, name) == getattr(e3, name)


def test_24k_work_chunking_does_not_change_results(sys24k):
    """Pair forces are rounded to fixed point per block of 8 list entries
    (counted from each supercluster's first entry) and each block's energy is
    one f64 value: the work shape only decides which wavefront evaluates which
    blocks, so every chunk (1 and 7 round up to 8; 16 and more switch the i
    forces from per-block records to atomics) gives the same bytes."""
    s, _ = sys24k
    say("[gpu] 24k chunk sizes (tile layout)")
    g, (x, y, z), e0, f0 = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params, layout="tiles")
    for chunk in (1, 7, 8, 16, 24, 32, 64, 100000, 0):
        g.set_work_chunk(chunk)
        e, *f = g.compute_forces(x, y, z)
        for u, v in zip(f, f0):
            assert u.tobytes() == v.tobytes(), f"chunk {chunk}"
        for name in ("bond", "angle", "dihedral", "lj", "lj14", "total"):
            assert getattr(e, name) == getattr(e0, name), f"chunk {chunk} {name}"


def displaced(x, y, z, amp, seed):
    """Every atom moved by a random direction times U[0, amp] (A)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(3, len(x)))
    d *= amp * rng.uniform(0.0, 1.0, len(x)) / np.linalg.norm(d, axis=0)
    return x + d[0], y + d[1], z + d[2]


@pytest.mark.parametrize("layout", ["partners", "tiles"])
def test_tile_pruning_is_exact(sys24k, layout):
    """Tiles whose closest listed pair at build time is provably beyond r_c
    after the atoms' largest displacement are skipped: results are byte-
    identical with pruning off, at the build positions, after displacements
    below half the skin (0.75 A; the bound shrinks what is pruned) and past it
    (nothing can be pruned).  At 0.7 A the forces from the stale list also match
    the oracle evaluating the same list (CpuBackend between rebuilds).  With
    partner lists, pruning skips the second part of every list while the
    largest displacement is below about 0.05 A (amplitude 0.02 A)."""
    s, ((wx, wy, wz), (pi, pj), _, _) = sys24k
    say(f"[gpu] 24k pruning, {layout}")
    g, (x, y, z), _, _ = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params, layout=layout)
    p = s.params
    def forces(xs, ys, zs):
        # past half the skin, random 1.2 A moves bring some atoms of this dense
        # system within ~0.3 A: those forces are past the fixed-point range and
        # come back as NaN (GM_ERR_NUMERICAL), the same ones either way
        try:
            return g.compute_forces(xs, ys, zs)
        except gh.GmError as err:
            assert err.status == gh.GM_ERR_NUMERICAL and amp == 1.2
            return err.result

    for amp in (0.0, 0.02, 0.2, 0.7, 1.2):
        xs, ys, zs = (x, y, z) if amp == 0.0 else displaced(x, y, z, amp, seed=int(amp * 10))
        g.set_tile_pruning(True)
        e1, *f1 = forces(xs, ys, zs)
        g.set_tile_pruning(False)
        e2, *f2 = forces(xs, ys, zs)
        g.set_tile_pruning(True)
        for u, v in zip(f1, f2):
            assert u.tobytes() == v.tobytes(), f"amp {amp}"
        for name in ("bond", "angle", "dihedral", "lj", "lj14", "total"):
            a, b = getattr(e1, name), getattr(e2, name)
            assert a == b or (np.isnan(a) and np.isnan(b)), f"amp {amp} {name}"
        if amp == 0.7:
            t, *fo = orc.cpu_backend_compute(xs, ys, zs, s.sigma, s.epsilon, s.topo, pi, pj, p.box,
                                             p.r_cutoff, p.r_switch, n_threads=8, scale14=p.scale_14,
                                             bonded_min_image=p.bonded_min_image, periodic=p.periodic)
            check_terms(e1, t, "24k displaced 0.7 A")
            check_forces(f1, fo, "24k displaced 0.7 A", synthetic=True)


def test_24k_layouts_agree_and_switch_in_place(sys24k):
    """Partner lists (the default) and cluster tiles evaluate the same list:
    both within the bars against the oracle and against each other; a layout
    chosen after the build gives the bytes of a build made with it; the pair
    list and the statistics are the same under either."""
    s, (_, (pi, pj), t, f) = sys24k
    say("[gpu] 24k layouts")
    gp, (x, y, z), ep, fp = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params)
    gt, _, et, ft = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params, layout="tiles")
    sp, st = gp.stats(), gt.stats()
    assert sp["pair_layout"] == 1 and st["pair_layout"] == 0 and sp["n_lj_types"] == 16
    assert sp["n_partner_entries"] >= 2 * len(pi) and st["n_partner_entries"] == 0
    for k in ("n_list_pairs", "n_pairs_cutoff", "n_tiles", "n_entries"):
        assert sp[k] == st[k], k
    check_terms(ep, t, "24k partners")
    check_forces(fp, f, "24k partners", synthetic=True)
    check_terms(et, t, "24k tiles")
    check_forces(ft, f, "24k tiles", synthetic=True)
    check_terms(ep, et, "partners vs tiles")
    check_forces(fp, ft, "partners vs tiles", synthetic=True)
    gp.set_pair_layout("tiles")
    e1, *f1 = gp.compute_forces(x, y, z)
    gt.set_pair_layout("partners")
    e2, *f2 = gt.compute_forces(x, y, z)
    for a, b, g, h in ((e1, f1, et, ft), (e2, f2, ep, fp)):
        assert a.total == g.total and a.lj == g.lj
        assert all(u.tobytes() == v.tobytes() for u, v in zip(b, h))
    for g in (gp, gt):
        qi, qj = g.pair_list()
        assert np.array_equal(qi, pi) and np.array_equal(qj, pj)


def test_more_than_256_lj_types_use_tiles():
    """300 distinct (sigma, epsilon) pairs: past the partner kernel's LDS type
    table, so the system evaluates tiles whichever layout is asked for, and
    matches the oracle."""
    n = 3000
    s = systems.protein_solvent(n, (n / 0.1) ** (1 / 3), 300, seed=41, n_types=300)
    assert len(set(zip(s.sigma.tolist(), s.epsilon.tolist()))) > 256
    say("[gpu] 300 LJ types")
    (wx, wy, wz), (pi, pj), t, f = oracle_reference(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params,
                                                    method="brute")
    g, _, e, fg = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params, layout="partners")
    st = g.stats()
    assert st["pair_layout"] == 0 and st["n_lj_types"] == 0
    check_terms(e, t, "300 types")
    check_forces(fg, f, "300 types", synthetic=True)


@pytest.mark.parametrize("layout", ["partners", "tiles"])
def test_rectangular_periodic_box(layout):
    """A periodic box with three different edges (the fixed-point positions
    then use per-axis steps): list, energies and forces against the oracle."""
    n = 6000
    s = systems.protein_solvent(n, (n / 0.1) ** (1 / 3), 600, seed=43)
    L = s.params.box[0]
    box = (L, 1.25 * L, 1.5 * L)
    p = gh.Params(box=box, periodic=True, r_cutoff=s.params.r_cutoff, r_switch=s.params.r_switch,
                  r_skin=s.params.r_skin, scale_14=s.params.scale_14, bonded_min_image=s.params.bonded_min_image)
    x, y, z = s.x.copy(), s.y * 1.25, s.z * 1.5
    say(f"[gpu] rectangular box {box}, {layout}")
    (wx, wy, wz), (pi, pj), t, f = oracle_reference(x, y, z, s.sigma, s.epsilon, s.topo, p, method="cells")
    g, (gx, gy, gz), e, fg = run_gpu(x, y, z, s.sigma, s.epsilon, s.topo, p, layout=layout)
    assert g.stats()["pair_layout"] == (1 if layout == "partners" else 0)
    assert gx.tobytes() == wx.tobytes() and gy.tobytes() == wy.tobytes() and gz.tobytes() == wz.tobytes()
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    check_terms(e, t, f"rectangular {layout}")
    check_forces(fg, f, f"rectangular {layout}", synthetic=True)


def test_reference_bench_lattice_with_exact_ties():
    """The workload of the reference's own benchmarks (benchmarks.rs:66-135):
    10,000 atoms on a 2.0 A lattice in a 44 A box, r_c/r_skin/r_sw 5/6/4.
    Lattice distances sit exactly on thresholds: pairs at r = 6.0 A
    (r2 == r_skin^2: in the list, neighbor.rs:38 tests <=) and at r = 4.0 A
    (r == r_switch: the unswitched branch).  The list is bit-exact, forces and energies match
    the oracle; a 0.25 A displacement of every atom keeps it so on the stale list."""
    s = systems.lj_lattice(10000)
    say("[gpu] reference bench lattice 10k")
    (wx, wy, wz), (pi, pj), t, f = oracle_reference(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params,
                                                    method="brute")
    L = s.params.box[0]
    dx = wx[pj] - wx[pi]
    dy = wy[pj] - wy[pi]
    dz = wz[pj] - wz[pi]
    d2 = (dx - L * np.round(dx / L)) ** 2 + (dy - L * np.round(dy / L)) ** 2 + (dz - L * np.round(dz / L)) ** 2
    assert (d2 == 36.0).sum() > 0 and (d2 == 16.0).sum() > 0
    g, (x, y, z), e, fg = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params)
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    assert g.stats()["n_list_pairs"] == len(pi)
    check_terms(e, t, "lattice")
    check_forces(fg, f, "lattice", synthetic=True)
    xs, ys, zs = displaced(x, y, z, 0.25, seed=3)
    e2, *f2 = g.compute_forces(xs, ys, zs)
    p = s.params
    t2, *fo = orc.cpu_backend_compute(xs, ys, zs, s.sigma, s.epsilon, s.topo, pi, pj, p.box, p.r_cutoff,
                                      p.r_switch, n_threads=8, periodic=True)
    check_terms(e2, t2, "lattice displaced")
    check_forces(f2, fo, "lattice displaced", synthetic=True)


def test_needs_rebuild_matches_oracle(sys24k):
    """neighbor.rs:60-72 at half-skin 0.75 A: 0.74 A -> no rebuild, 0.76 A -> rebuild."""
    s, ((wx, wy, wz), _, _, _) = sys24k
    say("[gpu] needs_rebuild")
    g = gh.System(s.sigma, s.epsilon, s.topo, s.params)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    g.build_neighbor_list(x, y, z)
    assert not g.needs_rebuild(x, y, z)
    for d, expect in ((0.74, False), (0.76, True)):
        x2 = x.copy()
        x2[12345] += d
        assert g.needs_rebuild(x2, y, z) == expect
        assert orc.needs_rebuild(x2, y, z, wx, wy, wz, s.params.r_skin, s.params.r_cutoff) == expect


@pytest.mark.parametrize("n", [500, 926, 2000])
def test_small_periodic_boxes_use_the_single_nearest_image(n):
    """Replica-sized boxes down to L = 17.1 A < 2 r_list: still one image per pair."""
    say(f"[gpu] small box n={n}")
    s = systems.protein_solvent(n, (n / 0.1) ** (1 / 3), n // 8, seed=11 + n)
    (wx, wy, wz), (pi, pj), t, f = oracle_reference(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params,
                                                    method="brute")
    g, _, e, fg = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params)
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    check_terms(e, t, f"n={n}")
    check_forces(fg, f, f"n={n}", synthetic=True)


def test_exclusions_at_every_index_gap():
    """Exclusion partners near in index (a bit of the atom's mask of partners
    i+1..i+64) and far (a binary search), including the gaps 63, 64 and 65
    on either side of the mask's end: the list equals the oracle's."""
    say("[gpu] exclusions at near and far index gaps")
    s = systems.protein_solvent(2000, (2000 / 0.1) ** (1 / 3), 250, seed=5)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    orc.wrap_positions(x, y, z, s.params.box)
    pi, pj = orc.neighbor_build(x, y, z, s.params.box, s.params.r_skin, s.topo.excl_i, s.topo.excl_j,
                                method="brute")
    gap = pj.astype(np.int64) - pi.astype(np.int64)
    rng = np.random.default_rng(17)
    pick = []
    for g0, g1 in ((63, 64), (64, 65), (65, 66), (1, 63), (66, 2000)):
        idx = np.nonzero((gap >= g0) & (gap < g1))[0]
        pick.append(rng.choice(idx, size=min(len(idx), 300), replace=False))
    pick = np.concatenate(pick)
    assert len(pick) > 600 and {63, 64, 65} <= set(gap[pick].tolist())
    topo = s.topo
    topo.excl_i = np.concatenate([topo.excl_i, pj[pick]]).astype(np.uint32)  # either order
    topo.excl_j = np.concatenate([topo.excl_j, pi[pick]]).astype(np.uint32)
    (wx, wy, wz), (ri, rj), t, f = oracle_reference(s.x, s.y, s.z, s.sigma, s.epsilon, topo, s.params,
                                                    method="brute")
    assert len(ri) == len(pi) - len(pick)
    g, _, e, fg = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, topo, s.params)
    qi, qj = g.pair_list()
    assert np.array_equal(qi, ri) and np.array_equal(qj, rj)
    check_terms(e, t, "exclusion gaps")
    check_forces(fg, f, "exclusion gaps", synthetic=True)


def _vacuum(s, shift, box=1000.0):
    """The synthetic system moved into a large non-periodic box."""
    x, y, z = (v + d for v, d in zip((s.x, s.y, s.z), shift))
    p = gh.Params(box=(box,) * 3, periodic=False, r_cutoff=s.params.r_cutoff, r_switch=s.params.r_switch,
                  r_skin=s.params.r_skin)
    return x, y, z, p


def test_non_periodic_system_in_a_large_box():
    """A dense 6,000-atom cluster in a 1000 A non-periodic box (partly at
    negative coordinates): the fixed-point span is fitted to the atoms at each
    build.  List, energies and forces against the oracle, reruns byte-identical,
    and a stale list after displacements up to 0.35 A against the oracle on
    the same list."""
    say("[gpu] non-periodic 6000-atom cluster, 1000 A box")
    s = systems.protein_solvent(6000, (6000 / 0.1) ** (1 / 3), 60, seed=21)
    x, y, z, p = _vacuum(s, (-20.0, 350.0, -480.0))
    (wx, wy, wz), (pi, pj), t, f = oracle_reference(x, y, z, s.sigma, s.epsilon, s.topo, p, method="brute")
    g, (gx, gy, gz), e, fg = run_gpu(x, y, z, s.sigma, s.epsilon, s.topo, p)
    assert gx.tobytes() == x.tobytes()  # not wrapped
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    check_terms(e, t, "vacuum cluster")
    check_forces(fg, f, "vacuum cluster", synthetic=True)
    e2, *f2 = g.compute_forces(x, y, z)
    assert all(u.tobytes() == v.tobytes() for u, v in zip(fg, f2)) and e2.total == e.total
    rng = np.random.default_rng(4)
    d = [rng.uniform(-0.2, 0.2, len(x)) for _ in range(3)]
    x2, y2, z2 = x + d[0], y + d[1], z + d[2]
    assert not g.needs_rebuild(x2, y2, z2)
    e3, *f3 = g.compute_forces(x2, y2, z2)
    t3, *r3 = orc.cpu_backend_compute(x2, y2, z2, s.sigma, s.epsilon, s.topo, pi, pj, p.box, p.r_cutoff,
                                      p.r_switch, n_threads=8, periodic=False)
    check_terms(e3, t3, "vacuum cluster, stale list")
    check_forces(f3, r3, "vacuum cluster, stale list", synthetic=True)


def test_non_periodic_atoms_spread_over_the_box():
    """Two clusters 600 A apart in a 1000 A non-periodic box: no cube of twice
    the extent fits, so the span is the box; the reference's min image with
    the box then applies to every pair, as in the oracle."""
    say("[gpu] non-periodic, two clusters 600 A apart")
    a = systems.protein_solvent(1000, (1000 / 0.1) ** (1 / 3), 100, seed=8)
    b = systems.protein_solvent(1000, (1000 / 0.1) ** (1 / 3), 100, seed=9)
    n = 1000
    x = np.concatenate([a.x + 10.0, b.x + 610.0])
    y = np.concatenate([a.y - 5.0, b.y + 20.0])
    z = np.concatenate([a.z, b.z + 3.0])
    sig = np.concatenate([a.sigma, b.sigma])
    eps = np.concatenate([a.epsilon, b.epsilon])
    ta, tb = a.topo, b.topo
    cat = lambda u, v: np.concatenate([u, v + n]).astype(np.uint32)  # noqa: E731
    topo = systems.Topology(cat(ta.bond_i, tb.bond_i), cat(ta.bond_j, tb.bond_j),
                            np.concatenate([ta.bond_k, tb.bond_k]), np.concatenate([ta.bond_r0, tb.bond_r0]),
                            cat(ta.angle_i, tb.angle_i), cat(ta.angle_j, tb.angle_j), cat(ta.angle_k, tb.angle_k),
                            np.concatenate([ta.angle_kth, tb.angle_kth]),
                            np.concatenate([ta.angle_th0, tb.angle_th0]),
                            cat(ta.dihed_i, tb.dihed_i), cat(ta.dihed_j, tb.dihed_j), cat(ta.dihed_k, tb.dihed_k),
                            cat(ta.dihed_l, tb.dihed_l), np.concatenate([ta.dihed_kphi, tb.dihed_kphi]),
                            np.concatenate([ta.dihed_n, tb.dihed_n]).astype(np.uint32),
                            np.concatenate([ta.dihed_delta, tb.dihed_delta]),
                            cat(ta.excl_i, tb.excl_i), cat(ta.excl_j, tb.excl_j))
    p = gh.Params(box=(1000.0,) * 3, periodic=False, r_cutoff=9.0, r_switch=7.5, r_skin=10.5)
    (wx, wy, wz), (pi, pj), t, f = oracle_reference(x, y, z, sig, eps, topo, p, method="brute")
    g, _, e, fg = run_gpu(x, y, z, sig, eps, topo, p)
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    assert np.all((pi < n) == (pj < n))  # no pair across the clusters
    check_terms(e, t, "two clusters")
    check_forces(fg, f, "two clusters", synthetic=True)


def test_empty_and_single_atom_systems():
    say("[gpu] empty / single")
    z32, zf = np.zeros(0, np.uint32), np.zeros(0)
    topo = systems.Topology(z32, z32, zf, zf, z32, z32, z32, zf, zf, z32, z32, z32, z32, zf, z32, zf, z32, z32)
    p = gh.Params(box=(20.0,) * 3, periodic=True, r_cutoff=9.0, r_switch=7.5, r_skin=10.5)
    g = gh.System(zf, zf, topo, p)
    g.build_neighbor_list(zf.copy(), zf.copy(), zf.copy())
    e, fx, fy, fz = g.compute_forces(zf, zf, zf)
    assert e.total == 0.0 and len(fx) == 0
    g1 = gh.System(np.array([3.0]), np.array([0.1]), topo, p)
    x, y, z = np.array([25.0]), np.array([-1.0]), np.array([3.0])
    g1.build_neighbor_list(x, y, z)
    assert x[0] == 5.0 and y[0] == 19.0
    e, fx, fy, fz = g1.compute_forces(x, y, z)
    assert e.total == 0.0 and fx[0] == 0.0


def test_all_pairs_excluded_gives_empty_list():
    say("[gpu] all excluded")
    (x, y, z), atoms, topo = load("water_box_4")
    ii, jj = np.triu_indices(12, 1)
    topo.excl_i, topo.excl_j = ii.astype(np.uint32), jj.astype(np.uint32)
    g, _, e, f = run_gpu(x, y, z, atoms.sigma, atoms.epsilon, topo, fixture_params(12))
    qi, _ = g.pair_list()
    assert len(qi) == 0 and e.lj == 0.0


def test_errors_are_reported_not_crashed():
    say("[gpu] error paths")
    (x, y, z), atoms, topo = load("water_box_4")
    g = gh.System(atoms.sigma, atoms.epsilon, topo, fixture_params(12))
    with pytest.raises(gh.GmError) as ei:
        g.compute_forces(x, y, z)
    assert ei.value.status == 7  # GM_ERR_NOT_BUILT
    bad = gh.Params(box=(100.0,) * 3, periodic=True, r_cutoff=12.0, r_switch=13.0, r_skin=14.0)
    with pytest.raises(gh.GmError) as ei:
        gh.System(atoms.sigma, atoms.epsilon, topo, bad)
    assert ei.value.status == 5  # GM_ERR_INVALID_ARG: r_switch < r_cutoff violated
    topo.bond_j = topo.bond_j.copy()
    topo.bond_j[0] = 99
    with pytest.raises(gh.GmError) as ei:
        gh.System(atoms.sigma, atoms.epsilon, topo, fixture_params(12))
    assert ei.value.status == 5


def test_device_resident_path_equals_host_path(sys24k):
    """gm_compute_forces_device on caller-owned device arrays, and the
    resident entry points, give the host path's bytes."""
    s, _ = sys24k
    say("[gpu] device-resident 24k")
    g, (x, y, z), e, f = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params)
    dx, dy, dz = (gh.DeviceArray.from_host(v) for v in (x, y, z))
    out = [gh.DeviceArray(len(x)) for _ in range(3)]
    g.compute_forces_device(dx.ptr, dy.ptr, dz.ptr, *[o.ptr for o in out])
    e2 = g.sync()
    for a, b in zip(out, f):
        assert a.to_host().tobytes() == b.tobytes()
    assert e2.total == e.total
    g.set_positions(x, y, z)
    g.compute_forces_resident()
    e3 = g.sync()
    for a, b in zip(g.get_forces(), f):
        assert a.tobytes() == b.tobytes()
    assert e3.total == e.total


def test_builds_from_alternating_device_arrays(sys24k):
    """Steady-state builds from two caller-owned sets of device arrays used in
    turn (each build's launch arguments differ from the previous one's), and
    then from one set again: every list and every evaluation is byte-identical
    to the host-array build, and the caller's arrays hold the wrapped
    positions."""
    s, _ = sys24k
    say("[gpu] builds from alternating device arrays, 24k")
    g, (x, y, z), e, f = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params)
    pi, pj = g.pair_list()
    sets = [[gh.DeviceArray.from_host(v) for v in (s.x, s.y, s.z)] for _ in range(2)]
    out = [gh.DeviceArray(len(x)) for _ in range(3)]
    for k in (0, 1, 0, 1, 0, 0, 0):
        d = sets[k]
        g.build_neighbor_list_device(*(a.ptr for a in d))
        for a, w in zip(d, (x, y, z)):
            assert a.to_host().tobytes() == w.tobytes()
        qi, qj = g.pair_list()
        assert np.array_equal(qi, pi) and np.array_equal(qj, pj), k
        g.compute_forces_device(*(a.ptr for a in d), *(o.ptr for o in out))
        e2 = g.sync()
        for a, b in zip(out, f):
            assert a.to_host().tobytes() == b.tobytes(), k
        assert e2.total == e.total


def test_rebuild_past_the_entry_capacity():
    """A rebuild whose list needs more entries than the buffers the first build
    sized (a quarter more than its own entries, plus 1024): the atoms start
    spread over a box 1.5 times wider, then return to their normal density in
    one corner of it (20,000 atoms: the xy columns, sized for the mean
    density, are then fine enough for the corner to hold several per axis).  The build finds its entry count past the capacity, sizes
    the buffers anew and runs again: list, energies and forces against the
    oracle, and a second build at the same positions byte-identical."""
    s = systems.protein_solvent(20000, (20000 / 0.1) ** (1 / 3), 2000, seed=31)
    L = s.params.box[0] * 1.5
    p = gh.Params(box=(L,) * 3, periodic=True, r_cutoff=s.params.r_cutoff, r_switch=s.params.r_switch,
                  r_skin=s.params.r_skin, scale_14=s.params.scale_14, bonded_min_image=s.params.bonded_min_image)
    say("[gpu] rebuild past the entry capacity")
    g = gh.System(s.sigma, s.epsilon, s.topo, p)
    x1, y1, z1 = s.x * 1.5, s.y * 1.5, s.z * 1.5
    g.build_neighbor_list(x1, y1, z1)
    ent1 = g.stats()["n_entries"]
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    g.build_neighbor_list(x, y, z)
    ent2 = g.stats()["n_entries"]
    assert ent2 > ent1 + ent1 // 4 + 1024, (ent1, ent2)  # the capacity of the first build was exceeded
    (wx, wy, wz), (pi, pj), t, f = oracle_reference(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, p, method="cells")
    assert x.tobytes() == wx.tobytes()
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    e, *fg = g.compute_forces(x, y, z)
    check_terms(e, t, "grown list")
    check_forces(fg, f, "grown list", synthetic=True)
    g.build_neighbor_list(x, y, z)
    e2, *fg2 = g.compute_forces(x, y, z)
    assert e2.total == e.total and all(a.tobytes() == b.tobytes() for a, b in zip(fg, fg2))


@pytest.fixture(scope="module")
def sys250k():
    s = systems.system_250k()
    say("[gpu] 250k oracle reference (cells)")
    ref = oracle_reference(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params, method="cells", threads=16)
    return s, ref


@pytest.mark.parametrize("layout", ["auto", "partners"])
def test_250k_full_size_properties(sys250k, layout):
    """At the full 250k size: the list equals the oracle's, energies and forces
    within the bars, and reruns byte-identical.  Its neighbouring cells lie
    more than 32,767 slots apart: the automatic layout evaluates tiles, and
    partner lists (asked for) hold 32-bit slot numbers."""
    s, ((wx, wy, wz), (pi, pj), t, f) = sys250k
    say(f"[gpu] 250k build + compute, {layout}")
    g, (x, y, z), e, fg = run_gpu(s.x, s.y, s.z, s.sigma, s.epsilon, s.topo, s.params, layout=layout)
    assert g.stats()["pair_layout"] == (0 if layout == "auto" else 1)
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    check_terms(e, t, f"250k {layout}")
    check_forces(fg, f, f"250k {layout}", synthetic=True)
    e2, *f2 = g.compute_forces(x, y, z)
    for u, v in zip(fg, f2):
        assert u.tobytes() == v.tobytes()
