"""Batched replicas on the GPU (no reference counterpart): every replica of a
batch against its own f64 oracle, against the same replica evaluated alone
(byte-identical), across graph replays, and across work shapes; in both
batch layouts (the default, partner lists of every replica, and tiles)."""
import sys

import numpy as np
import pytest

import geodesic_hip as gh
from geodesic_hip import systems
from oracle import oracle as orc

from _common import force_rms_rel

pytestmark = pytest.mark.gpu

SIZES = [500, 926, 1500, 2400, 3100, 5200]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_replicas(sizes, seed=21):
    reps = []
    for k, n in enumerate(sizes):
        reps.append(systems.protein_solvent(int(n), (n / 0.1) ** (1 / 3), int(n) // 8, seed=seed + k))
    return reps


def concat(reps):
    return (np.concatenate([r.x for r in reps]), np.concatenate([r.y for r in reps]),
            np.concatenate([r.z for r in reps]))


# the batch's layout setting -> the layout its replicas then evaluate
LAYOUT = {"auto": "partners", "tiles": "tiles"}


@pytest.fixture(scope="module", params=["auto", "tiles"])
def batch_case(request):
    reps = make_replicas(SIZES)
    say(f"[gpu] batch build, layout {request.param}")
    b = gh.Batch([(r.sigma, r.epsilon, r.topo, r.params) for r in reps])
    b.set_pair_layout(request.param)
    x, y, z = concat(reps)
    b.build_neighbor_lists(x, y, z)
    assert b.pair_layout() == LAYOUT[request.param]
    b.compute()
    e, ms = b.sync()
    f = b.get_forces()
    return reps, b, (x, y, z), e, f


def test_each_replica_matches_its_oracle(batch_case):
    reps, b, (x, y, z), e, (fx, fy, fz) = batch_case
    for k, r in enumerate(reps):
        say(f"[gpu] batch replica {k} n={r.n}")
        o0, o1 = b.offsets[k], b.offsets[k + 1]
        wx, wy, wz = r.x.copy(), r.y.copy(), r.z.copy()
        orc.wrap_positions(wx, wy, wz, r.params.box)
        assert x[o0:o1].tobytes() == wx.tobytes()
        pi, pj = orc.neighbor_build(wx, wy, wz, r.params.box, r.params.r_skin, r.topo.excl_i, r.topo.excl_j,
                                    method="brute")
        qi, qj = b.replica(k).pair_list()
        assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
        t, gx, gy, gz = orc.cpu_backend_compute(wx, wy, wz, r.sigma, r.epsilon, r.topo, pi, pj, r.params.box,
                                                r.params.r_cutoff, r.params.r_switch, n_threads=8,
                                                scale14=r.params.scale_14, bonded_min_image=True)
        for name in ("bond", "angle", "dihedral", "lj", "lj14", "total"):
            g, ref = getattr(e[k], name), getattr(t, name)
            assert abs(g - ref) <= 1e-6 * abs(ref), (k, name, g, ref)
        assert force_rms_rel(fx[o0:o1], fy[o0:o1], fz[o0:o1], gx, gy, gz) <= 1e-5


def test_batch_equals_each_replica_alone(batch_case):
    """A plain system in the layout the batch evaluates: same lists, same
    arithmetic, forces and energies byte-identical."""
    reps, b, _, e, (fx, fy, fz) = batch_case
    for k, r in enumerate(reps):
        say(f"[gpu] plain replica {k}, {b.pair_layout()}")
        g = gh.System(r.sigma, r.epsilon, r.topo, r.params)
        g.set_pair_layout(b.pair_layout())
        x, y, z = r.x.copy(), r.y.copy(), r.z.copy()
        g.build_neighbor_list(x, y, z)
        ep, px, py, pz = g.compute_forces(x, y, z)
        o0, o1 = b.offsets[k], b.offsets[k + 1]
        assert px.tobytes() == fx[o0:o1].tobytes()
        assert py.tobytes() == fy[o0:o1].tobytes() and pz.tobytes() == fz[o0:o1].tobytes()
        assert ep.total == e[k].total and ep.lj == e[k].lj


def test_graph_replays_are_byte_identical(batch_case):
    """Replays of the captured graph, with eager copies and syncs between them,
    reproduce the first evaluation exactly."""
    reps, b, (x, y, z), e0, f0 = batch_case
    for it in range(3):
        say(f"[gpu] graph replay {it}")
        b.set_positions(x, y, z)
        b.compute()
        e, ms = b.sync()
        f = b.get_forces()
        assert ms > 0.0
        for u, v in zip(f, f0):
            assert u.tobytes() == v.tobytes()
        assert [t.total for t in e] == [t.total for t in e0]


@pytest.mark.parametrize("entries,waves", [(6, 8), (8, 20), (16, 12), (32, 16)])
def test_work_shape_does_not_change_results(batch_case, entries, waves):
    """Tiles: work items of 8 entries (6 rounds up) leave i forces as per-block
    records, 16 and 32 as atomics (and the finish kernel then takes one
    wavefront per supercluster); the persistent grid varies too.  Forces are
    rounded per block of 8 entries either way: every replica's bytes equal
    the automatic shape's, and the replica evaluated alone.  Partner lists
    have no shape: the same calls leave every byte as it was."""
    reps, b, (x, y, z), e0, f0 = batch_case
    say(f"[gpu] batch shape {entries} entries per item, {waves} waves per CU")
    b.set_shape(entries, waves)
    b.set_positions(x, y, z)
    b.compute()
    e, _ = b.sync()
    f = b.get_forces()
    for u, v in zip(f, f0):
        assert u.tobytes() == v.tobytes()
    assert [t.total for t in e] == [t.total for t in e0] and [t.lj for t in e] == [t.lj for t in e0]
    fx, fy, fz = f
    for k, r in enumerate(reps[:2]):
        g = gh.System(r.sigma, r.epsilon, r.topo, r.params)
        g.set_pair_layout(b.pair_layout())
        px, py, pz = r.x.copy(), r.y.copy(), r.z.copy()
        g.build_neighbor_list(px, py, pz)
        g.set_work_chunk(entries)
        ep, gx, gy, gz = g.compute_forces(px, py, pz)
        o0, o1 = b.offsets[k], b.offsets[k + 1]
        assert gx.tobytes() == fx[o0:o1].tobytes() and gy.tobytes() == fy[o0:o1].tobytes()
        assert gz.tobytes() == fz[o0:o1].tobytes() and ep.total == e[k].total
    b.set_shape(0, 16)  # back to the automatic shape for the tests after this one


def test_autotuned_shape_is_pinned_and_reproducible(batch_case):
    reps, b, (x, y, z), e0, f0 = batch_case
    say("[gpu] autotune")
    c, w, ms = b.autotune()
    if b.pair_layout() == "tiles":
        assert c in (8, 16, 24, 32) and 1 <= w <= 64 and ms > 0.0
    else:  # partner lists: no shape to tune, the step timed
        assert c == 0 and w == 0 and ms > 0.0
    b.compute()
    e1, _ = b.sync()
    f1 = b.get_forces()
    b.compute()
    e2, _ = b.sync()
    f2 = b.get_forces()
    for u, v, r in zip(f1, f2, f0):
        # the autotuner's choice (timing-dependent) changes only the schedule
        assert u.tobytes() == v.tobytes() == r.tobytes()
    assert [a.total for a in e1] == [a.total for a in e2] == [a.total for a in e0]


def test_replica_handles_refuse_list_and_shape_changes(batch_case):
    """A replica's list and work items belong to its batch (the captured graph
    points at them): the borrowed handle exports and evaluates, but does not
    rebuild or reshape them."""
    reps, b, (x, y, z), e0, f0 = batch_case
    h = b.replica(0)
    r = reps[0]
    for call in (lambda: h.build_neighbor_list(r.x.copy(), r.y.copy(), r.z.copy()),
                 lambda: h.set_work_chunk(16)):
        with pytest.raises(gh.GmError) as ei:
            call()
        assert ei.value.status == 5
    b.compute()
    e, _ = b.sync()
    for u, v in zip(b.get_forces(), f0):
        assert u.tobytes() == v.tobytes()


def test_layout_switch_in_place(batch_case):
    """A built batch switched to the other layout and back: the other layout's
    bytes equal that layout evaluated alone (first replica), and switching
    back restores every byte."""
    reps, b, (x, y, z), e0, f0 = batch_case
    here = b.pair_layout()
    other = "tiles" if here == "partners" else "partners"
    say(f"[gpu] batch layout {here} -> {other} -> {here}")
    b.set_pair_layout(other)
    assert b.pair_layout() == other
    b.compute()
    e1, _ = b.sync()
    fx, fy, fz = b.get_forces()
    r = reps[0]
    g = gh.System(r.sigma, r.epsilon, r.topo, r.params)
    g.set_pair_layout(other)
    px, py, pz = r.x.copy(), r.y.copy(), r.z.copy()
    g.build_neighbor_list(px, py, pz)
    ep, gx, gy, gz = g.compute_forces(px, py, pz)
    o1 = b.offsets[1]
    assert gx.tobytes() == fx[:o1].tobytes() and gy.tobytes() == fy[:o1].tobytes()
    assert gz.tobytes() == fz[:o1].tobytes() and ep.total == e1[0].total
    b.set_pair_layout("auto" if here == "partners" else "tiles")
    assert b.pair_layout() == here
    b.compute()
    e2, _ = b.sync()
    for u, v in zip(b.get_forces(), f0):
        assert u.tobytes() == v.tobytes()
    assert [t.total for t in e2] == [t.total for t in e0]


def test_replica_exports_then_batched_evaluations():
    """Statistics and pair-list exports through replica handles make a
    partner-list replica's tile list on demand, which can reallocate the LJ
    energy records its groups write; the batch then makes its plan again
    before launching.  Partner-list evaluations after the exports equal those
    of a batch never exported from, and the same batch switched to tiles with
    a new work shape equals each replica evaluated alone in tiles."""
    reps = make_replicas([500, 1200, 2400, 3100, 4000, 5200, 800, 1800], seed=77)
    members = [(r.sigma, r.epsilon, r.topo, r.params) for r in reps]
    x, y, z = concat(reps)
    say("[gpu] batch without exports")
    b0 = gh.Batch(members)
    b0.build_neighbor_lists(x.copy(), y.copy(), z.copy())
    b0.compute()
    e0, _ = b0.sync()
    f0 = b0.get_forces()
    b0.close()
    say("[gpu] batch with statistics and exports before evaluating")
    b = gh.Batch(members)
    b.build_neighbor_lists(x, y, z)
    assert b.pair_layout() == "partners"
    n_cut = sum(b.replica(k).stats()["n_pairs_cutoff"] for k in range(len(reps)))
    assert n_cut > 0
    for it in range(3):
        b.compute()
        e, _ = b.sync()
        for u, v in zip(b.get_forces(), f0):
            assert u.tobytes() == v.tobytes(), it
        assert [t.total for t in e] == [t.total for t in e0] and [t.lj for t in e] == [t.lj for t in e0]
    say("[gpu] switched to tiles, 8 entries per item")
    b.set_pair_layout("tiles")
    b.set_shape(8, 8)
    b.compute()
    e, _ = b.sync()
    fx, fy, fz = b.get_forces()
    for k, r in enumerate(reps):
        g = gh.System(r.sigma, r.epsilon, r.topo, r.params)
        g.set_pair_layout("tiles")
        px, py, pz = r.x.copy(), r.y.copy(), r.z.copy()
        g.build_neighbor_list(px, py, pz)
        ep, gx, gy, gz = g.compute_forces(px, py, pz)
        o0, o1 = b.offsets[k], b.offsets[k + 1]
        assert gx.tobytes() == fx[o0:o1].tobytes() and gy.tobytes() == fy[o0:o1].tobytes(), k
        assert gz.tobytes() == fz[o0:o1].tobytes() and ep.total == e[k].total, k
    b.close()


def test_batch_errors():
    reps = make_replicas([500, 700])
    b = gh.Batch([(r.sigma, r.epsilon, r.topo, r.params) for r in reps])
    with pytest.raises(gh.GmError) as ei:
        b.compute()
    assert ei.value.status == 7
    r = reps[0]
    bad = gh.Params(box=(17.0, 18.0, 17.0), periodic=True, r_cutoff=9.0, r_switch=7.5, r_skin=10.5)
    with pytest.raises(gh.GmError) as ei:
        gh.Batch([(r.sigma, r.epsilon, r.topo, bad)])
    assert ei.value.status == 5
