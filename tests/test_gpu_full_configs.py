"""The benchmark configurations at their stated sizes, against the f64 oracle.

  - 512 batched replicas of log-uniform 500..20,000 atoms (2.68 M atoms), one
    persistent launch from a captured graph: every replica's Verlet list
    bit-exact against the oracle's cell-list build, per-term energies within
    1e-6, force RMS within 1e-5; a spread sample of replicas byte-identical to
    the same replica evaluated alone (geodesic-gpu/tests/forces_gpu_vs_cpu.rs:7-24
    and gpu_determinism.rs:7-20 are the pattern; paths relative to the
    reference repository root).
  - 250,000 atoms, 136 A box, under the rebuild protocol: 20 steps of seeded
    Gaussian moves of 0.02 A per coordinate, forces on the stale list against
    the oracle evaluating that same list, then a rebuild bit-exact against the
    oracle's list at the new positions and forces on it.
  - The shell-force surrogate on the 24,000-atom system: features, shell
    Hessian and shell force against gm_sur_oracle.c; trained on 20,000 frames;
    the 4:1 multiple-time-step force against the f64 oracle within 1e-3.
"""
import sys

import numpy as np
import pytest

import geodesic_hip as gh
from geodesic_hip import systems
from oracle import oracle as orc

from _common import force_rms_rel

pytestmark = pytest.mark.gpu

E_REL = 1e-6
F_RMS = 1e-5
THREADS = 16


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check_terms(e, t, tag):
    for name in ("bond", "angle", "dihedral", "lj", "lj14", "total"):
        g, r = getattr(e, name), getattr(t, name)
        if r == 0.0:
            assert g == 0.0, f"{tag}: {name} gpu={g} oracle=0"
        else:
            assert abs(g - r) <= E_REL * abs(r), f"{tag}: {name} gpu={g!r} oracle={r!r}"


def oracle_eval(x, y, z, s, p, pair_list=None):
    """Wrap (periodic), cell-list build (unless a list is given) and
    CpuBackend::compute_forces in f64."""
    x, y, z = x.copy(), y.copy(), z.copy()
    if pair_list is None:
        orc.wrap_positions(x, y, z, p.box)
        pair_list = orc.neighbor_build(x, y, z, p.box, p.r_skin, s.topo.excl_i, s.topo.excl_j, method="cells")
    pi, pj = pair_list
    t, fx, fy, fz = orc.cpu_backend_compute(x, y, z, s.sigma, s.epsilon, s.topo, pi, pj, p.box, p.r_cutoff,
                                            p.r_switch, n_threads=THREADS, scale14=p.scale_14,
                                            bonded_min_image=p.bonded_min_image, periodic=p.periodic)
    return (x, y, z), (pi, pj), t, (fx, fy, fz)


# ---- 512 replicas -----------------------------------------------------------

def test_512_replicas_against_the_oracle_and_alone():
    """The 512-replica batch in its default layout (partner lists), then
    switched in place to tiles with the autotuner's shape: every list
    bit-exact, every replica within the bars of its f64 oracle, and a spread
    sample byte-identical to the replica evaluated alone in the same layout."""
    reps = systems.replica_set(512)
    n_tot = sum(r.n for r in reps)
    say(f"[gpu] 512 replicas, {n_tot} atoms: batch build + compute")
    b = gh.Batch([(r.sigma, r.epsilon, r.topo, r.params) for r in reps])
    x = np.concatenate([r.x for r in reps])
    y = np.concatenate([r.y for r in reps])
    z = np.concatenate([r.z for r in reps])
    b.build_neighbor_lists(x, y, z)
    assert b.pair_layout() == "partners"
    ref = []
    sample = sorted({0, 1, 137, 255, 384, 511, int(np.argmax([r.n for r in reps])), int(np.argmin([r.n for r in reps]))})
    for layout in ("partners", "tiles"):
        if layout == "tiles":
            b.set_pair_layout("tiles")
            chunk, waves, _ = b.autotune()
            say(f"[gpu] tiles, autotuned shape: {chunk} entries per item, {waves} waves per CU")
        assert b.pair_layout() == layout
        b.compute()
        e, _ = b.sync()
        fx, fy, fz = b.get_forces()
        worst_rms, worst_e = 0.0, 0.0
        for k, r in enumerate(reps):
            o0, o1 = b.offsets[k], b.offsets[k + 1]
            if layout == "partners":
                (wx, wy, wz), (pi, pj), t, g3 = oracle_eval(r.x, r.y, r.z, r, r.params)
                assert x[o0:o1].tobytes() == wx.tobytes(), k  # wrapped in place as neighbor.rs:13-15
                qi, qj = b.replica(k).pair_list()
                assert np.array_equal(qi, pi) and np.array_equal(qj, pj), f"replica {k} (n={r.n}): list"
                ref.append((t, g3))
            t, (gx, gy, gz) = ref[k]
            check_terms(e[k], t, f"{layout}, replica {k}")
            rms = force_rms_rel(fx[o0:o1], fy[o0:o1], fz[o0:o1], gx, gy, gz)
            assert rms <= F_RMS, f"{layout}, replica {k}: force rms {rms:.3e}"
            worst_rms = max(worst_rms, rms)
            worst_e = max(worst_e, abs(e[k].total - t.total) / abs(t.total))
            if k % 128 == 0:
                say(f"[gpu] {layout}: replica {k}/512 (n={r.n}) ok; worst so far: force rms {worst_rms:.2e}, "
                    f"energy {worst_e:.2e}")
        for k in sample:
            r = reps[k]
            g = gh.System(r.sigma, r.epsilon, r.topo, r.params)
            g.set_pair_layout(layout)
            px, py, pz = r.x.copy(), r.y.copy(), r.z.copy()
            g.build_neighbor_list(px, py, pz)
            ep, gx, gy, gz = g.compute_forces(px, py, pz)  # (tiles: its own automatic work shape)
            o0, o1 = b.offsets[k], b.offsets[k + 1]
            assert gx.tobytes() == fx[o0:o1].tobytes() and gy.tobytes() == fy[o0:o1].tobytes(), k
            assert gz.tobytes() == fz[o0:o1].tobytes() and ep.total == e[k].total and ep.lj == e[k].lj, k
        say(f"[gpu] {layout}: batch vs alone byte-identical on replicas {sample}; worst force rms {worst_rms:.2e}")


# ---- 250k rebuild protocol ----------------------------------------------------

def test_250k_rebuild_protocol():
    s = systems.system_250k()
    p = s.params
    g = gh.System(s.sigma, s.epsilon, s.topo, p)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    say("[gpu] 250k: build")
    g.build_neighbor_list(x, y, z)
    (wx, wy, wz), (pi, pj), _, _ = oracle_eval(s.x, s.y, s.z, s, p)
    assert x.tobytes() == wx.tobytes()
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi) and np.array_equal(qj, pj)
    rng = np.random.default_rng(250)
    for step in range(20):
        d = rng.normal(0.0, 0.02, (3, s.n))
        x, y, z = x + d[0], y + d[1], z + d[2]
        e, fx, fy, fz = g.compute_forces(x, y, z)
    assert not g.needs_rebuild(x, y, z)  # 20 steps stay inside half the skin
    say("[gpu] 250k: forces after 20 steps on the stale list")
    _, _, t, f = oracle_eval(x, y, z, s, p, pair_list=(pi, pj))
    check_terms(e, t, "250k stale list")
    rms = force_rms_rel(fx, fy, fz, *f)
    assert rms <= F_RMS, rms
    say(f"[gpu] 250k stale: force rms {rms:.3e}")
    say("[gpu] 250k: rebuild at step 20")
    g.build_neighbor_list(x, y, z)
    (wx, wy, wz), (pi2, pj2), t2, f2 = oracle_eval(x, y, z, s, p)
    assert x.tobytes() == wx.tobytes() and y.tobytes() == wy.tobytes() and z.tobytes() == wz.tobytes()
    qi, qj = g.pair_list()
    assert np.array_equal(qi, pi2) and np.array_equal(qj, pj2)
    e2, *fg2 = g.compute_forces(x, y, z)
    check_terms(e2, t2, "250k rebuilt")
    rms2 = force_rms_rel(*fg2, *f2)
    assert rms2 <= F_RMS, rms2
    say(f"[gpu] 250k rebuilt: {len(pi2)} listed pairs, force rms {rms2:.3e}")


# ---- surrogate on the 24k system ------------------------------------------------

def test_surrogate_on_the_24k_system():
    s = systems.system_24k()
    p = gh.Params(box=s.params.box, periodic=True, r_cutoff=12.0, r_switch=7.5, r_skin=13.5,
                  scale_14=s.params.scale_14, bonded_min_image=s.params.bonded_min_image)
    cls = s.types % 4
    sur = gh.Surrogate(s.sigma, s.epsilon, s.topo, p, cls, seed=5)
    d = tuple(gh.DeviceArray.from_host(a) for a in (s.x, s.y, s.z))
    say("[gpu] surrogate 24k: relax")
    sur.minimize(*d, steps=300)
    sur.build(*d)
    x, y, z = (a.to_host() for a in d)
    say("[gpu] surrogate 24k: features, K and shell force against the oracle")
    f = sur.features(*d)
    ref = orc.sur_features(x, y, z, p.box, cls)
    err = np.abs(f - ref)
    assert np.all(err <= 1e-4 * np.maximum(1.0, np.abs(ref))), float(err.max())
    assert np.sqrt(np.mean(err ** 2)) <= 1e-5 * np.sqrt(np.mean(ref ** 2))
    K = sur.hessian()
    pi, pj = orc.neighbor_build(x, y, z, p.box, p.r_skin, s.topo.excl_i, s.topo.excl_j, method="cells")
    fsh_ref, Kref = orc.sur_shell(x, y, z, s.sigma, s.epsilon, pi, pj, p.box)
    kerr = np.abs(K - Kref)
    assert np.sqrt(np.mean(kerr ** 2)) <= 1e-5 * np.sqrt(np.mean(Kref ** 2))
    fs = [gh.DeviceArray(s.n) for _ in range(3)]
    sur.shell_forces(*d, *fs)
    sur.sync()
    fsh = np.stack([a.to_host() for a in fs])
    rms = np.sqrt(np.mean((fsh - fsh_ref) ** 2)) / np.sqrt(np.mean(fsh_ref ** 2))
    assert rms <= 1e-5, rms
    say("[gpu] surrogate 24k: training on 20,000 frames")
    hist, rep = sur.train(*d, n_frames=20000, seed=101, lr=1e-3, batch=8192)
    say(f"[gpu] training: {rep}")
    assert rep["frames"] == 20000 and np.all(np.isfinite(hist))
    say("[gpu] surrogate 24k: 4:1 multiple-time-step force against the f64 oracle")
    q = tuple(gh.DeviceArray.from_host(a) for a in (x, y, z))
    sur.build(*q)
    fm = [gh.DeviceArray(s.n) for _ in range(3)]
    sur.mts_reset()
    for k in range(4):
        if k > 0:
            sur.bd_step(*q, *fm, seed=303, step=k - 1)
        sur.mts_step(*q, *fm, mode=gh.MTS_NET)
    sur.sync()
    got = np.stack([a.to_host() for a in fm])
    qx, qy, qz = (a.to_host() for a in q)
    _, _, _, fo = oracle_eval(qx, qy, qz, s, p)
    fo = np.stack(fo)
    err = np.sqrt(np.mean((got - fo) ** 2)) / np.sqrt(np.mean(fo ** 2))
    say(f"[gpu] 24k MTS force RMS error vs f64 at the third inner step: {err:.3e}")
    assert err <= 1e-3
