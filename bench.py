"""Force-evaluation throughput on one MI355X.

Workload: the 24,000-atom periodic system (62 A box, 16 LJ types, 3,000-atom
chain + 7,000 rigid triplets; r_c 9, r_sw 7.5, list 10.5 A).  One step = one
force evaluation (nonbonded partner lists, or cluster tiles with --layout
tiles, + bonds/angles/dihedrals/scaled 1-4 + fixed-point reduction + f64
output) at positions already resident in HBM.  The Verlet list is built once
before timing (its device time is reported).

Metric: pair interactions / s, where a pair interaction is an unordered,
non-excluded pair with r <= r_cutoff, counted once.

Prints one JSON line.  cpu_baseline: the C restatement of the reference's
CpuBackend::compute_forces (Rayon static strips, f64) on the host cores.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "geodesic-m_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

FLOP_PER_PAIR = 40          # canonical count per pair interaction (FMA = 2)
PEAK_FP32_TFLOPS = 157.3    # MI355X published FP32 peak
PEAK_HBM_GBS = 8000.0
PROFILE_EVERY = 20          # timed region: HIP events on every 20th evaluation
FORCE_KERNEL = {"partners": "k_forces_pl<cubic> (partner lists + bonded workgroups)",
                "tiles": "k_forces<cubic> (tiles + bonded workgroups)"}
TRAFFIC_FILE = {"partners": "nonbonded_pl_24k_traffic.json", "tiles": "nonbonded_24k_traffic.json"}


def layout_name(stats):
    """The layout a system's evaluations use (gm_stats.pair_layout)."""
    return "partners" if stats["pair_layout"] == 1 else "tiles"


def traffic_from_profiles(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(s, n_pairs_cut, evals=20, warm=3):
    from oracle import oracle as orc
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    orc.wrap_positions(x, y, z, s.params.box)
    t0 = time.perf_counter()
    pi, pj = orc.neighbor_build(x, y, z, s.params.box, s.params.r_skin, s.topo.excl_i, s.topo.excl_j,
                                method="brute")
    list_s = time.perf_counter() - t0
    times = []
    for k in range(warm + evals):
        t0 = time.perf_counter()
        orc.cpu_backend_compute(x, y, z, s.sigma, s.epsilon, s.topo, pi, pj, s.params.box,
                                s.params.r_cutoff, s.params.r_switch, n_threads=cores,
                                scale14=s.params.scale_14, bonded_min_image=s.params.bonded_min_image)
        if k >= warm:
            times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": n_pairs_cut / med, "unit": "pair-interactions/s", "cores": cores, "kind": "port",
            "sample": f"{evals} force evaluations (median) of the same 24k system after {warm} warm-ups, "
                      f"T={cores} static strips; serial O(N^2) list build timed separately",
            "ms_per_eval": med * 1e3, "list_build_s": list_s}


def build_wall_ms(g, x, y, z, n=10):
    """Wall-clock time of a list build from positions resident on the device
    (build_neighbor_list_device: the build kernels and the one read-back of
    the counts at its end), median of n."""
    import geodesic_hip as gh
    d = [gh.DeviceArray.from_host(a) for a in (x, y, z)]
    g.build_neighbor_list_device(*(a.ptr for a in d))
    ts = []
    for _ in range(n):
        t0 = time.perf_counter()
        g.build_neighbor_list_device(*(a.ptr for a in d))
        ts.append((time.perf_counter() - t0) * 1e3)
    for a in d:
        a.free()
    return statistics.median(ts)


def time_resident(g, steps, warmup, every=PROFILE_EVERY):
    """Average wall ms per evaluation at resident positions over the timed
    region, the HIP-event profile of that region (every `every`-th evaluation
    carries events around its kernels), and the wall ms per evaluation of the
    same loop with events on every evaluation."""
    for _ in range(warmup):
        g.compute_forces_resident()
    g.profile_begin(1)  # create the event pool outside the timed region
    for _ in range(min(steps, 8192)):
        g.compute_forces_resident()
    g.profile_end()
    g.sync()
    g.profile_begin(every)
    t0 = time.perf_counter()
    for _ in range(steps):
        g.compute_forces_resident()
    e = g.sync()
    t1 = time.perf_counter()
    prof = g.profile_end()
    g.profile_begin(1)
    t2 = time.perf_counter()
    for _ in range(steps):
        g.compute_forces_resident()
    g.sync()
    t3 = time.perf_counter()
    g.profile_end()
    return (t1 - t0) * 1e3 / steps, prof, e, (t3 - t2) * 1e3 / steps


ROOF_LAUNCHES = 200         # the roofline pass times this many force-kernel launches


def roofline_pass(g, n_pairs, launches=ROOF_LAUNCHES):
    """A fixed-length pass beside the timed region: every evaluation's force
    kernel timed by its own dispatch-stamped start and end events.  Returns
    the kernel's average and minimum ms over `launches` launches and the
    achieved FP32 rate (canonical 40 FLOP per pair interaction)."""
    g.profile_begin(1)
    for _ in range(launches):
        g.compute_forces_resident()
    g.sync()
    prof = g.profile_end()
    avg = prof["ms_nonbonded_total"] / max(prof["n_evals"], 1)
    achieved = n_pairs * FLOP_PER_PAIR / (avg * 1e-3) / 1e12
    return {"bound": "valu", "kernel_ms_avg": avg, "kernel_ms_min": prof["ms_nonbonded_min"],
            "timed_launches": prof["n_evals"], "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
            "eval_ms_avg": prof["ms_eval_total"] / max(prof["n_evals"], 1)}


def bench_reference_workloads(cpu=True):
    """The reference's own benchmarks (geodesic/benches/benchmarks.rs:97-135):
    bench_lj_inner_loop (nonbonded throughput in atom-pairs/s over the Verlet
    list) and bench_neighbor_rebuild (list build time), on its 10,000-atom
    lattice (2.0 A spacing, sigma 3, eps 0.1, r_c/r_skin/r_sw 5/6/4), with the
    oracle's restatement of the reference's single-threaded routines beside
    them (compute_pair_forces over the list; the serial O(N^2) neighbor::build)."""
    import geodesic_hip as gh
    from geodesic_hip import systems
    s = systems.lj_lattice(10000)
    g = gh.System(s.sigma, s.epsilon, s.topo, s.params)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    g.build_neighbor_list(x, y, z)
    builds = []
    for _ in range(5):
        g.build_neighbor_list(x, y, z)
        builds.append(g.stats()["ms_build"])
    t0 = time.perf_counter()
    for _ in range(5):
        g.build_neighbor_list(x, y, z)
    build_wall_ms = (time.perf_counter() - t0) * 1e3 / 5
    st = g.stats()
    n_list, n_cut = st["n_list_pairs"], st["n_pairs_cutoff"]
    ms, _, _, _ = time_resident(g, 200, 20)
    out = {"workload": "10,000-atom lattice (2.0 A spacing), 44 A periodic box, r_c/r_skin/r_sw 5/6/4, "
                       "sigma 3 / eps 0.1, no topology",
           "n_list_pairs": n_list, "n_pairs_cutoff": n_cut,
           "bench_lj_inner_loop": {"ms_per_eval": ms, "atom_pairs_per_s": n_list / (ms * 1e-3),
                                   "pair_interactions_per_s": n_cut / (ms * 1e-3)},
           "bench_neighbor_rebuild": {"ms_device": statistics.median(builds), "ms_wall": build_wall_ms}}
    if cpu:
        from oracle import oracle as orc
        pi, pj = orc.neighbor_build(x, y, z, s.params.box, s.params.r_skin, s.topo.excl_i, s.topo.excl_j,
                                    method="brute")
        t0 = time.perf_counter()
        orc.neighbor_build(x, y, z, s.params.box, s.params.r_skin, s.topo.excl_i, s.topo.excl_j, method="brute")
        cpu_build_s = time.perf_counter() - t0
        times = []
        for k in range(8):
            t0 = time.perf_counter()
            orc.pair_forces(x, y, z, s.sigma, s.epsilon, pi, pj, s.params.r_cutoff, s.params.r_switch,
                            s.params.box)
            if k >= 2:
                times.append(time.perf_counter() - t0)
        cms = statistics.median(times) * 1e3
        out["cpu_baseline"] = {"cores": 1, "kind": "port",
                               "sample": "6 evaluations (median) of compute_pair_forces over the list after 2 "
                                         "warm-ups; one serial O(N^2) list build",
                               "bench_lj_inner_loop_ms": cms, "atom_pairs_per_s": n_list / (cms * 1e-3),
                               "bench_neighbor_rebuild_ms": cpu_build_s * 1e3}
        out["speedup_lj_inner_loop"] = cms / ms
        out["speedup_neighbor_rebuild_wall"] = cpu_build_s * 1e3 / build_wall_ms
    return out


def bench_build(repeats=20, layout="auto"):
    """The Verlet list build alone, at steady state (buffers sized, the build
    replayed as one graph), on the 24,000- and 250,000-atom systems: device
    time between the events around the build's kernels, and wall time of
    build_neighbor_list_device on resident positions (the kernels plus the
    one read-back of the counts)."""
    import geodesic_hip as gh
    from geodesic_hip import systems
    out = {}
    for name, mk in (("24k", systems.system_24k), ("250k", systems.system_250k)):
        s = mk()
        g = gh.System(s.sigma, s.epsilon, s.topo, s.params)
        g.set_pair_layout(layout)
        d = [gh.DeviceArray.from_host(a) for a in (s.x, s.y, s.z)]
        for _ in range(3):
            g.build_neighbor_list_device(*(a.ptr for a in d))
        wall, dev = [], []
        for _ in range(repeats):
            t0 = time.perf_counter()
            g.build_neighbor_list_device(*(a.ptr for a in d))
            wall.append((time.perf_counter() - t0) * 1e3)
            dev.append(g.stats()["ms_build"])
        st = g.stats()
        out[name] = {"n_atoms": s.n, "ms_device": statistics.median(dev), "ms_wall": statistics.median(wall),
                     "n_superclusters": st["n_superclusters"], "n_entries": st["n_entries"],
                     "n_list_pairs": st["n_list_pairs"]}
        for a in d:
            a.free()
    return out


def bench_250k(steps=100, warmup=10, layout="auto"):
    """250,000 atoms, 136 A box: force evaluation at resident positions and the
    list build, amortised over the 20-step rebuild interval; and the workload's
    motion (protocol_250k), which evaluates displaced positions."""
    import geodesic_hip as gh
    from geodesic_hip import systems
    s = systems.system_250k()
    g = gh.System(s.sigma, s.epsilon, s.topo, s.params)
    g.set_pair_layout(layout)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    g.build_neighbor_list(x, y, z)
    builds = []
    for _ in range(3):
        g.build_neighbor_list(x, y, z)
        builds.append(g.stats()["ms_build"])
    st = g.stats()
    wall_ms = build_wall_ms(g, x, y, z)
    ms, prof, e, ms_all_events = time_resident(g, steps, warmup)
    build_ms = statistics.median(builds)
    n_cut = st["n_pairs_cutoff"]
    roof = roofline_pass(g, n_cut, 50)
    proto = protocol_250k(g, s)
    return {"workload": "force evaluation, 250,000 atoms, 136 A periodic box, rebuild every 20 steps",
            "kernel": FORCE_KERNEL[layout_name(st)], "n_pairs_cutoff": n_cut, "ms_per_eval": ms, "ms_per_eval_events_every_eval": ms_all_events,
            "protocol": proto,
            "kernel_ms_avg": roof["kernel_ms_avg"], "roofline": roof,
            "list_build_ms": build_ms, "list_build_ms_wall": wall_ms,
            "ms_per_step_amortised": ms + max(build_ms, wall_ms) / 20.0,
            "pair_interactions_per_s": n_cut / (ms * 1e-3),
            "pair_interactions_per_s_amortised": n_cut / ((ms + build_ms / 20.0) * 1e-3)}


def protocol_250k(g, s, interval=20, sigma=0.02, cycles=3, seed=7):
    """The 250k workload's motion: Gaussian displacements of sigma A per
    component per step, the list rebuilt every `interval` steps.  One cycle =
    a build from the first frame, then an evaluation at each of the interval's
    frames (precomputed, resident in HBM); timed over whole cycles (wall),
    and the force kernel over one cycle with dispatch-stamped events."""
    import geodesic_hip as gh
    rng = np.random.default_rng(seed)
    pos = np.stack([s.x, s.y, s.z]).astype(np.float64)
    frames = []
    for _ in range(interval):
        frames.append([gh.DeviceArray.from_host(np.ascontiguousarray(v)) for v in pos])
        pos = pos + rng.normal(0.0, sigma, pos.shape)
    out = [gh.DeviceArray(s.n) for _ in range(3)]

    def cycle():
        for k, fr in enumerate(frames):
            if k == 0:
                g.build_neighbor_list_device(*(a.ptr for a in fr))
            g.compute_forces_device(*(a.ptr for a in fr), *(o.ptr for o in out))
        g.sync()

    cycle()  # warm-up
    t0 = time.perf_counter()
    for _ in range(cycles):
        cycle()
    ms = (time.perf_counter() - t0) * 1e3 / (cycles * interval)
    g.profile_begin(1)
    cycle()
    prof = g.profile_end()
    kern = prof["ms_nonbonded_total"] / max(prof["n_evals"], 1)
    d = np.sqrt(((pos - np.stack([s.x, s.y, s.z])) ** 2).sum(axis=0)).max()
    for fr in frames:
        for a in fr:
            a.free()
    for o in out:
        o.free()
    return {"interval": interval, "sigma_A_per_step": sigma, "ms_per_step_build_included": ms,
            "kernel_ms_avg": kern, "max_displacement_at_interval_end_A": float(d)}


def bench_batched(n_replicas=512, steps=20, layout="auto"):
    """512 independent replicas (log-uniform 500..20000 atoms): one graph launch
    per batched step vs the plain path evaluating the replicas one by one.  The
    batch's default layout (partner lists of every replica) is the headline;
    the same batch switched in place to tiles, with the autotuner's shape, is
    reported beside it."""
    import geodesic_hip as gh
    from geodesic_hip import systems
    t0 = time.perf_counter()
    reps = systems.replica_set(n_replicas)
    t_gen = time.perf_counter() - t0
    b = gh.Batch([(r.sigma, r.epsilon, r.topo, r.params) for r in reps])
    b.set_pair_layout(layout)
    x = np.concatenate([r.x for r in reps])
    y = np.concatenate([r.y for r in reps])
    z = np.concatenate([r.z for r in reps])
    t0 = time.perf_counter()
    b.build_neighbor_lists(x, y, z)
    t_build = time.perf_counter() - t0
    n_cut = sum(b.replica(k).stats()["n_pairs_cutoff"] for k in range(len(reps)))

    def measure():
        chunk, waves, tune_ms = b.autotune()
        for _ in range(3):
            b.compute()
        b.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.compute()
        _, last_ms = b.sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        return ms, last_ms, {"entries_per_item": chunk, "waves_per_cu": waves, "ms": tune_ms}

    main_layout = b.pair_layout()
    batched_ms, last_ms, shape = measure()
    tiles = None
    if main_layout != "tiles":
        b.set_pair_layout("tiles")
        t_ms, t_last, t_shape = measure()
        tiles = {"ms_per_batched_step": t_ms, "device_ms_last_step": t_last,
                 "pair_interactions_per_s": n_cut / (t_ms * 1e-3), "autotuned_shape": t_shape}
    b.close()
    # plain path: every replica as its own system (own stream, automatic
    # layout and work shape) with its own launches (prep, forces, finish): all
    # 512 evaluations launched, then all waited for
    plain = []
    for r in reps:
        g = gh.System(r.sigma, r.epsilon, r.topo, r.params)
        px, py, pz = r.x.copy(), r.y.copy(), r.z.copy()
        g.build_neighbor_list(px, py, pz)
        plain.append(g)
    for g in plain:
        g.compute_forces_resident()
    for g in plain:
        g.sync()
    sweeps = 3
    t0 = time.perf_counter()
    for _ in range(sweeps):
        for g in plain:
            g.compute_forces_resident()
    for g in plain:
        g.sync()
    plain_ms = (time.perf_counter() - t0) * 1e3 / sweeps
    for g in plain:
        g.close()
    if tiles is not None:
        tiles["speedup_vs_plain"] = plain_ms / tiles["ms_per_batched_step"]
    step_tflops = n_cut * FLOP_PER_PAIR / (last_ms * 1e-3) / 1e12
    kern = "partner-list groups of every replica" if main_layout == "partners" else "persistent tiles"
    return {"workload": f"{len(reps)} replicas, {sum(r.n for r in reps)} atoms, log-uniform 500..20000 atoms, "
                        f"one launch per step from a captured hipGraph ({kern})",
            "pair_layout": main_layout,
            "roofline": {"bound": "valu", "achieved": step_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": step_tflops / PEAK_FP32_TFLOPS,
                         "timing": "device time of the whole graph step (prep, bonded, nonbonded, finish)"},
            "n_atoms": sum(r.n for r in reps), "n_pairs_cutoff": n_cut, "ms_per_batched_step": batched_ms,
            "device_ms_last_step": last_ms, "ms_plain_path_all_replicas": plain_ms,
            "pair_interactions_per_s": n_cut / (batched_ms * 1e-3),
            "plain_pair_interactions_per_s": n_cut / (plain_ms * 1e-3),
            "speedup_vs_plain": plain_ms / batched_ms,
            "autotuned_shape": shape,
            "tile_layout": tiles,
            "setup_s": {"generate": t_gen, "build_lists": t_build}}


def bench_surrogate(frames=20000, steps=400, traj_cycles=25, min_steps=300):
    """Shell-force surrogate on the 24k system with LJ switched over 7.5-12 A
    (r_c 12, list 13.5): steepest descent, then training on the GPU along a
    seeded Brownian trajectory; held-out loss against the plain multiple-time-
    step method (shell force held from the last exact step); force error along
    a Brownian trajectory driven by the multiple-time-step forces; throughput
    of the 4:1 driver against the exact model evaluated every step."""
    import geodesic_hip as gh
    from geodesic_hip import systems
    s = systems.system_24k()
    p = gh.Params(box=s.params.box, periodic=True, r_cutoff=12.0, r_switch=7.5, r_skin=13.5,
                  scale_14=s.params.scale_14, bonded_min_image=s.params.bonded_min_image)
    sur = gh.Surrogate(s.sigma, s.epsilon, s.topo, p, s.types % 4, seed=5)
    d = tuple(gh.DeviceArray.from_host(a) for a in (s.x, s.y, s.z))
    t0 = time.perf_counter()
    fmax = sur.minimize(*d, steps=min_steps)
    min_s = time.perf_counter() - t0
    relaxed = [a.to_host() for a in d]
    dev = lambda: tuple(gh.DeviceArray.from_host(a) for a in relaxed)  # noqa: E731
    hist, rep = sur.train(*d, n_frames=frames, seed=101, lr=1e-3, batch=8192)
    _, held = sur.train(*dev(), n_frames=300, seed=202, lr=0.0, batch=8192, update=False)
    n = s.n
    # force error at inner steps along a Brownian trajectory driven by the
    # multiple-time-step forces, against the exact model on the GPU
    ex = sur.system(0)
    f = [gh.DeviceArray(n) for _ in range(3)]
    g = [gh.DeviceArray(n) for _ in range(3)]
    errs = {}
    modes = {"network": gh.MTS_NET, "held_shell_force": gh.MTS_HELD, "taylor": gh.MTS_TAYLOR}
    for name, mode in modes.items():
        q = dev()
        sur.build(*q)
        sur.mts_reset()
        num = den = 0.0
        for k in range(4 * traj_cycles):
            if k % 4 == 0:
                sur.maybe_rebuild(*q)
            if sur.mts_step(*q, *f, mode=mode) is False:
                ex.compute_forces_device(*(a.ptr for a in q), *(a.ptr for a in g))
                ex.sync()
                fm = np.stack([a.to_host() for a in f])
                fe = np.stack([a.to_host() for a in g])
                num += float(np.sum((fm - fe) ** 2))
                den += float(np.sum(fe ** 2))
            sur.bd_step(*q, *f, seed=303, step=k)
        errs[name] = (num / den) ** 0.5
    # throughput at fixed positions: exact model every step vs the 4:1 driver
    q = dev()
    sur.build(*q)
    st = ex.stats()
    for _ in range(8):
        ex.compute_forces_device(*(a.ptr for a in q), *(a.ptr for a in f))
    ex.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ex.compute_forces_device(*(a.ptr for a in q), *(a.ptr for a in f))
    ex.sync()
    plain_ms = (time.perf_counter() - t0) * 1e3 / steps
    sur.mts_reset()
    for _ in range(8):
        sur.mts_step(*q, *f)
    sur.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        sur.mts_step(*q, *f)
    sur.sync()
    mts_ms = (time.perf_counter() - t0) * 1e3 / steps
    n_cut = st["n_pairs_cutoff"]
    return {"workload": "24,000-atom system, LJ switched 7.5-12 A (r_c 12, list 13.5); inner model r_c 7.5 "
                        "switched from 6.5 (list 9.0); 96 radial features, 96-128-128-3 SiLU, BF16 MFMA; 4:1 MTS",
            "relax": {"steps": min_steps, "seconds": min_s, "max_force_after": fmax},
            "train": {"frames": rep["frames"], "steps": rep["steps"], "seconds": rep["ms_total"] / 1e3,
                      "trajectory": "Brownian, 300 K, 0.02 A noise per coordinate per step",
                      "loss_first": rep["loss_first"], "loss_last": rep["loss_last"],
                      "loss_mean_network": rep["loss_net"], "loss_mean_held_shell": rep["loss_stale"],
                      "loss_mean_taylor": rep["loss_taylor"]},
            "held_out": {"frames": held["frames"], "loss_network": held["loss_net"],
                         "loss_held_shell": held["loss_stale"], "loss_taylor": held["loss_taylor"]},
            "force_rms_rel_err_inner_steps": dict(errs, against="exact model on the GPU (itself within 1e-5 of f64)"),
            "n_pairs_cutoff_exact": n_cut, "ms_per_step_exact_every_step": plain_ms, "ms_per_step_mts": mts_ms,
            "speedup_vs_plain": plain_ms / mts_ms,
            "pair_interactions_per_s_equivalent": n_cut / (mts_ms * 1e-3)}


KB_KCAL = 0.0019872041      # kcal/mol/K
KCAL_PER_AMU = 418.4        # (kcal/mol)/amu in A^2/ps^2


def md_system(k=10, spacing=16.0):
    """k^3 copies of the ala_dipeptide fixture (prmtop read by the host
    layer, bonds to hydrogen promoted to constraints) in a periodic box."""
    from geodesic_hip import host, systems
    fix = os.path.join(ROOT, "tests", "golden", "fixtures")
    t = host.Topology(os.path.join(fix, "ala_dipeptide.prmtop"))
    t.promote_hydrogen_bonds()
    fields = t.as_dict()
    pos, _ = host.read_inpcrd(os.path.join(fix, "ala_dipeptide.inpcrd"), len(fields["sigma"]))
    return systems.replicated_molecule(fields, pos, k, spacing)


def md_velocities(mass, temperature=300.0, seed=7):
    """Seeded Maxwell-Boltzmann velocities (A/ps)."""
    from geodesic_hip import systems
    rng = systems.Rng(seed)
    sd = np.sqrt(KB_KCAL * temperature * KCAL_PER_AMU / mass)
    return [np.ascontiguousarray(rng.normal(len(mass)) * sd) for _ in range(3)]


def bench_md(steps=200, warmup=20, frame_interval=50, cpu_steps=20, cpu=True):
    """Device-resident MD (gm_md_run: the reference run loop, lib.rs:255-330,
    with B / A + SHAKE / rebuild when needed / forces / B / RATTLE on device
    arrays): 1,000 ala_dipeptide molecules, 22,000 atoms, 12,000 constraints,
    160 A periodic box, r_c 12 / r_sw 10 / list 14 A, dt 1 fs, 300 K start.
    Beside it (cpu=True), the oracle's restatement of the same loop on the
    host cores: the reference's CPU path for this row."""
    import geodesic_hip as gh
    x, y, z, sig, eps, mass, topo, box = md_system()
    n, nc = len(x), len(topo.constr_i)
    p = gh.Params(box=(box,) * 3, periodic=True, r_cutoff=12.0, r_switch=10.0, r_skin=14.0)
    g = gh.System(sig, eps, topo, p)
    g.dynamics_set(mass, topo.constr_i, topo.constr_j, topo.constr_dsq)
    vx, vy, vz = md_velocities(mass)
    d = [gh.DeviceArray.from_host(np.ascontiguousarray(a, np.float64)) for a in (x, y, z, vx, vy, vz)]
    g.md_run(*d, n_steps=warmup, dt=0.001, frame_interval=warmup, max_iter=100, tol=1e-8)
    t0 = time.perf_counter()
    frames, epot, ekin = g.md_run(*d, n_steps=steps, dt=0.001, frame_interval=frame_interval, max_iter=100,
                                  tol=1e-8)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    etot = epot + ekin
    out = {"workload": "device-resident MD, 1,000 ala_dipeptide copies (22,000 atoms, 12,000 H-bond constraints), "
                       "160 A periodic box, r_c 12 / r_sw 10 / list 14 A, dt 1 fs, 300 K start",
           "n_atoms": n, "n_constraints": nc, "steps": steps, "ms_per_step": ms, "steps_per_s": 1e3 / ms,
           "frames": int(len(frames)), "total_energy_first_last": [float(etot[0]), float(etot[-1])]}
    if cpu:
        out["cpu_baseline"] = md_cpu_baseline(x, y, z, vx, vy, vz, sig, eps, mass, topo, box, cpu_steps)
        out["speedup_vs_cpu"] = out["cpu_baseline"]["ms_per_step"] / ms
    return out


def md_cpu_baseline(x, y, z, vx, vy, vz, sig, eps, mass, topo, box, steps):
    """The oracle's restatement of the reference loop (integrator.rs,
    constraint.rs, cpu_backend.rs with T static strips) on the same system:
    ms per step over `steps` steps after the first list build and forces."""
    from oracle import oracle as orc
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    x, y, z, vx, vy, vz = (a.copy() for a in (x, y, z, vx, vy, vz))
    bx = (box,) * 3
    rc, rs, rsw, dt, mi, tol = 12.0, 14.0, 10.0, 0.001, 100, 1e-8
    orc.wrap_positions(x, y, z, bx)
    pi, pj = orc.neighbor_build(x, y, z, bx, rs, topo.excl_i, topo.excl_j, method="cells")
    rx, ry, rz = x.copy(), y.copy(), z.copy()
    t, fx, fy, fz = orc.cpu_backend_compute(x, y, z, sig, eps, topo, pi, pj, bx, rc, rsw, n_threads=cores,
                                            periodic=True)
    orc.constrain_velocities(x, y, z, vx, vy, vz, topo, mass, mi, tol)
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.half_kick(vx, vy, vz, fx, fy, fz, mass, dt / 2)
        orc.drift_and_constrain(x, y, z, vx, vy, vz, topo, mass, dt, mi, tol)
        if orc.needs_rebuild(x, y, z, rx, ry, rz, rs, rc):
            orc.wrap_positions(x, y, z, bx)
            pi, pj = orc.neighbor_build(x, y, z, bx, rs, topo.excl_i, topo.excl_j, method="cells")
            rx, ry, rz = x.copy(), y.copy(), z.copy()
        t, fx, fy, fz = orc.cpu_backend_compute(x, y, z, sig, eps, topo, pi, pj, bx, rc, rsw,
                                                n_threads=cores, periodic=True)
        orc.half_kick(vx, vy, vz, fx, fy, fz, mass, dt / 2)
        orc.constrain_velocities(x, y, z, vx, vy, vz, topo, mass, mi, tol)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return {"ms_per_step": ms, "cores": cores, "kind": "port",
            "sample": f"{steps} steps of the same loop after the first list build and forces"}


def pruning_study(g, s, x, y, z, e, args):
    """Tile pruning skips tiles that lie beyond r_c at the current positions,
    bounded by the largest displacement since the build: the same evaluation
    with pruning off, and at positions drifted as after a whole 20-step
    rebuild interval of 0.02 A Gaussian moves per coordinate (the 250k
    workload's protocol), where the bound has grown the most."""
    g.set_tile_pruning(False)
    ms_off, prof_off, e_off, _ = time_resident(g, args.steps, args.warmup)
    g.set_tile_pruning(True)
    assert e_off.total == e.total
    rng = np.random.default_rng(20)
    d = rng.normal(0.0, 0.02 * 20 ** 0.5, (3, s.n))
    g.set_positions(x + d[0], y + d[1], z + d[2])
    ms_drift, prof_drift, _, _ = time_resident(g, args.steps, args.warmup)
    g.set_tile_pruning(False)
    ms_drift_off, _, _, _ = time_resident(g, args.steps, args.warmup)
    g.set_tile_pruning(True)
    g.set_positions(x, y, z)
    return {"ms_per_step_pruning_off": ms_off,
            "kernel_ms_pruning_off": prof_off["ms_nonbonded_total"] / max(prof_off["n_evals"], 1),
            "drift_max_displacement_A": float(np.sqrt((d ** 2).sum(axis=0)).max()),
            "ms_per_step_drifted": ms_drift,
            "kernel_ms_drifted": prof_drift["ms_nonbonded_total"] / max(prof_drift["n_evals"], 1),
            "ms_per_step_drifted_pruning_off": ms_drift_off}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--chunk", type=int, default=-1, help="entries per work item (tuning, tile layout)")
    ap.add_argument("--layout", choices=("auto", "partners", "tiles"), default="auto",
                    help="nonbonded list layout (gm_set_pair_layout)")
    ap.add_argument("--no-250k", action="store_true")
    ap.add_argument("--no-reference-workloads", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-md", action="store_true")
    ap.add_argument("--no-pruning-study", action="store_true",
                    help="skip the pruning-off and drifted-position timings (profiling)")
    ap.add_argument("--only", choices=("250k", "batched", "surrogate", "md", "reference", "build"), default=None,
                    help="run one secondary section alone and print its JSON (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the force path would shard by replica or by spatial domain across processes)", file=sys.stderr)
        sys.exit(2)

    if args.only == "250k":
        print(json.dumps(bench_250k(layout=args.layout)))
        return
    if args.only == "batched":
        print(json.dumps(bench_batched(layout=args.layout)))
        return
    if args.only == "surrogate":
        print(json.dumps(bench_surrogate()))
        return
    if args.only == "md":
        print(json.dumps(bench_md(cpu=not args.no_cpu_baseline)))
        return
    if args.only == "reference":
        print(json.dumps(bench_reference_workloads(cpu=not args.no_cpu_baseline)))
        return
    if args.only == "build":
        print(json.dumps(bench_build(layout=args.layout)))
        return

    import geodesic_hip as gh
    from geodesic_hip import systems

    s = systems.system_24k()
    g = gh.System(s.sigma, s.epsilon, s.topo, s.params)
    g.set_pair_layout(args.layout)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    g.build_neighbor_list(x, y, z)
    first_build_ms = g.stats()["ms_build"]
    builds = []
    for _ in range(3):  # steady-state build time (the first one also allocates)
        g.build_neighbor_list(x, y, z)
        builds.append(g.stats()["ms_build"])
    build_wall = build_wall_ms(g, x, y, z)
    if args.chunk >= 0:
        g.set_work_chunk(args.chunk)
    st = g.stats()
    n_cut, n_list = st["n_pairs_cutoff"], st["n_list_pairs"]

    ms_step, prof, e, ms_all_events = time_resident(g, args.steps, args.warmup)
    # PCIe-inclusive rate of the host-buffer entry point (positions in, forces out)
    hx, hy, hz = s.x.copy(), s.y.copy(), s.z.copy()
    for _ in range(5):
        g.compute_forces(hx, hy, hz)
    n_host = 100
    t0 = time.perf_counter()
    for _ in range(n_host):
        g.compute_forces(hx, hy, hz)
    ms_host = (time.perf_counter() - t0) * 1e3 / n_host
    ev_ms = prof["ms_eval_total"] / max(prof["n_evals"], 1)
    prep_ms = prof["ms_prep_total"] / max(prof["n_evals"], 1)
    roof = roofline_pass(g, n_cut)
    pruning = None if args.no_pruning_study else pruning_study(g, s, x, y, z, e, args)

    out = {
        "metric": "pair-interactions/s (24k-atom force evaluation)",
        "value": n_cut / (ms_step * 1e-3),
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair math, int64 fixed-point accumulation, f64 bonded",
        "data": "synthetic (seeded jittered lattice, 3000-atom chain + 7000 triplets)",
        "config": {"workload": "force evaluation, 24,000 atoms, 62 A periodic box, r_c 9 / r_sw 7.5 / list 10.5 A",
                   "n_atoms": s.n, "n_list_pairs": n_list, "n_pairs_cutoff": n_cut,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "compute": "fp32 VALU (no matrix work in this kernel)",
                     "achieved": roof["achieved"], "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": roof["frac"],
                     "traffic": traffic_from_profiles(TRAFFIC_FILE[layout_name(st)]),
                     "kernel": FORCE_KERNEL[layout_name(st)], "kernel_ms_avg": roof["kernel_ms_avg"],
                     "kernel_ms_min": roof["kernel_ms_min"], "flop_per_pair": FLOP_PER_PAIR,
                     "timed_launches": roof["timed_launches"],
                     "timing": f"dispatch-stamped events on each of {ROOF_LAUNCHES} launches after the timed region"},
        "device_ms_per_eval": ev_ms,
        "ms_to_force_kernel_start": prep_ms,
        "tile_pruning": pruning,
        "ms_per_step_events_every_step": ms_all_events,
        "ms_per_eval_host_buffers": ms_host,
        "list_build_ms": statistics.median(builds),
        "list_build_ms_first": first_build_ms,
        "list_build_ms_wall": build_wall,
        "list": {k: st[k] for k in ("n_superclusters", "n_entries", "n_tiles", "n_tiles_cutoff",
                                    "n_work_items", "n_partner_entries", "pair_layout", "n_lj_types")},
        "energy_total": e.total,
    }
    if not args.no_reference_workloads:
        out["reference_benchmarks"] = bench_reference_workloads(cpu=not args.no_cpu_baseline)
    if not args.no_250k:
        out["system_250k"] = bench_250k(layout=args.layout)
    if not args.no_batched:
        bb = bench_batched(layout=args.layout)
        out["batched"] = bb
    if not args.no_surrogate:
        out["surrogate"] = bench_surrogate()
    if not args.no_md:
        out["md"] = bench_md(cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(s, n_cut)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
