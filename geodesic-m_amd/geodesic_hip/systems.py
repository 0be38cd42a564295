"""Seeded synthetic systems for the force-evaluation workloads.

The reference commits no workload at these sizes, so these generators define
them (their only randomness is a counter-based splitmix64 stream, so a seed
gives the same system on every platform):

  protein_solvent(24000, 62.0, 3000)    24k atoms, 62 A periodic box
  protein_solvent(250000, 136.0, 30000) 250k atoms, 136 A periodic box
  replica_set(512)                      log-uniform 500..20000-atom replicas
  lj_lattice(10000)                     the reference's own benchmark system
                                        (geodesic/benches/benchmarks.rs:66-95)

Layout: a jittered cubic lattice (about 0.1 atoms/A^3, jitter <= 0.3 A per
coordinate) walked in a space-filling snake.  The first n_chain sites form a
bonded chain (bonds, angles, proper dihedrals of periodicity 1-3), the rest
are 3-site rigid triplets.  1-2/1-3/1-4 chain pairs and intra-triplet pairs are
excluded from the Verlet list; chain 1-4 pairs carry LJ scaled by 1/2.  The
whole system is shifted so the chain crosses the periodic boundary, so the
bonded terms need minimum-image displacements.  16 LJ types with sigma
U[2.0, 3.8] A and epsilon U[0.01, 0.25] kcal/mol, Lorentz-Berthelot combining.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import Params

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Rng:
    """Counter-based splitmix64 stream: value k = mix(seed + (k+1) * golden)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.ctr = 0

    def u64(self, n):
        idx = np.arange(self.ctr + 1, self.ctr + n + 1, dtype=np.uint64)
        self.ctr += n
        with np.errstate(over="ignore"):
            return splitmix64(self.seed + idx * _GOLDEN)

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def integers(self, n, hi):
        return np.minimum((self.uniform(n) * hi).astype(np.int64), hi - 1)

    def normal(self, n):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)

    def permutation(self, n):
        return np.argsort(self.u64(n), kind="stable")


@dataclass
class Topology:
    """SoA bonded topology with the field names of geodesic-core BondedTopology
    (topology.rs:4-37) plus the scaled 1-4 pair list."""
    bond_i: np.ndarray
    bond_j: np.ndarray
    bond_k: np.ndarray
    bond_r0: np.ndarray
    angle_i: np.ndarray
    angle_j: np.ndarray
    angle_k: np.ndarray
    angle_kth: np.ndarray
    angle_th0: np.ndarray
    dihed_i: np.ndarray
    dihed_j: np.ndarray
    dihed_k: np.ndarray
    dihed_l: np.ndarray
    dihed_kphi: np.ndarray
    dihed_n: np.ndarray
    dihed_delta: np.ndarray
    excl_i: np.ndarray
    excl_j: np.ndarray
    constr_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    constr_j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    constr_dsq: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float64))
    pair14_i: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    pair14_j: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))


@dataclass
class SyntheticSystem:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    sigma: np.ndarray
    epsilon: np.ndarray
    types: np.ndarray
    topo: Topology
    params: Params
    seed: int
    name: str = ""

    @property
    def n(self):
        return len(self.x)


def _snake_sites(m):
    """All m^3 lattice sites ordered so consecutive sites are lattice neighbours."""
    out = []
    row = 0
    for k in range(m):
        js = range(m) if k % 2 == 0 else range(m - 1, -1, -1)
        for j in js:
            is_ = range(m) if row % 2 == 0 else range(m - 1, -1, -1)
            out.extend((i, j, k) for i in is_)
            row += 1
    return np.array(out, dtype=np.int64)


def protein_solvent(n_atoms: int, box: float, n_chain: int, seed: int = 1, n_types: int = 16,
                    r_cutoff=9.0, r_switch=7.5, r_skin=10.5, name="") -> SyntheticSystem:
    rng = Rng(seed)
    m = int(np.ceil(round(n_atoms ** (1.0 / 3.0), 9)))
    while m ** 3 < n_atoms:
        m += 1
    a = box / m
    sites = _snake_sites(m)
    n_chain = min(n_chain, n_atoms)
    # vacancies only after the chain, chosen uniformly among the remaining sites
    n_vac = m ** 3 - n_atoms
    tail = np.arange(n_chain, m ** 3)
    vac = np.sort(tail[rng.permutation(len(tail))[:n_vac]])
    keep = np.ones(m ** 3, bool)
    keep[vac] = False
    sites = sites[keep]
    jit = rng.uniform(3 * n_atoms, -0.3, 0.3).reshape(3, n_atoms)
    shift = np.array([0.37, 0.41, 0.43]) * box
    xyz = (sites.T.astype(np.float64) + 0.5) * a + jit + shift[:, None]
    xyz = xyz - box * np.floor(xyz / box)
    x, y, z = (np.ascontiguousarray(v) for v in xyz)

    tsig = rng.uniform(n_types, 2.0, 3.8)
    teps = rng.uniform(n_types, 0.01, 0.25)
    types = rng.integers(n_atoms, n_types).astype(np.int32)
    sigma, epsilon = tsig[types].copy(), teps[types].copy()

    u32 = lambda v: np.ascontiguousarray(v, dtype=np.uint32)  # noqa: E731
    c = np.arange(n_chain)
    b_i, b_j = c[:-1], c[1:]
    a_i, a_j, a_k = c[:-2], c[1:-1], c[2:]
    d_i, d_j, d_k, d_l = c[:-3], c[1:-2], c[2:-1], c[3:]
    # equilibrium angles: the undisplaced lattice angle (90 or 180 degrees)
    s = sites[:n_chain].astype(np.float64)
    v1 = s[a_i] - s[a_j]
    v2 = s[a_k] - s[a_j]
    cos0 = np.sum(v1 * v2, axis=1) / (np.linalg.norm(v1, axis=1) * np.linalg.norm(v2, axis=1))
    th0 = np.arccos(np.clip(cos0, -1.0, 1.0))
    # 180-degree equilibria make sin(theta) vanish at rest; use 170 degrees there
    th0 = np.where(th0 > 3.0, np.deg2rad(170.0), th0)
    nd = len(d_i)
    dn = (1 + rng.integers(nd, 3)).astype(np.uint32)
    ddelta = np.where(rng.uniform(nd) < 0.5, 0.0, np.pi)
    dk = rng.uniform(nd, 0.2, 2.0)

    trip = np.arange(n_chain, n_atoms)
    n_trip = len(trip) // 3
    t0, t1, t2 = trip[0:3 * n_trip:3], trip[1:3 * n_trip:3], trip[2:3 * n_trip:3]
    ex_i = np.concatenate([b_i, a_i, d_i, t0, t0, t1])
    ex_j = np.concatenate([b_j, a_k, d_l, t1, t2, t2])
    topo = Topology(
        bond_i=u32(b_i), bond_j=u32(b_j), bond_k=np.full(len(b_i), 300.0), bond_r0=np.full(len(b_i), a),
        angle_i=u32(a_i), angle_j=u32(a_j), angle_k=u32(a_k), angle_kth=np.full(len(a_i), 50.0),
        angle_th0=np.ascontiguousarray(th0),
        dihed_i=u32(d_i), dihed_j=u32(d_j), dihed_k=u32(d_k), dihed_l=u32(d_l), dihed_kphi=dk,
        dihed_n=dn, dihed_delta=ddelta,
        excl_i=u32(ex_i), excl_j=u32(ex_j),
        constr_i=u32(np.concatenate([t0, t0, t1])), constr_j=u32(np.concatenate([t1, t2, t2])),
        constr_dsq=np.full(3 * n_trip, a * a),
        pair14_i=u32(d_i), pair14_j=u32(d_l),
    )
    params = Params(box=(box, box, box), periodic=True, r_cutoff=r_cutoff, r_switch=r_switch,
                    r_skin=r_skin, scale_14=0.5, bonded_min_image=True)
    return SyntheticSystem(x, y, z, sigma, epsilon, types, topo, params, seed, name)


def system_24k(seed=1):
    """24,000 atoms, 62 A box, 3,000-atom chain; r_c 9, r_sw 7.5, list 10.5 A."""
    return protein_solvent(24000, 62.0, 3000, seed=seed, name="24k")


def system_250k(seed=2):
    """250,000 atoms, 136 A box, 30,000-atom chain; r_c 9, r_sw 7.5, list 10.5 A."""
    return protein_solvent(250000, 136.0, 30000, seed=seed, name="250k")


def replica_sizes(n_replicas=512, lo=500, hi=20000, seed=3):
    """Log-uniform atom counts in [lo, hi]."""
    rng = Rng(seed)
    u = rng.uniform(n_replicas)
    return np.round(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


def replica_set(n_replicas=512, lo=500, hi=20000, seed=3, density=0.1):
    """Independent replicas generated like the 24k system at their own size and
    box L = (N / density)^(1/3); 1/8 of each replica's atoms form its chain."""
    sizes = replica_sizes(n_replicas, lo, hi, seed)
    out = []
    for k, n in enumerate(sizes):
        box = (n / density) ** (1.0 / 3.0)
        out.append(protein_solvent(int(n), float(box), int(n) // 8, seed=seed * 100003 + k,
                                   name=f"replica{k}"))
    return out


_TOPO_INDEX = ("bond_i", "bond_j", "angle_i", "angle_j", "angle_k", "dihed_i", "dihed_j", "dihed_k", "dihed_l",
               "excl_i", "excl_j", "constr_i", "constr_j", "pair14_i", "pair14_j")
_TOPO_VALUE = ("bond_k", "bond_r0", "angle_kth", "angle_th0", "dihed_kphi", "dihed_n", "dihed_delta",
               "constr_dsq")


def replicated_molecule(fields, pos, k, spacing):
    """k^3 copies of one molecule on a cubic grid of the given spacing (A),
    in a periodic box of edge k * spacing.  `fields`: the molecule's AtomData
    and BondedTopology arrays by field name (sigma, epsilon, mass and the
    topology fields, as the host layer's prmtop reader gives them); `pos`:
    its positions, 3 x n.  Returns (x, y, z, sigma, epsilon, mass, Topology,
    box).  Atom indices of copy c are the molecule's plus c n; every copy sits
    inside its own cell, so no bonded term crosses the box boundary."""
    n = len(fields["sigma"])
    nc = k ** 3
    p = np.asarray(pos, np.float64)
    centre = 0.5 * (p.min(axis=1) + p.max(axis=1))
    g = np.stack(np.meshgrid(np.arange(k), np.arange(k), np.arange(k), indexing="ij"), -1).reshape(-1, 3)
    shift = (g + 0.5) * spacing - centre  # (nc, 3)
    x, y, z = (np.ascontiguousarray((p[a][None, :] + shift[:, a:a + 1]).reshape(-1)) for a in range(3))
    rep = lambda v: np.ascontiguousarray(np.tile(np.asarray(v), nc))  # noqa: E731
    off = lambda v: np.ascontiguousarray(  # noqa: E731
        (np.asarray(v, np.int64)[None, :] + (np.arange(nc, dtype=np.int64) * n)[:, None]).reshape(-1).astype(np.uint32))
    t = {}
    for name in _TOPO_INDEX:
        t[name] = off(fields.get(name, np.zeros(0, np.uint32)))
    for name in _TOPO_VALUE:
        v = fields.get(name, np.zeros(0))
        t[name] = rep(v).astype(np.uint32 if name == "dihed_n" else np.float64)
    topo = Topology(**t)
    return (x, y, z, rep(fields["sigma"]).astype(np.float64), rep(fields["epsilon"]).astype(np.float64),
            rep(fields["mass"]).astype(np.float64), topo, float(k * spacing))


def _empty_topology():
    z32, zf = np.zeros(0, np.uint32), np.zeros(0, np.float64)
    return Topology(z32, z32, zf, zf, z32, z32, z32, zf, zf, z32, z32, z32, z32, zf, z32, zf, z32, z32)


def lj_lattice(n=10000):
    """The workload of the reference's bench_lj_inner_loop and
    bench_neighbor_rebuild (geodesic/benches/benchmarks.rs:49-95; paths
    relative to the reference repository root): n atoms on a cubic lattice of
    2.0 A spacing, side ceil(n^(1/3)) sites, filled x-major then y then z
    (the last plane partly), in a periodic box of side * 2.0 A; sigma 3.0,
    epsilon 0.1 for every atom; no topology; r_cutoff 5, r_skin 6, r_switch 4
    (SimParams of lj_params).  Lattice distances hit the thresholds exactly:
    pairs at r = 6.0 A (r2 == r_skin^2, in the list) and r = 4.0 A (r ==
    r_switch, unswitched)."""
    side = int(np.ceil(round(n ** (1.0 / 3.0), 12)))
    ix, iy, iz = np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij")
    sites = np.stack([ix.reshape(-1), iy.reshape(-1), iz.reshape(-1)], 1)[:n]
    x, y, z = (np.ascontiguousarray(sites[:, a].astype(np.float64) * 2.0) for a in range(3))
    box = side * 2.0
    params = Params(box=(box, box, box), periodic=True, r_cutoff=5.0, r_switch=4.0, r_skin=6.0)
    return SyntheticSystem(x, y, z, np.full(n, 3.0), np.full(n, 0.1), np.zeros(n, np.int64), _empty_topology(),
                           params, 0, "lj_lattice")
