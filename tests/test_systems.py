"""The seeded workload generators: determinism, sizes, topology, geometry."""
import numpy as np

from geodesic_hip import systems


def test_splitmix64_known_values():
    # splitmix64 stream from seed 0 (Vigna's reference implementation)
    r = systems.Rng(0)
    assert [int(v) for v in r.u64(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_generator_is_deterministic():
    a = systems.protein_solvent(3000, 31.07, 375, seed=5)
    b = systems.protein_solvent(3000, 31.07, 375, seed=5)
    for k in ("x", "y", "z", "sigma", "epsilon"):
        assert getattr(a, k).tobytes() == getattr(b, k).tobytes()
    assert a.topo.dihed_n.tobytes() == b.topo.dihed_n.tobytes()


def test_24k_system_shape():
    s = systems.system_24k()
    assert s.n == 24000 and s.params.box == (62.0, 62.0, 62.0)
    assert len(s.topo.bond_i) == 2999 and len(s.topo.angle_i) == 2998 and len(s.topo.dihed_i) == 2997
    assert len(s.topo.pair14_i) == 2997 and s.params.scale_14 == 0.5
    assert set(np.unique(s.topo.dihed_n)) == {1, 2, 3}
    assert len(np.unique(s.types)) == 16
    assert 2.0 <= s.sigma.min() and s.sigma.max() <= 3.8
    assert 0.01 <= s.epsilon.min() and s.epsilon.max() <= 0.25
    for v in (s.x, s.y, s.z):
        assert v.min() >= 0.0 and v.max() < 62.0
    # the chain crosses the periodic boundary: some raw bond length exceeds half the box
    d = np.abs(s.x[s.topo.bond_j] - s.x[s.topo.bond_i])
    assert d.max() > 31.0
    # 7,000 rigid triplets: 3 exclusions each, plus 1-2/1-3/1-4 along the chain
    assert len(s.topo.excl_i) == 2999 + 2998 + 2997 + 3 * 7000


def test_replica_sizes_are_log_uniform_in_range():
    n = systems.replica_sizes(512)
    assert n.min() >= 500 and n.max() <= 20000
    assert abs(np.median(np.log(n)) - 0.5 * (np.log(500) + np.log(20000))) < 0.3
