"""gm_compute_forces_device_noalloc (no reference counterpart; for callers
that cannot allocate): through the C-ABI, byte-identical to
gm_compute_forces_device on two calls in a row, the second on a smaller
system, with the static scratch array keeping the first call's contents in
between; a list that needs more scratch than the array holds is refused."""
import sys

import numpy as np
import pytest

import geodesic_hip as gh
from geodesic_hip import systems

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def built(s, layout=None):
    g = gh.System(s.sigma, s.epsilon, s.topo, s.params)
    if layout is not None:
        g.set_pair_layout(layout)
    x, y, z = s.x.copy(), s.y.copy(), s.z.copy()
    g.build_neighbor_list(x, y, z)
    return g, [gh.DeviceArray.from_host(a) for a in (x, y, z)]


def both_ways(g, d, n):
    """(forces, energies) of the allocation-free entry point, then of
    gm_compute_forces_device, on the same device positions."""
    out = []
    for fn in (g.compute_forces_device_noalloc, g.compute_forces_device):
        f = [gh.DeviceArray(n) for _ in range(3)]
        fn(*(a.ptr for a in d), *(a.ptr for a in f))
        e = g.sync()
        out.append(([a.to_host() for a in f], e))
    return out


@pytest.mark.parametrize("layout", ["partners", "tiles"])
def test_noalloc_equals_the_allocating_entry_point_twice(layout):
    cases = [systems.system_24k(), systems.protein_solvent(3000, (3000 / 0.1) ** (1 / 3), 375, seed=7)]
    for s in cases:
        say(f"[gpu] noalloc: {s.n} atoms, {layout}")
        g, d = built(s, layout)
        (fa, ea), (fb, eb) = both_ways(g, d, s.n)
        for u, v in zip(fa, fb):
            assert u.tobytes() == v.tobytes()
        for name in ("bond", "angle", "dihedral", "lj", "lj14", "total"):
            assert getattr(ea, name) == getattr(eb, name), name
        assert np.isfinite(fa[0]).all()
        g.close()


def test_noalloc_records_mode_and_refusal():
    """On the tile layout, small work items keep per-block i-force records in
    the array too; a 250k list with such records needs more than 32 MiB and is
    refused before any launch, and still evaluates with atomics (work items of
    32)."""
    s = systems.system_24k()
    g, d = built(s, "tiles")
    for chunk in (8, 32):
        g.set_work_chunk(chunk)
        (fa, ea), (fb, eb) = both_ways(g, d, s.n)
        for u, v in zip(fa, fb):
            assert u.tobytes() == v.tobytes()
        assert ea.total == eb.total
    big = systems.system_250k()
    g2, d2 = built(big, "tiles")
    g2.set_work_chunk(8)
    f = [gh.DeviceArray(big.n) for _ in range(3)]
    with pytest.raises(gh.GmError) as ei:
        g2.compute_forces_device_noalloc(*(a.ptr for a in d2), *(a.ptr for a in f))
    assert ei.value.status == 8 and "32" in str(ei.value)
    g2.set_work_chunk(32)
    (fa, ea), (fb, eb) = both_ways(g2, d2, big.n)
    for u, v in zip(fa, fb):
        assert u.tobytes() == v.tobytes()
    assert ea.total == eb.total
