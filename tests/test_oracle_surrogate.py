"""Pins of the f64 surrogate oracle (oracle/gm_sur_oracle.c), CPU only.

The surrogate has no reference counterpart, so its oracle is pinned by
mathematics: central finite differences of the loss for the hand-derived
gradient, a closed form and a direct numpy evaluation for the features, and
symmetries (permutation, translation) the definitions imply."""
import numpy as np
import pytest

from oracle import oracle as orc

NP = 29315
OFF_W1, OFF_B1 = 0, 128 * 96
OFF_W2 = OFF_B1 + 128
OFF_B2 = OFF_W2 + 128 * 128
OFF_W3 = OFF_B2 + 128
OFF_B3 = OFF_W3 + 3 * 128


def rand_params(rng, scale=1.0):
    p = np.zeros(NP, np.float32)
    p[OFF_W1:OFF_B1] = rng.uniform(-1, 1, 128 * 96) * np.sqrt(6 / 224) * scale
    p[OFF_B1:OFF_W2] = rng.uniform(-0.1, 0.1, 128)
    p[OFF_W2:OFF_B2] = rng.uniform(-1, 1, 128 * 128) * np.sqrt(6 / 256) * scale
    p[OFF_B2:OFF_W3] = rng.uniform(-0.1, 0.1, 128)
    p[OFF_W3:OFF_B3] = rng.uniform(-1, 1, 3 * 128) * 0.3
    p[OFF_B3:] = rng.uniform(-0.1, 0.1, 3)
    return p


def rand_problem(rng, n):
    feat = rng.gamma(2.0, 2.0, (n, 96)).astype(np.float32)
    fmean = feat.mean(axis=0).astype(np.float32)
    finv = (1.0 / np.sqrt(feat.var(axis=0) + 1e-4)).astype(np.float32)
    basis = rng.normal(size=(n, 9)).astype(np.float32)
    target = (basis[:, 0:3] * 1.1 + 0.2 * basis[:, 3:6] + 0.05 * rng.normal(size=(n, 3))).astype(np.float32)
    return feat, fmean, finv, basis, target


def np_forward(feat, fmean, finv, p):
    x = ((feat - fmean) * finv).astype(np.float32).astype(np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(128, 96).astype(np.float64)
    W2 = p[OFF_W2:OFF_B2].reshape(128, 128).astype(np.float64)
    W3 = p[OFF_W3:OFF_B3].reshape(3, 128).astype(np.float64)
    silu = lambda z: z / (1.0 + np.exp(-z))  # noqa: E731
    h1 = silu(x @ W1.T + p[OFF_B1:OFF_W2])
    h2 = silu(h1 @ W2.T + p[OFF_B2:OFF_W3])
    return h2 @ W3.T + p[OFF_B3:]


def test_forward_matches_numpy():
    rng = np.random.default_rng(1)
    feat, fmean, finv, _, _ = rand_problem(rng, 17)
    p = rand_params(rng)
    o = orc.sur_forward(feat, fmean, finv, p, round_bf16=False)
    ref = np_forward(feat, fmean, finv, p)
    assert np.allclose(o, ref, rtol=1e-12, atol=1e-12)


def bf16_np(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_forward_bf16_rounding_points():
    """Rounding x, weights, h1, h2 to BF16 (nearest even) reproduces the
    oracle's rounded forward pass."""
    rng = np.random.default_rng(2)
    feat, fmean, finv, _, _ = rand_problem(rng, 9)
    p = rand_params(rng)
    o = orc.sur_forward(feat, fmean, finv, p, round_bf16=True)
    x = bf16_np(((feat - fmean) * finv).astype(np.float32))
    W1 = bf16_np(p[OFF_W1:OFF_B1]).reshape(128, 96)
    W2 = bf16_np(p[OFF_W2:OFF_B2]).reshape(128, 128)
    W3 = bf16_np(p[OFF_W3:OFF_B3]).reshape(3, 128)
    silu = lambda z: z / (1.0 + np.exp(-z))  # noqa: E731
    h1 = bf16_np(silu(x @ W1.T + p[OFF_B1:OFF_W2]))
    h2 = bf16_np(silu(h1 @ W2.T + p[OFF_B2:OFF_W3]))
    ref = h2 @ W3.T + p[OFF_B3:]
    assert np.allclose(o, ref, rtol=1e-9, atol=1e-9)
    # and rounding matters: the unrounded pass differs
    assert not np.allclose(o, orc.sur_forward(feat, fmean, finv, p, round_bf16=False), rtol=1e-6, atol=0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    feat, fmean, finv, basis, target = rand_problem(rng, 6)
    p = rand_params(rng)
    loss, g = orc.sur_loss_grad(feat, fmean, finv, p, basis, target, round_bf16=False)
    idx = list(rng.choice(np.arange(OFF_W1, OFF_B1), 12, replace=False))
    idx += list(rng.choice(np.arange(OFF_B1, OFF_W2), 4, replace=False))
    idx += list(rng.choice(np.arange(OFF_W2, OFF_B2), 12, replace=False))
    idx += list(rng.choice(np.arange(OFF_B2, OFF_W3), 4, replace=False))
    idx += list(rng.choice(np.arange(OFF_W3, OFF_B3), 6, replace=False)) + [OFF_B3, OFF_B3 + 1, OFF_B3 + 2]
    checked = 0
    for k in idx:
        h = 1e-3 * max(1.0, abs(float(p[k])))
        pp, pm = p.copy(), p.copy()
        pp[k] = np.float32(p[k] + h)
        pm[k] = np.float32(p[k] - h)
        dh = float(pp[k]) - float(pm[k])  # the perturbation float32 actually represents
        lp, _ = orc.sur_loss_grad(feat, fmean, finv, pp, basis, target, round_bf16=False)
        lm, _ = orc.sur_loss_grad(feat, fmean, finv, pm, basis, target, round_bf16=False)
        fd = (lp - lm) / dh
        assert abs(fd - g[k]) <= 1e-5 * max(abs(g[k]), 1e-3 * np.abs(g).max()), (k, fd, g[k])
        checked += 1
    assert checked == len(idx)


def test_loss_definition_and_zero_output():
    """With W3 = b3 = 0 the outputs are zero and the loss is |B1 - T|^2 averaged."""
    rng = np.random.default_rng(4)
    feat, fmean, finv, basis, target = rand_problem(rng, 11)
    p = rand_params(rng)
    p[OFF_W3:] = 0.0
    loss, g = orc.sur_loss_grad(feat, fmean, finv, p, basis, target)
    ref = np.mean(np.sum((basis[:, 0:3].astype(np.float64) - target) ** 2, axis=1))
    assert abs(loss - ref) <= 1e-12 * abs(ref)
    # gradient of b3 at o = 0: (2/n) sum_i r_i . B_k
    r = basis[:, 0:3].astype(np.float64) - target
    for k in range(3):
        gk = 2.0 * np.mean(np.sum(r * basis[:, 3 * k:3 * k + 3], axis=1))
        assert abs(g[OFF_B3 + k] - gk) <= 1e-9 * max(1.0, abs(gk))


def np_features(x, y, z, box, cls, r_lo=7.5, r_hi=12.0, fade=0.5):
    n = len(x)
    P = np.stack([x, y, z], axis=1)
    d = P[None, :, :] - P[:, None, :]
    d -= np.asarray(box) * np.round(d / np.asarray(box))
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, -1.0)
    w = (r_hi - r_lo) / 23
    u = lambda v: np.clip(v, 0, 1) ** 3 * (10 - 15 * np.clip(v, 0, 1) + 6 * np.clip(v, 0, 1) ** 2)  # noqa: E731
    inside = (r >= r_lo) & (r <= r_hi)
    env = np.where(inside, u((r - r_lo) / fade) * u((r_hi - r) / fade), 0.0)
    out = np.zeros((n, 96))
    t = (r - r_lo) / w
    for k in range(24):
        g = env * np.exp(-0.5 * (t - k) ** 2)
        for c in range(4):
            out[:, c * 24 + k] = (g * (cls[None, :] == c)).sum(axis=1)
    return out


def test_features_match_numpy():
    rng = np.random.default_rng(5)
    n, L = 80, 30.0
    x, y, z = (rng.uniform(0, L, n) for _ in range(3))
    cls = rng.integers(0, 4, n).astype(np.uint8)
    f = orc.sur_features(x, y, z, (L, L, L), cls)
    ref = np_features(x, y, z, (L, L, L), cls)
    assert np.allclose(f, ref, rtol=1e-12, atol=1e-13)
    assert f.max() > 0.5


def test_features_closed_form_pair():
    """Two atoms 9.3 A apart, classes 1 and 3: atom 0's class-3 features are
    exp(-(t - k)^2 / 2), t = (9.3 - 7.5) / w (no fade inside [8, 11.5])."""
    r = 9.3
    x, y, z = np.array([1.0, 1.0 + r]), np.array([2.0, 2.0]), np.array([3.0, 3.0])
    f = orc.sur_features(x, y, z, (100.0, 100.0, 100.0), np.array([1, 3], np.uint8))
    w = 4.5 / 23
    t = (r - 7.5) / w
    k = np.arange(24)
    assert np.allclose(f[0, 72:96], np.exp(-0.5 * (t - k) ** 2), rtol=1e-13, atol=1e-15)
    assert np.allclose(f[1, 24:48], np.exp(-0.5 * (t - k) ** 2), rtol=1e-13, atol=1e-15)
    assert np.count_nonzero(f[0, :72]) == 0 and np.count_nonzero(f[1, :24]) == 0
    # beyond the window: nothing
    f2 = orc.sur_features(np.array([0.0, 12.5]), np.zeros(2), np.zeros(2), (100.0,) * 3, np.array([0, 0], np.uint8))
    assert np.count_nonzero(f2) == 0


def test_features_invariances():
    rng = np.random.default_rng(6)
    n, L = 60, 28.0
    x, y, z = (rng.uniform(0, L, n) for _ in range(3))
    cls = rng.integers(0, 4, n).astype(np.uint8)
    f = orc.sur_features(x, y, z, (L,) * 3, cls)
    perm = rng.permutation(n)
    fp = orc.sur_features(x[perm], y[perm], z[perm], (L,) * 3, cls[perm])
    assert np.allclose(fp, f[perm], rtol=1e-12, atol=1e-13)
    fs = orc.sur_features(x + 7.3, y - 11.1, z + 30.0, (L,) * 3, cls)  # translated, partly out of the box
    assert np.allclose(fs, f, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("n", [1, 2])
def test_features_tiny(n):
    f = orc.sur_features(np.zeros(n), np.zeros(n), np.zeros(n), (50.0,) * 3, np.zeros(n, np.uint8))
    assert f.shape == (n, 96) and np.count_nonzero(f) == 0


def shell_system(rng, n=120, L=26.0, dmin=2.6):
    """Random atoms at least dmin apart, 16-type-like LJ parameters."""
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, L, 3)
        if all(np.linalg.norm((p - q) - L * np.round((p - q) / L)) >= dmin for q in pts):
            pts.append(p)
    P = np.array(pts)
    sig = rng.uniform(2.0, 3.8, n)
    eps = rng.uniform(0.01, 0.25, n)
    return P[:, 0].copy(), P[:, 1].copy(), P[:, 2].copy(), sig, eps, (L, L, L)


def test_shell_force_is_exact_minus_inner():
    """The shell force with weight S_12 - S_in equals the exact model's
    switched LJ minus the inner model's, pair by pair (golden-pinned pair
    routine, two parameter sets)."""
    rng = np.random.default_rng(11)
    x, y, z, sig, eps, box = shell_system(rng)
    pi, pj = orc.neighbor_build(x, y, z, box, 12.5, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    fsh, _ = orc.sur_shell(x, y, z, sig, eps, pi, pj, box)
    n = len(x)
    fe = [np.zeros(n) for _ in range(3)]
    fi = [np.zeros(n) for _ in range(3)]
    orc.pair_forces(x, y, z, sig, eps, pi, pj, 12.0, 7.5, box, *fe)
    orc.pair_forces(x, y, z, sig, eps, pi, pj, 7.5, 6.5, box, *fi)
    ref = np.stack(fe) - np.stack(fi)
    assert np.abs(ref).max() > 1e-3
    assert np.allclose(fsh, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())


def test_shell_hessian_matches_finite_differences():
    """K_i = -dF_i/dx_i of the shell force (central differences, f64)."""
    rng = np.random.default_rng(12)
    x, y, z, sig, eps, box = shell_system(rng, n=80, L=24.0)
    pi, pj = orc.neighbor_build(x, y, z, box, 12.5, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    _, K = orc.sur_shell(x, y, z, sig, eps, pi, pj, box)
    h = 1e-5
    for i in (0, 7, 33, 61):
        J = np.zeros((3, 3))
        for d, arr in enumerate((x, y, z)):
            arr[i] += h
            fp, _ = orc.sur_shell(x, y, z, sig, eps, pi, pj, box)
            arr[i] -= 2 * h
            fm, _ = orc.sur_shell(x, y, z, sig, eps, pi, pj, box)
            arr[i] += h
            J[:, d] = -(fp[:, i] - fm[:, i]) / (2 * h)
        k = K[i]
        Km = np.array([[k[0], k[3], k[4]], [k[3], k[1], k[5]], [k[4], k[5], k[2]]])
        assert np.abs(Km).max() > 0.0
        assert np.allclose(Km, J, rtol=1e-6, atol=1e-7 * np.abs(Km).max()), (i, Km, J)
