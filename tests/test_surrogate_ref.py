"""Pins the CPU reference of the Born-radius surrogate (oracle/surrogate.py)
by mathematics: B-spline partition of unity and derivatives, descriptor
gradients and MLP gradients against central finite differences, bf16
rounding against the IEEE definition, and descriptors against a brute-force
double loop."""
import numpy as np

from oracle import surrogate as sg


def test_bf16_rounding_matches_ieee_round_to_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-3], np.float32)
    r = sg.bf16(x)
    # bf16 keeps 8 significant bits; the result is the nearest, ties to even
    for v, b in zip(x.astype(np.float64), r):
        m, e = np.frexp(v)
        grid = np.ldexp(np.round(m * 256.0), e - 8)
        lo, hi = np.ldexp(np.floor(m * 256.0), e - 8), np.ldexp(np.ceil(m * 256.0), e - 8)
        assert b in (lo, hi) and abs(b - v) <= abs(grid - v) + 1e-30
    assert sg.bf16(np.float32(1.00390625)) == 1.0          # tie -> even mantissa


def test_bspline_basis_partition_and_derivative():
    d = np.linspace(1.6, 10.0, 500)                          # inside the shells, below the taper
    w, wp = sg.shell_weights(d)
    assert np.allclose(w.sum(1), 1.0, atol=1e-12)
    assert np.allclose(wp.sum(1), 0.0, atol=1e-12)
    h = 1e-6
    wpl, _ = sg.shell_weights(d - h)
    wph, _ = sg.shell_weights(d + h)
    assert np.allclose((wph - wpl) / (2 * h), wp, atol=1e-6)
    # smooth cut-off: everything vanishes at 12 A with zero slope
    w12, wp12 = sg.shell_weights(np.array([12.0, 12.5]))
    assert np.all(w12 == 0.0) and np.all(wp12 == 0.0)


def test_taper_is_c2_smooth_step():
    """T = 1 up to 10.5 A and 0 from 12 A, with value, slope and curvature
    continuous at both ends; the stated derivative is the derivative."""
    h = 1e-6
    assert sg.taper(10.5) == 1.0 and sg.taper(12.0) == 0.0 and sg.taper(13.0) == 0.0
    for e in (10.5, 12.0):
        assert abs(sg.taperp(e)) < 1e-12
        curv = (sg.taperp(e + 1e-4) - sg.taperp(e - 1e-4)) / 2e-4
        assert abs(curv) < 1e-3
    d = np.linspace(10.51, 11.99, 50)
    fd = (sg.taper(d + h) - sg.taper(d - h)) / (2 * h)
    assert np.allclose(fd, sg.taperp(d), atol=1e-7)
    assert np.all(np.diff(sg.taper(d)) < 0)
    assert abs(sg.taper(11.25) - 0.5) < 1e-12                      # symmetric about the midpoint


def small_system(seed, n=40):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-6, 6, (n, 3))
    R = rng.choice([2.2, 2.05, 2.02, 2.3, 1.7], n)
    heavy = R > 1.9
    q = rng.normal(0, 0.3, n)
    gamma = rng.uniform(0.1, 0.9, n)
    return pos, R, q, heavy, gamma


def test_descriptors_match_brute_force():
    pos, R, q, heavy, gamma = small_system(1)
    F = sg.features(pos, R, q, heavy, gamma)
    for i in range(len(R)):
        f = np.zeros(64)
        for j in range(len(R)):
            if j == i or not heavy[j]:
                continue
            d = np.linalg.norm(pos[j] - pos[i])
            if d >= 12.0:
                continue
            c = int(sg.channel(R[j]))
            for k in range(16):
                f[c * 16 + k] += gamma[j] * sg.b3((d - sg.MU[k]) / sg.H) * sg.taper(d)
        assert np.allclose(F[i, :64], f, atol=1e-12)
    assert np.array_equal(F[:, 64], R) and np.array_equal(F[:, 65], q)


def test_unit_weights_count_neighbours():
    """With unit weights the shell sums of one channel add up to the number of
    its heavy neighbours between 1.125 and 10.5 A (B-spline partition of
    unity, taper 1); neighbours beyond 12 A add nothing."""
    rng = np.random.default_rng(8)
    d = np.concatenate([rng.uniform(1.2, 10.4, 9), [12.5, 14.0]])
    u = rng.normal(size=(d.size, 3))
    pos = np.vstack([np.zeros(3), u / np.linalg.norm(u, axis=1)[:, None] * d[:, None]])
    R = np.array([1.7] + [2.2] * d.size)                  # channel 2 neighbours
    heavy = R > 1.9
    F = sg.features(pos, R, np.zeros(len(R)), heavy)
    assert abs(F[0, 32:48].sum() - 9.0) < 1e-12
    assert np.all(F[0, :32] == 0.0) and np.all(F[0, 48:64] == 0.0)


def test_descriptor_gradient_matches_finite_differences():
    pos, R, q, heavy, gamma = small_system(2)
    G = np.random.default_rng(3).normal(size=(len(R), 64))
    g, dgam = sg.feature_gradient(pos, R, heavy, G, gamma)
    h = 1e-5
    for a in (0, 7, 19):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[a, c] += h
            p2[a, c] -= h
            e1 = (G * sg.features(p1, R, q, heavy, gamma)[:, :64]).sum()
            e2 = (G * sg.features(p2, R, q, heavy, gamma)[:, :64]).sum()
            assert abs((e1 - e2) / (2 * h) - g[a, c]) < 1e-6 * max(1.0, abs(g[a, c]))
    # the features are linear in the weights
    for a in np.flatnonzero(heavy)[:4]:
        g1 = gamma.copy()
        g1[a] += 0.25
        de = (G * sg.features(pos, R, q, heavy, g1)[:, :64]).sum() - (G * sg.features(pos, R, q, heavy, gamma)[:, :64]).sum()
        assert abs(de / 0.25 - dgam[a]) < 1e-9 * max(1.0, abs(dgam[a]))


def random_params(seed):
    rng = np.random.default_rng(seed)
    s = lambda *sh: rng.normal(0, 0.2, sh)
    return sg.Params(s(66, 128), s(128), s(128, 128), s(128), s(128, 128), s(128), s(128), 0.1,
                     rng.normal(0, 1, 66), rng.uniform(0.5, 2, 66), 5.0, 1.6)


def test_mlp_input_gradient_matches_finite_differences():
    P = random_params(4)
    F = np.random.default_rng(5).normal(size=(6, 66))
    s, cache = sg.forward(P, F, sg.exact, sg.exact)
    dF = sg.input_gradient(P, cache, sg.exact)
    h = 1e-6
    for k in (0, 13, 40, 64, 65):
        F1, F2 = F.copy(), F.copy()
        F1[:, k] += h
        F2[:, k] -= h
        fd = (sg.forward(P, F1, sg.exact, sg.exact)[0] - sg.forward(P, F2, sg.exact, sg.exact)[0]) / (2 * h)
        assert np.allclose(fd, dF[:, k], rtol=1e-5, atol=1e-8)


def test_mlp_weight_gradients_match_finite_differences():
    P = random_params(6)
    rng = np.random.default_rng(7)
    F = rng.normal(size=(9, 66))
    t = rng.normal(5.0, 1.0, 9)
    for w_abs, w_diff in ((1.0, 0.0), (0.5, 30.0)):         # values only; values and pair differences
        loss, g = sg.loss_and_weight_gradients(P, F, t, sg.exact, sg.exact, w_abs, w_diff)
        h = 1e-6
        for name, idx in (("W1", (3, 5)), ("b1", (7,)), ("W2", (10, 100)), ("W3", (127, 0)), ("b3", (64,)),
                          ("w4", (12,)), ("b4", ())):
            if name == "b4":
                old = P.b4
                P.b4 = old + h
                l1 = sg.loss_and_weight_gradients(P, F, t, sg.exact, sg.exact, w_abs, w_diff)[0]
                P.b4 = old - h
                l2 = sg.loss_and_weight_gradients(P, F, t, sg.exact, sg.exact, w_abs, w_diff)[0]
                P.b4 = old
            else:
                arr = getattr(P, name)
                old = arr[idx]
                arr[idx] = old + h
                l1 = sg.loss_and_weight_gradients(P, F, t, sg.exact, sg.exact, w_abs, w_diff)[0]
                arr[idx] = old - h
                l2 = sg.loss_and_weight_gradients(P, F, t, sg.exact, sg.exact, w_abs, w_diff)[0]
                arr[idx] = old
            assert abs((l1 - l2) / (2 * h) - g[name][idx]) < 1e-6 * max(1.0, abs(g[name][idx])), (name, w_diff)


def test_pair_loss_is_the_squared_error_of_differences():
    """With w_abs = 0 the loss sees only the differences within pairs: adding
    the same constant to both targets of every pair leaves it unchanged, and
    a network that is off by a constant has zero loss."""
    P = random_params(8)
    rng = np.random.default_rng(9)
    F = rng.normal(size=(10, 66))
    s, _ = sg.forward(P, F, sg.exact, sg.exact)
    shift = np.repeat(rng.normal(0, 3, 5), 2)
    l0 = sg.loss_and_weight_gradients(P, F, s + 0.7, sg.exact, sg.exact, 0.0, 1.0)[0]
    assert l0 < 1e-20
    t = rng.normal(5.0, 1.0, 10)
    l1 = sg.loss_and_weight_gradients(P, F, t, sg.exact, sg.exact, 0.0, 1.0)[0]
    l2 = sg.loss_and_weight_gradients(P, F, t + shift, sg.exact, sg.exact, 0.0, 1.0)[0]
    assert abs(l1 - l2) < 1e-12 * l1


def test_bf16_pair_carries_sixteen_bits():
    rng = np.random.default_rng(12)
    v = rng.normal(0, 3, 10000).astype(np.float32).astype(np.float64)
    hi = sg.bf16(v)
    assert np.all(np.abs(hi - v) <= 2.0 ** -8 * np.abs(v))          # one bf16: 8 significant bits
    assert np.all(np.abs(sg.split(v) - v) <= 2.0 ** -16 * np.abs(v))   # the pair: 16
    assert np.array_equal(sg.split(hi), hi)                           # bf16 values are carried exactly
