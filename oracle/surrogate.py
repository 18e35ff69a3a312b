"""CPU reference of the learned Born-radius surrogate (test infrastructure:
only tests/ may import it).

The surrogate has no counterpart in the reference; this module is the
checker for the GPU kernels of agbnp3_amd/csrc/k_surrogate.hip:

- local-environment descriptors: for atom i, four heavy-atom channels (by
  radius, R = r + 0.5) x 16 radial shells up to 12 A,
      F[c, k] = sum_{j heavy, j != i, channel(j) = c} g_j B3((d_ij - mu_k) / h) T(d_ij)
  with h = 0.75 A, mu_k = h (k + 1/2), B3 the cubic B-spline (support
  |t| < 2), T a quintic smoothstep taper from 1 at 10.5 A to 0 at 12 A and g_j the
  neighbour's self-volume fraction V'_j / V_j (the descreening's factor);
  then R_i and q_i (66 features);
- the MLP 66 -> 128 -> 128 -> 128 -> 1 with SiLU, evaluated as the GPU does:
  bf16 weights; normalised inputs and hidden activations carried as a pair
  of bf16 values (hi = bf16(v), lo = bf16(v - hi)); products accumulated
  exactly (here in FP64; FP32 on the GPU), FP32 biases;
- its input gradient (bf16 deltas) and, for training, the weight gradients
  of a mean squared error on the values and on the differences between two
  conformations of the same atom (bf16 activations and deltas in the
  products).

Everything is plain numpy in FP64 on bf16-rounded values.
"""
import numpy as np

H = 0.75
NSHELL = 16
NCHAN = 4
RCUT = 12.0
TAPER0 = 10.5
MU = H * (np.arange(NSHELL) + 0.5)
CH_EDGES = (2.035, 2.125, 2.25)     # R thresholds between the four heavy channels
NFEAT = NCHAN * NSHELL + 2           # 66
KPAD = 96                            # first-layer input padded to a multiple of 32
HID = 128


def bf16(a):
    """Round to bfloat16 (nearest even) and back to float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def b3(t):
    a = np.abs(t)
    out = np.where(a < 1.0, 2.0 / 3.0 - a * a + 0.5 * a ** 3, np.where(a < 2.0, (2.0 - a) ** 3 / 6.0, 0.0))
    return out


def b3p(t):
    a = np.abs(t)
    s = np.sign(t)
    out = np.where(a < 1.0, -2.0 * a + 1.5 * a * a, np.where(a < 2.0, -0.5 * (2.0 - a) ** 2, 0.0))
    return s * out


def taper(d):
    """1 up to 10.5 A, 0 from 12 A, the quintic smoothstep between (C2)."""
    x = np.clip((np.asarray(d, np.float64) - TAPER0) / (RCUT - TAPER0), 0.0, 1.0)
    return 1.0 - x ** 3 * (10.0 - 15.0 * x + 6.0 * x * x)


def taperp(d):
    x = np.clip((np.asarray(d, np.float64) - TAPER0) / (RCUT - TAPER0), 0.0, 1.0)
    return -30.0 * x * x * (1.0 - x) ** 2 / (RCUT - TAPER0)


def channel(R):
    R = np.asarray(R)
    return np.where(R < CH_EDGES[0], 0, np.where(R < CH_EDGES[1], 1, np.where(R < CH_EDGES[2], 2, 3)))


def shell_weights(d):
    """phi_k(d) = B3((d - mu_k)/h) T(d) and its derivative, shape (..., 16)."""
    d = np.asarray(d, np.float64)[..., None]
    t = (d - MU) / H
    T, Tp = taper(d), taperp(d)
    w = b3(t) * T
    wp = b3p(t) / H * T + b3(t) * Tp
    return w, wp


def features(pos, R, q, heavy, gamma=None):
    """pos (n, 3), R (n,), q (n,), heavy (n,) bool, gamma (n,) neighbour
    weights (None: 1) -> F (n, 66), in FP64."""
    pos = np.asarray(pos, np.float64)
    n = len(R)
    gamma = np.ones(n) if gamma is None else np.asarray(gamma, np.float64)
    F = np.zeros((n, NFEAT))
    hv = np.flatnonzero(heavy)
    ch = channel(R[hv])
    for i in range(n):
        dv = pos[hv] - pos[i]
        d = np.sqrt((dv * dv).sum(1))
        m = (hv != i) & (d < RCUT)
        w, _ = shell_weights(d[m])
        w = w * gamma[hv[m]][:, None]
        for c in range(NCHAN):
            F[i, c * NSHELL:(c + 1) * NSHELL] = w[ch[m] == c].sum(0)
    F[:, 64] = R
    F[:, 65] = q
    return F


def feature_gradient(pos, R, heavy, G, gamma=None):
    """Given G (n, 64) = dE/dF for the shell features, return dE/dpos (n, 3)
    at fixed weights and dE/dgamma (n,)."""
    pos = np.asarray(pos, np.float64)
    n = len(R)
    gamma = np.ones(n) if gamma is None else np.asarray(gamma, np.float64)
    hv = np.flatnonzero(heavy)
    ch = channel(R[hv])
    g = np.zeros((n, 3))
    dgam = np.zeros(n)
    for i in range(n):
        dv = pos[hv] - pos[i]
        d = np.sqrt((dv * dv).sum(1))
        m = (hv != i) & (d < RCUT)
        w, wp = shell_weights(d[m])
        gi = G[i].reshape(NCHAN, NSHELL)[ch[m]]     # dE/dF of i in each neighbour's channel
        coef = (gi * wp).sum(1) * gamma[hv[m]] / d[m]   # dE/dd / d
        contrib = coef[:, None] * dv[m]             # along x_j - x_i
        g[i] -= contrib.sum(0)
        np.add.at(g, hv[m], contrib)
        np.add.at(dgam, hv[m], (gi * w).sum(1))
    return g, dgam


class Params:
    """MLP parameters: FP32 master values and the normalisation."""

    def __init__(self, W1, b1, W2, b2, W3, b3_, w4, b4, xmu, xsd, ymu, ysd):
        self.W1, self.b1, self.W2, self.b2, self.W3, self.b3 = W1, b1, W2, b2, W3, b3_
        self.w4, self.b4 = w4, b4
        self.xmu, self.xsd, self.ymu, self.ysd = xmu, xsd, ymu, ysd

    @classmethod
    def from_flat(cls, flat, layout):
        parts = {}
        o = 0
        for name, shape in layout:
            size = int(np.prod(shape))
            parts[name] = np.asarray(flat[o:o + size], np.float64).reshape(shape)
            o += size
        return cls(parts["W1"], parts["b1"], parts["W2"], parts["b2"], parts["W3"], parts["b3"], parts["w4"],
                   parts["b4"][0], parts["xmu"], parts["xsd"], parts["ymu"][0], parts["ysd"][0])


def silu(z):
    return z / (1.0 + np.exp(-z))


def silup(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def exact(a):
    """No rounding: for finite-difference checks of the formulas."""
    return np.asarray(a, np.float64)


def split(a):
    """The value the bf16 pair (hi, lo) carries: hi = bf16(v), lo = bf16(v - hi)."""
    a = np.asarray(a, np.float32).astype(np.float64)
    hi = bf16(a)
    return hi + bf16(a - hi)


def normalise(P, F, act=split):
    X = np.zeros((len(F), KPAD))
    X[:, :NFEAT] = (F - P.xmu) / P.xsd
    return act(X)


def forward(P, F, rnd=bf16, act=split):
    """-> (s, cache) with s the descreening sum predicted for each atom;
    rnd rounds the weights, act the inputs and hidden activations."""
    X = normalise(P, F, act)
    W1 = rnd(np.vstack([P.W1, np.zeros((KPAD - NFEAT, HID))]))
    z1 = X @ W1 + P.b1
    a1 = act(silu(z1))
    z2 = a1 @ rnd(P.W2) + P.b2
    a2 = act(silu(z2))
    z3 = a2 @ rnd(P.W3) + P.b3
    a3 = silu(z3)
    y = a3 @ P.w4 + P.b4
    return y * P.ysd + P.ymu, (X, z1, a1, z2, a2, z3, a3, y)


def input_gradient(P, cache, rnd=bf16):
    """ds/dF (n, 66) for every atom (shell features first)."""
    X, z1, a1, z2, a2, z3, a3, y = cache
    d3 = rnd(P.w4[None, :] * silup(z3))
    d2 = rnd((d3 @ rnd(P.W3).T) * silup(z2))
    d1 = rnd((d2 @ rnd(P.W2).T) * silup(z1))
    dX = d1 @ rnd(P.W1).T
    return dX * P.ysd / P.xsd


def loss_and_weight_gradients(P, F, target, rnd=bf16, act=split, w_abs=1.0, w_diff=0.0):
    """Loss of the normalised output and its parameter gradients; the
    weight-gradient products take the activations rounded by rnd.

    Rows 2k and 2k+1 form a pair (the same atom in two conformations); with
    e = y - t the loss is mean over rows of w_abs e_r^2 + w_diff (e_r - e_p)^2,
    p the row's partner (none for a last, odd row)."""
    s, (X, z1, a1, z2, a2, z3, a3, y) = forward(P, F, rnd, act)
    X, a1, a2 = rnd(X), rnd(a1), rnd(a2)
    t = (target - P.ymu) / P.ysd
    n = len(t)
    e = y - t
    partner = np.arange(n) ^ 1
    partner[partner >= n] = np.arange(n)[partner >= n]
    de = e - e[partner]
    loss = float(np.mean(w_abs * e * e + w_diff * de * de))
    dy = (2.0 * w_abs * e + 4.0 * w_diff * de) / n
    g = {"w4": a3.T @ dy, "b4": dy.sum()}
    d3 = rnd(dy[:, None] * P.w4[None, :] * silup(z3))
    g["W3"], g["b3"] = a2.T @ d3, d3.sum(0)
    d2 = rnd((d3 @ rnd(P.W3).T) * silup(z2))
    g["W2"], g["b2"] = a1.T @ d2, d2.sum(0)
    d1 = rnd((d2 @ rnd(P.W2).T) * silup(z1))
    g["W1"], g["b1"] = (X.T @ d1)[:NFEAT], d1.sum(0)
    return loss, g
