"""Oracle (test infrastructure only) of the learned SGS closure.

The closure replaces the dynamic procedure of the compressible DSM
(src/DSMSGS.hpp:124-444) by a small network evaluated per point:

    features f(grad u, Delta)  ->  MLP 7-64-64-64-2 (tanh-form GELU)
                               ->  (C, C / Pr_t)  ->  mu_sgs, k_sgs

with the DSM's own definitions of the outputs:
    mu_sgs = rho C Delta^2 |S|                      (src/DSMSGS.hpp:319, with
                                                     C = LM / (Delta^2 MM), :334)
    k_sgs  = cp (C / Pr_t) rho Delta^2 |S|          (:406-408 with
                                                     Pr_t = -C Delta^2 NN / NK, :440)
    Delta  = (dxi deta dzeta / J)^(1/3)             (:332-333)
    |S|    = sqrt(2 S_ij S_ij)                      (:127-168)
The second output is C / Pr_t = -NK / (Delta^2 NN) rather than Pr_t itself:
the DSM's Pr_t is singular where NK -> 0, its ratio is not.

Numerics the GPU kernels follow and this file restates:
- features in FP64, then rounded to FP32 and standardised in FP32;
- every matrix product takes BF16 operands (round to nearest even) and
  accumulates in FP32 (here: exactly, in FP64, then rounded once);
- biases, GELU, the loss and Adam in FP32.
This file is NumPy only and is the checker; nothing in the product imports it.
"""
import numpy as np

N_IN, N_HID, N_OUT = 7, 64, 2
LAYERS = [(N_HID, N_IN), (N_HID, N_HID), (N_HID, N_HID), (N_OUT, N_HID)]
N_PARAMS = sum(o * i + o for o, i in LAYERS)  # 8962
GELU_K0 = np.float32(0.7978845608028654)      # sqrt(2 / pi)
GELU_K1 = np.float32(0.044715)
TINY = 1e-300


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).reshape(x.shape)


def unpack(params):
    """Flat float32 parameters -> [(W, b)] per layer; W is (out, in)."""
    out, o = [], 0
    for n_out, n_in in LAYERS:
        W = params[o:o + n_out * n_in].reshape(n_out, n_in)
        o += n_out * n_in
        b = params[o:o + n_out]
        o += n_out
        out.append((W, b))
    return out


def init_params(seed):
    """He-style uniform init (weights U(-s, s), s = sqrt(3 / fan_in) * gain), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for n_out, n_in in LAYERS:
        s = np.sqrt(3.0 / n_in)
        parts.append(rng.uniform(-s, s, size=n_out * n_in))
        parts.append(np.zeros(n_out))
    return np.concatenate(parts).astype(np.float32)


def gelu(z):
    """tanh-form GELU in FP32: 0.5 z (1 + tanh(k0 (z + k1 z^3)))."""
    z = z.astype(np.float32)
    u = GELU_K0 * (z + GELU_K1 * z * z * z)
    return (np.float32(0.5) * z * (np.float32(1) + np.tanh(u))).astype(np.float32)


def gelu_grad(z):
    z = z.astype(np.float64)
    u = 0.7978845608028654 * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    du = 0.7978845608028654 * (1 + 3 * 0.044715 * z * z)
    return (0.5 * (1 + t) + 0.5 * z * (1 - t * t) * du).astype(np.float32)


def _mm(a, W):
    """a (n, k) @ W.T with BF16 operands, exact products, one FP32 rounding."""
    return (bf16(a).astype(np.float64) @ bf16(W).astype(np.float64).T).astype(np.float32)


def forward(params, x, keep=False):
    """x: (n, 7) standardised FP32 inputs -> (n, 2) FP32 outputs."""
    h = x.astype(np.float32)
    cache = []
    layers = unpack(params)
    for li, (W, b) in enumerate(layers):
        z = (_mm(h, W) + b).astype(np.float32)
        cache.append((h, z))
        h = gelu(z) if li < len(layers) - 1 else z
    return (h, cache) if keep else h


def loss_and_grad(params, x, t):
    """L = sum((y - t)^2) / (n * N_OUT) over a batch of n samples, and its
    gradient (flat, parameter order) with BF16 operands in every product;
    dL/dy = 2 (y - t) / (n * N_OUT) in FP32."""
    n = x.shape[0]
    y, cache = forward(params, x, keep=True)
    r = (y.astype(np.float64) - t)
    L = float(np.sum(r * r) / (n * N_OUT))
    dz = (2.0 * r / (n * N_OUT)).astype(np.float32)  # dL/dy
    layers = unpack(params)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        h, z = cache[li]
        if li < len(layers) - 1:
            dz = (dz * gelu_grad(z)).astype(np.float32)
        dW = (bf16(dz).astype(np.float64).T @ bf16(h).astype(np.float64)).astype(np.float32)
        db = np.sum(dz.astype(np.float64), axis=0).astype(np.float32)
        grads[li] = (dW, db)
        if li > 0:
            dz = (bf16(dz).astype(np.float64) @ bf16(W).astype(np.float64)).astype(np.float32)
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads]).astype(np.float32)
    return L, flat


class Adam:
    """Adam (Kingma & Ba 2015) in FP32 with bias correction."""

    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n, np.float32)
        self.v = np.zeros(n, np.float32)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = np.float32(lr), np.float32(b1), np.float32(b2), np.float32(eps)

    def step(self, p, g):
        """One update; the bias corrections 1 - beta^t are formed in FP64 and
        rounded to FP32, everything else is FP32 with one rounding per operation."""
        self.t += 1
        b1, b2 = self.b1, self.b2
        one = np.float32(1)
        self.m = (b1 * self.m + (one - b1) * g).astype(np.float32)
        self.v = (b2 * self.v + (one - b2) * (g * g)).astype(np.float32)
        c1 = np.float32(1.0 - float(b1) ** self.t)
        c2 = np.float32(1.0 - float(b2) ** self.t)
        mh = self.m / c1
        vh = self.v / c2
        return (p - self.lr * mh / (np.sqrt(vh) + self.eps)).astype(np.float32)


def invariants(A):
    """A (..., 3, 3) FP64 velocity gradient d u_i / d x_j -> (theta, s2, w2, s3, sw2, s2w2)."""
    S = 0.5 * (A + np.swapaxes(A, -1, -2))
    W = 0.5 * (A - np.swapaxes(A, -1, -2))
    theta = np.trace(A, axis1=-2, axis2=-1)
    s2 = np.sum(S * S, axis=(-2, -1))
    w2 = np.sum(W * W, axis=(-2, -1))
    SS = S @ S
    WW = W @ W
    s3 = np.trace(SS @ S, axis1=-2, axis2=-1)
    sw2 = np.trace(S @ WW, axis1=-2, axis2=-1)
    s2w2 = np.trace(SS @ WW, axis1=-2, axis2=-1)
    return theta, s2, w2, s3, sw2, s2w2


def features(A, vol_J):
    """7 closure inputs (FP64): theta/a, (s2 - w2)/a^2, s3/a^3, sw2/a^3,
    s2w2/a^4, log a, log Delta, with a^2 = s2 + w2 = |grad u|^2 and
    log Delta = log(vol / J) / 3.  A vanishing gradient gives zeros and
    log a = log(1e-300) (outside any training range)."""
    theta, s2, w2, s3, sw2, s2w2 = invariants(A)
    a2 = s2 + w2
    a = np.sqrt(a2)
    ok = a2 > 0
    inv = np.where(ok, 1.0 / np.where(ok, a, 1.0), 0.0)
    f = np.stack([theta * inv, (s2 - w2) * inv * inv, s3 * inv ** 3, sw2 * inv ** 3, s2w2 * inv ** 4,
                  np.log(np.where(ok, a, TINY)), np.log(vol_J) / 3.0 * np.ones_like(a)], axis=-1)
    return f


def line_average(f, axis):
    """Mean of f along one grid direction, summed as the solver sums it: f / N
    added in line order (AbstractSGS::doAveraging, src/AbstractSGS.cpp:13-164)."""
    N = f.shape[axis]
    acc = np.zeros_like(np.take(f, 0, axis=axis))
    for i in range(N):
        acc = acc + np.take(f, i, axis=axis) / N
    return np.broadcast_to(np.expand_dims(acc, axis), f.shape)


def averaged_features(A, vol_J, avg):
    """Closure inputs under LESAVERAGING avg: the features rounded to FP32 (as
    the solver stores them), then line means along x, y or z (XI1_AVG, XI2_AVG,
    XI3_AVG) or the global mean (GLOBAL) of (Nz, Ny, Nx) grids."""
    f = features(A, vol_J).astype(np.float32).astype(np.float64)
    if avg == "GLOBAL":
        return np.broadcast_to(f.reshape(-1, f.shape[-1]).mean(axis=0), f.shape)
    axis = {"XI1_AVG": 2, "XI2_AVG": 1, "XI3_AVG": 0}[avg]
    return np.stack([line_average(f[..., k], axis) for k in range(f.shape[-1])], axis=-1)


def smag(A):
    """|S| = sqrt(2 S_ij S_ij)."""
    S = 0.5 * (A + np.swapaxes(A, -1, -2))
    return np.sqrt(2.0 * np.sum(S * S, axis=(-2, -1)))


def standardise(f, mean, std):
    """FP64 features -> FP32 network inputs (f32(f) - mean) / std in FP32."""
    return ((f.astype(np.float32) - mean.astype(np.float32)) / std.astype(np.float32)).astype(np.float32)


def closure_outputs(y, tmean, tstd, rho, Smag, vol_J, cp):
    """Network outputs -> (mu_sgs, k_sgs) in FP64; the de-standardised targets
    C = y0 std0 + mean0 and D = C / Pr_t are FP32 (one rounding per operation)."""
    y = y.astype(np.float32)
    C = ((y[..., 0] * np.float32(tstd[0])).astype(np.float32) + np.float32(tmean[0])).astype(np.float32)
    D = ((y[..., 1] * np.float32(tstd[1])).astype(np.float32) + np.float32(tmean[1])).astype(np.float32)
    d = np.cbrt(vol_J)
    base = rho * (d * d) * Smag
    return base * C.astype(np.float64), cp * base * D.astype(np.float64)
