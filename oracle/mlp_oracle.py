"""CPU restatement of the learned initial guess (numpy, float64 arithmetic on
the BF16-rounded operands the device uses).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
network (features, forward, loss, gradients).  The product path never imports
this module.

The learned guess has no counterpart in the reference: Meca::solve seeds the
solver with vSOL = 0 (meca.cc:704), and this network only changes that seed.
What it computes is fixed by the library's C header (cymg_guess_* in
include/cymeca.h):
  features of a fiber of np <= 31 points, s = |rhs_f|_inf:
    [0, 93)    (x_p - centroid) / 5, three slots per point (z = 0 in 2D)
    [93, 186)  rhs_p / s
    [186, 279) previous solution_p / s
    [279, 372) z1_p / s, z1 = M^-1 (b - A M^-1 b) (the block-Jacobi correction:
               what the linked fibers add to the fiber's own solution M^-1 b)
    [372, 384) zero
  every feature rounded to BF16 (round to nearest even);
  layers q = 0..3: Z_q = H_q W_q + b_q with BF16 H_q and W_q and exact (here
  float64; device FP32) accumulation, H_{q+1} = bf16(gelu(Z_q)), output Z_3;
  loss = sum over rows and valid outputs (o < T[row][95]) of (Z_3 - T)^2 / B;
  backward: dZ_3 = 2 (Z_3 - T) / B on the valid outputs, gW_q = H_q^T bf16(dZ_q),
  gb_q = column sums of dZ_q, dZ_{q-1} = bf16(dZ_q) bf16(W_q)^T * gelu'(Z_{q-1}).
Pinned by: finite differences of the loss (rounding off), the closed form of
GELU and its derivative, BF16 rounding on values with known results.
"""
import numpy as np
from scipy.special import erf

POINTS, INPUTS, HIDDEN, OUTPUTS = 31, 384, 256, 96
LAYER_IN = (INPUTS, HIDDEN, HIDDEN, HIDDEN)
LAYER_OUT = (HIDDEN, HIDDEN, HIDDEN, OUTPUTS)
WEIGHTS = sum(i * o + o for i, o in zip(LAYER_IN, LAYER_OUT))


def bf16(x):
    """Round float32 values to BF16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, u & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(w):
    """Flat weights -> [(W [in][out], b [out])] layer by layer."""
    w = np.asarray(w)
    assert w.size == WEIGHTS
    out, off = [], 0
    for i, o in zip(LAYER_IN, LAYER_OUT):
        W = w[off:off + i * o].reshape(i, o)
        off += i * o
        b = w[off:off + o]
        off += o
        out.append((W, b))
    return out


def init(seed):
    """Weights with the device initialiser's distribution (not its draws)."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, o in zip(LAYER_IN, LAYER_OUT):
        a = 1.0 / np.sqrt(i)
        parts.append(rng.uniform(-a, a, i * o))
        parts.append(np.zeros(o))
    return np.concatenate(parts).astype(np.float32)


def _r(x, rounding):
    return bf16(x).astype(np.float64) if rounding else np.asarray(x, dtype=np.float64)


def forward(w, X, rounding=True):
    """Pre-activations Z_q and layer inputs H_q of rows X [B][384]."""
    layers = unpack(w)
    H = [_r(X, rounding)]
    Z = []
    for q, (W, b) in enumerate(layers):
        z = H[q] @ _r(W, rounding) + np.asarray(b, dtype=np.float64)
        Z.append(z)
        if q + 1 < len(layers):
            H.append(_r(gelu(z).astype(np.float32) if rounding else gelu(z), rounding))
    return Z, H


def loss_and_grad(w, X, T, rounding=True):
    """Masked MSE loss of rows X against targets T [B][96] and its gradient."""
    layers = unpack(w)
    B = X.shape[0]
    Z, H = forward(w, X, rounding)
    T = np.asarray(T, dtype=np.float64)
    valid = np.arange(OUTPUTS)[None, :] < T[:, OUTPUTS - 1:OUTPUTS]
    e = np.where(valid, Z[-1] - T, 0.0)
    loss = float((e * e).sum() / B)
    dz = 2.0 * e / B
    grads = [None] * len(layers)
    for q in range(len(layers) - 1, -1, -1):
        W, _ = layers[q]
        dzr = _r(dz.astype(np.float32), rounding) if rounding else dz
        gW = H[q].T @ dzr
        gb = dz.sum(axis=0)
        grads[q] = (gW, gb)
        if q == 0:
            break
        dh = dzr @ _r(W, rounding).T
        dz = dh * gelu_grad(Z[q - 1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def features(dim, nb_points, pts, rhs, prev, corr=None):
    """Features [nfib][384] (BF16 values as float32) and scales [nfib] (float32);
    corr: the block-Jacobi correction z1 (zero if None)."""
    if corr is None:
        corr = np.zeros_like(np.asarray(rhs, dtype=np.float64))
    nb_points = np.asarray(nb_points)
    F = nb_points.size
    X = np.zeros((F, INPUTS), np.float32)
    S = np.zeros(F, np.float32)
    off = 0
    for f, n in enumerate(nb_points):
        n = int(n)
        sl = slice(dim * off, dim * (off + n))
        off += n
        if n > POINTS:
            continue
        p = np.asarray(pts[sl], dtype=np.float64).reshape(n, dim)
        r = np.asarray(rhs[sl], dtype=np.float64).reshape(n, dim)
        q = np.asarray(prev[sl], dtype=np.float64).reshape(n, dim)
        z = np.asarray(corr[sl], dtype=np.float64).reshape(n, dim)
        s = np.abs(r).max()
        inv = 1.0 / s if s > 0 else 0.0
        c = p.mean(axis=0)
        for k, block in enumerate(((p - c) * 0.2, r * inv, q * inv, z * inv)):
            g = np.zeros((n, 3))
            g[:, :dim] = block
            X[f, 3 * POINTS * k:3 * POINTS * k + 3 * n] = g.ravel().astype(np.float32)
        S[f] = np.float32(s) if s > 0 else 0.0
    return bf16(X), S


TARGET_CLIP = 32.0


def targets(dim, nb_points, x, y, scale):
    """Training targets (x - y)/s [nfib][96]; T[.][95] = valid outputs (3 np),
    0 for a fiber with a target beyond TARGET_CLIP in magnitude (not trained on)."""
    F = len(nb_points)
    T = np.zeros((F, OUTPUTS), np.float32)
    off = 0
    for f, n in enumerate(nb_points):
        n = int(n)
        sl = slice(dim * off, dim * (off + n))
        off += n
        s = float(scale[f])
        if n > POINTS or s <= 0:
            continue
        g = np.zeros((n, 3))
        g[:, :dim] = ((np.asarray(x[sl]) - np.asarray(y[sl])) / s).reshape(n, dim)
        T[f, :3 * n] = g.ravel().astype(np.float32)
        T[f, OUTPUTS - 1] = 3 * n if np.abs(T[f, :3 * n]).max() <= TARGET_CLIP else 0
    return T


def guess(dim, nb_points, y, out, scale):
    """x0 = y + s * out on the points of fibers with a learned guess."""
    x0 = np.array(y, dtype=np.float64)
    off = 0
    for f, n in enumerate(nb_points):
        n = int(n)
        sl = slice(dim * off, dim * (off + n))
        off += n
        s = float(scale[f])
        if n > POINTS or s <= 0:
            continue
        o = np.asarray(out[f], dtype=np.float64)[:3 * n].reshape(n, 3)[:, :dim]
        x0[sl] += s * o.ravel()
    return x0
