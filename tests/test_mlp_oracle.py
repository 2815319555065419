"""Pins of the CPU restatement of the learned initial guess (oracle/mlp_oracle.py):
BF16 rounding on values with known results and against torch's conversion,
GELU against torch's exact-erf GELU, the analytic gradient against torch
autograd and central finite differences, the feature layout on hand-built
fibers, and the guess/target round trip."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as mo


def test_weight_count_matches_header(pkg):
    assert mo.WEIGHTS == pkg.GUESS_WEIGHTS == 254816
    assert (mo.INPUTS, mo.OUTPUTS, mo.POINTS) == (pkg.GUESS_INPUTS, pkg.GUESS_OUTPUTS, pkg.GUESS_POINTS)


def test_bf16_known_values():
    e = 2.0 ** -8
    x = np.array([1.0, 1 + e, 1 + 3 * e, 1 + e + 2 ** -20, -(1 + 3 * e), 0.0, -0.0, np.inf, -np.inf,
                  3.0e38], np.float32)
    want = np.array([1.0, 1.0, 1 + 4 * e, 1 + 2 * e, -(1 + 4 * e), 0.0, -0.0, np.inf, -np.inf,
                     np.float32(3.0e38).item()], np.float64)
    got = mo.bf16(x).astype(np.float64)
    # 3e38 rounds to a BF16 neighbour within one BF16 ulp
    assert np.array_equal(got[:9], want[:9])
    assert abs(got[9] - 3.0e38) <= 3.0e38 * 2 ** -8
    assert np.signbit(got[6])
    assert np.isnan(mo.bf16(np.array([np.nan], np.float32))[0])


def test_bf16_matches_torch_conversion():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(20000), rng.standard_normal(20000) * 1e-30,
                        rng.standard_normal(20000) * 1e30]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mo.bf16(x), ref)


def test_gelu_closed_form():
    z = torch.linspace(-8, 8, 4001, dtype=torch.float64, requires_grad=True)
    g = torch.nn.functional.gelu(z, approximate="none")
    g.sum().backward()
    zn = z.detach().numpy()
    assert np.allclose(mo.gelu(zn), g.detach().numpy(), rtol=0, atol=1e-15)
    assert np.allclose(mo.gelu_grad(zn), z.grad.numpy(), rtol=0, atol=1e-14)


def _problem(B=32, seed=3):
    rng = np.random.default_rng(seed)
    w = mo.init(seed).astype(np.float64)
    # non-zero biases so the bias paths are exercised
    for W, b in mo.unpack(w):
        b[:] = rng.uniform(-0.1, 0.1, b.size)
    X = rng.standard_normal((B, mo.INPUTS))
    T = rng.standard_normal((B, mo.OUTPUTS)) * 0.3
    T[:, -1] = rng.integers(0, 94, B)    # valid outputs per row (0 masks a row out)
    return w, X, T


def _torch_loss(w, X, T):
    layers = []
    off = 0
    for i, o in zip(mo.LAYER_IN, mo.LAYER_OUT):
        layers.append((w[off:off + i * o].view(i, o), w[off + i * o:off + i * o + o]))
        off += i * o + o
    h = X
    for q, (W, b) in enumerate(layers):
        h = h @ W + b
        if q + 1 < len(layers):
            h = torch.nn.functional.gelu(h, approximate="none")
    valid = torch.arange(mo.OUTPUTS)[None, :] < T[:, -1:]
    e = torch.where(valid, h - T, torch.zeros_like(h))
    return (e * e).sum() / X.shape[0]


def test_gradient_matches_torch_autograd():
    w, X, T = _problem()
    loss, g = mo.loss_and_grad(w, X, T, rounding=False)
    wt = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    lt = _torch_loss(wt, torch.tensor(X), torch.tensor(T))
    lt.backward()
    assert abs(loss - lt.item()) <= 1e-12 * abs(lt.item())
    ref = wt.grad.numpy()
    assert np.abs(g - ref).max() <= 1e-12 * np.abs(ref).max()


def test_gradient_matches_finite_differences():
    w, X, T = _problem(B=16, seed=7)
    _, g = mo.loss_and_grad(w, X, T, rounding=False)
    rng = np.random.default_rng(11)
    # coordinates in every layer's weights and biases
    offs = np.cumsum([0] + [i * o + o for i, o in zip(mo.LAYER_IN, mo.LAYER_OUT)])
    picks = []
    for q, (i, o) in enumerate(zip(mo.LAYER_IN, mo.LAYER_OUT)):
        picks += list(offs[q] + rng.integers(0, i * o, 4))
        picks += list(offs[q] + i * o + rng.integers(0, o, 2))
    h = 1e-6
    for k in picks:
        wp = w.copy(); wp[k] += h
        wm = w.copy(); wm[k] -= h
        lp, _ = mo.loss_and_grad(wp, X, T, rounding=False)
        lm, _ = mo.loss_and_grad(wm, X, T, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_rounded_gradient_near_exact_gradient():
    # BF16 operands change the gradient by about their relative precision
    w, X, T = _problem()
    l0, g0 = mo.loss_and_grad(w, X, T, rounding=False)
    l1, g1 = mo.loss_and_grad(w, X, T, rounding=True)
    assert abs(l1 - l0) <= 2e-2 * l0
    assert np.linalg.norm(g1 - g0) <= 3e-2 * np.linalg.norm(g0)


def test_masked_rows_have_no_gradient():
    w, X, T = _problem(B=16)
    T[:, -1] = 0
    loss, g = mo.loss_and_grad(w, X, T)
    assert loss == 0.0 and not np.any(g)


@pytest.mark.parametrize("dim", [2, 3])
def test_feature_layout(dim):
    rng = np.random.default_rng(dim)
    nb = np.array([5, 31, 32, 2])
    n = int(nb.sum())
    pts = rng.standard_normal(dim * n) * 4
    rhs = rng.standard_normal(dim * n)
    prev = rng.standard_normal(dim * n)
    corr = rng.standard_normal(dim * n)
    rhs[dim * (5 + 31 + 32):] = 0.0              # a fiber without load
    X, S = mo.features(dim, nb, pts, rhs, prev, corr)
    z0 = corr[:dim * 5].reshape(5, dim)
    p0 = pts[:dim * 5].reshape(5, dim)
    r0 = rhs[:dim * 5].reshape(5, dim)
    q0 = prev[:dim * 5].reshape(5, dim)
    s0 = np.abs(r0).max()
    assert S[0] == np.float32(s0)
    c = p0.mean(axis=0)
    for p in range(5):
        for d in range(dim):
            assert X[0, 3 * p + d] == mo.bf16(np.float32((p0[p, d] - c[d]) / 5))
            assert X[0, 93 + 3 * p + d] == mo.bf16(np.float32(r0[p, d] / s0))
            assert X[0, 186 + 3 * p + d] == mo.bf16(np.float32(q0[p, d] / s0))
            assert X[0, 279 + 3 * p + d] == mo.bf16(np.float32(z0[p, d] / s0))
        if dim == 2:
            assert X[0, 3 * p + 2] == 0 and X[0, 93 + 3 * p + 2] == 0 and X[0, 186 + 3 * p + 2] == 0
            assert X[0, 279 + 3 * p + 2] == 0
    assert not np.any(X[0, 15:93]) and not np.any(X[0, 294:372]) and not np.any(X[0, 372:])
    # the largest rhs component of a fiber is +-1 after scaling
    assert np.abs(X[1, 93:186]).max() == 1.0
    assert not np.any(X[2]) and S[2] == 0           # 32 points: no learned guess
    assert S[3] == 0 and not np.any(X[3, 93:])      # unloaded fiber


@pytest.mark.parametrize("dim", [2, 3])
def test_guess_inverts_targets(dim):
    rng = np.random.default_rng(5)
    nb = np.array([7, 31, 40])
    n = int(nb.sum())
    x = rng.standard_normal(dim * n)
    y = rng.standard_normal(dim * n)
    S = np.array([0.5, 2.0, 1.0], np.float32)
    T = mo.targets(dim, nb, x, y, S)
    assert T[0, -1] == 21 and T[1, -1] == 93 and T[2, -1] == 0
    x0 = mo.guess(dim, nb, y, T, S)
    m = dim * (7 + 31)
    assert np.allclose(x0[:m], x[:m], rtol=0, atol=1e-6)
    assert np.array_equal(x0[m:], y[m:])             # fibers beyond 31 points keep y


def test_targets_beyond_the_clip_are_not_samples():
    """A fiber whose correction is more than 32 times its load scale gives no
    training output (valid count 0), so it cannot dominate the squared loss."""
    nb = np.array([4, 4])
    x = np.zeros(3 * 8)
    y = np.zeros(3 * 8)
    x[2] = 10.0          # fiber 0: 10 / 0.5 = 20 <= 32
    x[12 + 5] = 40.0     # fiber 1: 40 / 1 = 40 > 32
    T = mo.targets(3, nb, x, y, np.array([0.5, 1.0], np.float32))
    assert T[0, -1] == 12 and T[0, 2] == 20.0
    assert T[1, -1] == 0
