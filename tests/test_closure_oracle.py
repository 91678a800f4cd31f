"""Pins of the learned-closure oracle (oracle/orc_closure.py) against
independent references: torch's bfloat16 conversion, torch's tanh GELU,
torch autograd, torch.optim.Adam, and identities of the velocity-gradient
invariants (Cayley-Hamilton, the vorticity-vector forms, rotation invariance).
"""
import numpy as np
import pytest
import torch

from oracle import orc_closure as C


def rng(seed=0):
    return np.random.default_rng(seed)


def test_bf16_rounding_matches_torch():
    r = rng(1)
    x = np.concatenate([r.standard_normal(20000) * 10.0 ** r.integers(-20, 20, 20000),
                        np.array([0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 65504.0, 1e-38])]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(C.bf16(x), ref)


def test_gelu_is_torch_tanh_gelu():
    z = np.linspace(-12, 12, 20001).astype(np.float32)
    ref = torch.nn.functional.gelu(torch.from_numpy(z), approximate="tanh").numpy()
    assert np.max(np.abs(C.gelu(z) - ref)) <= 2e-6 * np.max(np.abs(ref))


def test_gelu_grad_is_the_derivative():
    z = torch.linspace(-8, 8, 4001, dtype=torch.float64, requires_grad=True)
    torch.nn.functional.gelu(z, approximate="tanh").sum().backward()
    assert np.max(np.abs(C.gelu_grad(z.detach().numpy()) - z.grad.numpy())) < 1e-6


def _torch_net(params, x, t):
    """The same network in float64 autograd; BF16 rounding of every product
    operand is applied in the forward pass with a straight-through gradient."""
    def bf(a):
        return a + (torch.from_numpy(C.bf16(a.detach().numpy().astype(np.float32))).double() - a).detach()
    p = torch.tensor(params.astype(np.float64), requires_grad=True)
    h = torch.from_numpy(x.astype(np.float64))
    o = 0
    for li, (n_out, n_in) in enumerate(C.LAYERS):
        W = p[o:o + n_out * n_in].reshape(n_out, n_in)
        o += n_out * n_in
        b = p[o:o + n_out]
        o += n_out
        z = bf(h) @ bf(W).T + b
        h = torch.nn.functional.gelu(z, approximate="tanh") if li < 3 else z
    L = ((h - torch.from_numpy(t)) ** 2).sum() / (x.shape[0] * C.N_OUT)
    L.backward()
    return L.item(), p.grad.numpy(), h.detach().numpy()


@pytest.mark.parametrize("n", [1, 37, 256])
def test_forward_loss_and_gradient_match_autograd(n):
    r = rng(n)
    params = C.init_params(n)
    params[C.N_PARAMS - 2:] = [0.3, -0.2]  # non-zero output bias
    x = r.standard_normal((n, C.N_IN)).astype(np.float32)
    t = r.standard_normal((n, C.N_OUT)).astype(np.float32)
    L, g = C.loss_and_grad(params, x, t)
    Lr, gr, yr = _torch_net(params, x, t)
    y = C.forward(params, x)
    # FP32 vs FP64 pre-activations round to BF16 alike except near a rounding
    # boundary, where one activation moves by a BF16 step: rare, bounded misses
    err = np.max(np.abs(y - yr), axis=1)
    assert np.quantile(err, 0.99) <= 1e-5 and np.max(err) <= 1e-2
    assert abs(L - Lr) <= 1e-3 * Lr
    # the oracle also rounds the backward operands (dz, W, h) to BF16; the
    # straight-through reference does not: agreement to that rounding, while a
    # transposed weight, a missing GELU' or a wrong scale is an O(1) miss
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    assert rel < 2e-2, rel
    for (lo, hi) in [(0, 448), (512, 4608), (4672, 8768), (8832, 8960)]:  # each weight block
        assert np.linalg.norm(g[lo:hi] - gr[lo:hi]) <= 3e-2 * np.linalg.norm(gr[lo:hi])
    bias = [(448, 512), (4608, 4672), (8768, 8832), (8960, 8962)]
    for (lo, hi) in bias:
        assert np.linalg.norm(g[lo:hi] - gr[lo:hi]) <= 3e-2 * np.linalg.norm(gr[lo:hi]) + 1e-12


def test_adam_matches_torch():
    r = rng(3)
    p0 = r.standard_normal(500).astype(np.float32)
    ad = C.Adam(500, lr=1e-3)
    tp = torch.tensor(p0.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p = p0.copy()
    for k in range(25):
        g = r.standard_normal(500).astype(np.float32) * (10.0 ** (k % 5 - 2))
        p = ad.step(p, g)
        opt.zero_grad()
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
    assert np.max(np.abs(p - tp.detach().numpy())) <= 1e-6 * np.max(np.abs(p0)) * 10


def _rand_grads(n, seed):
    return rng(seed).standard_normal((n, 3, 3))


def test_invariants_satisfy_the_vorticity_and_cayley_hamilton_forms():
    A = _rand_grads(500, 4)
    theta, s2, w2, s3, sw2, s2w2 = C.invariants(A)
    S = 0.5 * (A + np.swapaxes(A, 1, 2))
    om = np.stack([A[:, 2, 1] - A[:, 1, 2], A[:, 0, 2] - A[:, 2, 0], A[:, 1, 0] - A[:, 0, 1]], axis=1)
    trS = np.trace(S, axis1=1, axis2=2)
    # W^2 = (om om^T - |om|^2 I) / 4
    np.testing.assert_allclose(w2, 0.5 * np.sum(om * om, axis=1), rtol=1e-12)
    np.testing.assert_allclose(sw2, 0.25 * (np.einsum("ni,nij,nj->n", om, S, om) - np.sum(om * om, 1) * trS),
                               rtol=1e-10, atol=1e-12)
    S2 = S @ S
    np.testing.assert_allclose(s2w2, 0.25 * (np.einsum("ni,nij,nj->n", om, S2, om) - np.sum(om * om, 1) * s2),
                               rtol=1e-10, atol=1e-12)
    # traceless S: tr S^3 = 3 det S
    D = S - trS[:, None, None] / 3 * np.eye(3)
    t2 = C.invariants(D + 0.5 * (A - np.swapaxes(A, 1, 2)))
    np.testing.assert_allclose(t2[3], 3 * np.linalg.det(D), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(theta, np.trace(A, axis1=1, axis2=2), rtol=1e-14)


def test_features_are_rotation_invariant_and_bounded():
    A = _rand_grads(300, 5)
    Q, _ = np.linalg.qr(rng(6).standard_normal((3, 3)))
    f = C.features(A, 1e-6)
    fr = C.features(Q @ A @ Q.T, 1e-6)
    np.testing.assert_allclose(fr, f, rtol=1e-10, atol=1e-12)
    assert np.all(np.abs(f[:, 1]) <= 1 + 1e-12)          # (s2 - w2) / (s2 + w2)
    assert np.all(np.abs(f[:, 0]) <= np.sqrt(3) + 1e-12)  # |tr A| <= sqrt(3) |S|
    np.testing.assert_allclose(f[:, 6], np.log(1e-6) / 3)


def test_pure_strain_and_pure_rotation():
    S = np.diag([1.0, -0.5, -0.5])[None]
    Wm = np.array([[[0, 1.0, 0], [-1.0, 0, 0], [0, 0, 0]]])
    fs = C.features(S, 1.0)[0]
    fw = C.features(Wm, 1.0)[0]
    assert fs[1] == pytest.approx(1.0) and fw[1] == pytest.approx(-1.0)
    assert fs[3] == 0 and fs[4] == 0           # no rotation: tr S W^2 = tr S^2 W^2 = 0
    assert fw[2] == 0 and fw[3] == 0           # no strain
    assert fs[5] == pytest.approx(np.log(np.sqrt(1.5)))
    z = C.features(np.zeros((1, 3, 3)), 1.0)[0]
    assert np.all(z[:5] == 0) and z[5] < -600   # vanishing gradient: out of any range


def test_outputs_are_the_dsm_forms():
    """mu = rho C Delta^2 |S| and k = cp (C / Pr_t) rho Delta^2 |S| with
    Delta^2 = (vol / J)^(2/3), the DSM's definitions."""
    y = np.array([[0.5, -1.0]], np.float32)
    tm, ts = np.array([0.1, 0.2], np.float32), np.array([0.02, 0.05], np.float32)
    mu, k = C.closure_outputs(y, tm, ts, 1.2, 3.0, 8e-6, 1004.5)
    Cc, D = 0.5 * 0.02 + 0.1, -1.0 * 0.05 + 0.2
    d2 = (8e-6) ** (2 / 3)
    assert mu[0] == pytest.approx(1.2 * Cc * d2 * 3.0, rel=1e-6)
    assert k[0] == pytest.approx(1004.5 * 1.2 * D * d2 * 3.0, rel=1e-6)
    np.testing.assert_allclose(C.smag(np.diag([1.0, 2.0, 3.0])[None]), np.sqrt(2 * 14.0))
