"""Pin the oracle's own functions: sine, KNN, surrogate MLP and Adam."""
import math

import numpy as np
import pytest

import oracle


def test_sine_error_bound():
    xs = np.concatenate([np.linspace(-300, 300, 200001, dtype=np.float32),
                         np.float32([0.0, math.pi / 2, math.pi, 1e-7, -1e-7, 4 * math.pi * 0.4998])])
    got = np.array([oracle.sinf(x) for x in xs], np.float64)
    want = np.sin(xs.astype(np.float64))
    assert np.abs(got - want).max() <= 2e-7
    assert oracle.sinf(0.0) == 0.0
    # a wrong quadrant or sign would show up as an O(1) error
    for q in range(8):
        x = np.float32(q * math.pi / 2 + 0.3)
        assert abs(oracle.sinf(x) - math.sin(float(x))) < 1e-6


def brute_knn(pos, K):
    pos = pos.astype(np.float32)
    n = len(pos)
    d = pos[:, None, :] - pos[None, :, :]
    yz = (d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2]).astype(np.float32)
    dist = np.sqrt((d[..., 0] * d[..., 0] + yz).astype(np.float32)).astype(np.float32)
    dist[dist == 0] = np.inf
    nbr = np.zeros((n, K), np.int64)
    for i in range(n):
        order = sorted(range(n), key=lambda j: (dist[i, j], j))
        nbr[i] = order[:K]
    return nbr, np.take_along_axis(dist, nbr, 1)


def test_knn_matches_brute_force():
    rng = np.random.default_rng(4)
    pos = (rng.random((60, 3)) * 0.5).astype(np.float32)
    pos[7] = pos[3]  # a coincident pair: zero distance counts as +inf
    pos[11, :] = pos[12, :] + np.float32([0.01, 0, 0])
    nbr, dist = oracle.knn(pos, 10)
    bn, bd = brute_knn(pos, 10)
    assert np.array_equal(nbr, bn)
    assert np.array_equal(dist.view(np.uint32), bd.view(np.uint32))
    assert 7 not in nbr[3] and 3 not in nbr[7]


def make_net(rng, sizes):
    W = [rng.normal(0, 1 / math.sqrt(i), (o, i)) for i, o in zip(sizes[:-1], sizes[1:])]
    b = [rng.normal(0, 0.1, o) for o in sizes[1:]]
    return W, b


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(5)
    W, b = make_net(rng, [5, 7, 6, 1])
    x = rng.normal(size=(9, 5))
    t = rng.normal(size=(9, 1))
    loss, gW, gb = oracle.mlp_loss_grad(W, b, x, t)
    h = 1e-6
    for l in range(3):
        r, c = W[l].shape
        for idx in [(0, 0), (r - 1, c - 1), (r // 2, c // 2)]:
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            num = (oracle.mlp_loss_grad(Wp, b, x, t)[0] - oracle.mlp_loss_grad(Wm, b, x, t)[0]) / (2 * h)
            assert gW[l][idx] == pytest.approx(num, rel=1e-5, abs=1e-9)
        bp = [v.copy() for v in b]
        bm = [v.copy() for v in b]
        bp[l][0] += h
        bm[l][0] -= h
        num = (oracle.mlp_loss_grad(W, bp, x, t)[0] - oracle.mlp_loss_grad(W, bm, x, t)[0]) / (2 * h)
        assert gb[l][0] == pytest.approx(num, rel=1e-5, abs=1e-9)


def test_mlp_one_layer_is_linear_regression():
    rng = np.random.default_rng(6)
    W, b = make_net(rng, [4, 1])
    x = rng.normal(size=(16, 4))
    t = rng.normal(size=(16, 1))
    y = oracle.mlp_forward(W, b, x)
    assert np.allclose(y, x @ W[0].T + b[0])
    loss, gW, gb = oracle.mlp_loss_grad(W, b, x, t)
    e = y - t
    assert loss == pytest.approx(0.5 * np.mean(e ** 2))
    assert np.allclose(gW[0], e.T @ x / 16)
    assert np.allclose(gb[0], e.mean(0))


def _bf16(v):
    """float64 -> float32 -> bf16 (nearest even), back to float64."""
    u = np.ascontiguousarray(v, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_mlp_bf16_operands_mode_matches_numpy_restatement():
    """bf16_operands rounds the activations entering W a, both factors of
    the weight gradient and the delta entering W^T delta; bias gradients sum
    unrounded deltas.  Restated here with numpy matrix products; the
    rounding changes the results by far more than the two disagree, so a
    dropped or misplaced rounding fails."""
    gelu = lambda z: 0.5 * z * (1 + np.tanh(math.sqrt(2 / math.pi) * (z + 0.044715 * z ** 3)))

    def dgelu(z):
        th = np.tanh(math.sqrt(2 / math.pi) * (z + 0.044715 * z ** 3))
        return 0.5 * (1 + th) + 0.5 * z * (1 - th * th) * math.sqrt(2 / math.pi) * (1 + 3 * 0.044715 * z * z)

    rng = np.random.default_rng(8)
    W, b = make_net(rng, [12, 20, 16, 1])
    W = [_bf16(w) for w in W]
    x = rng.normal(size=(30, 12))
    t = rng.normal(size=(30, 1))
    a, Z, A = x, [], [x]
    for l in range(3):
        z = _bf16(a) @ W[l].T + b[l]
        Z.append(z)
        a = gelu(z) if l < 2 else z
        A.append(a)
    y = oracle.mlp_forward(W, b, x, bf16_operands=True)
    assert np.allclose(y, A[-1], rtol=1e-12, atol=1e-12)
    assert np.abs(y - oracle.mlp_forward(W, b, x)).max() > 1e-4
    d = (A[-1] - t) / len(x)
    want = []
    for l in (2, 1, 0):
        want = [(_bf16(d).T @ _bf16(A[l])), d.sum(0)] + want
        if l > 0:
            d = (_bf16(d) @ W[l]) * dgelu(Z[l - 1])
    loss, gW, gb = oracle.mlp_loss_grad(W, b, x, t, bf16_operands=True)
    assert loss == pytest.approx(0.5 * np.mean((A[-1] - t) ** 2), rel=1e-12)
    for l in range(3):
        assert np.allclose(gW[l], want[2 * l], rtol=1e-10, atol=1e-14)
        assert np.allclose(gb[l], want[2 * l + 1], rtol=1e-10, atol=1e-14)
    _, gW0, _ = oracle.mlp_loss_grad(W, b, x, t)
    assert max(np.abs(gW0[l] - gW[l]).max() for l in range(3)) > 1e-6


def test_gelu_hidden_layer_values():
    W = [np.array([[1.0]]), np.array([[1.0]])]
    b = [np.zeros(1), np.zeros(1)]
    for v in (-3.0, -0.5, 0.0, 0.7, 2.0):
        u = math.sqrt(2 / math.pi) * (v + 0.044715 * v ** 3)
        assert oracle.mlp_forward(W, b, [[v]])[0, 0] == pytest.approx(0.5 * v * (1 + math.tanh(u)))


def test_adam_first_step_is_lr_sign():
    g = np.array([0.3, -2.0, 1e-3, -5e-2])
    p = np.zeros(4)
    m = np.zeros(4)
    v = np.zeros(4)
    oracle.adam(p, g, m, v, 0.01, 0.9, 0.999, 1e-12, 1)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-6)
    oracle.adam(p, g, m, v, 0.01, 0.9, 0.999, 1e-12, 2)
    assert np.allclose(p, -0.02 * np.sign(g), rtol=1e-6)
