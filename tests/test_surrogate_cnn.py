"""Convolutional genome surrogate (mh_cnn): forward and gradient against a
float64 PyTorch-CPU restatement of the same network with the same
(bf16-rounded) weights, and training on a target that only spatial
structure explains (GPU)."""
import numpy as np
import pytest

from mercury_amd import evolve, surrogate


def bf16_round(w):
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def torch_reference(net, params, genomes):
    """The network in float64 torch: conv weights [out][k*C + c] with
    k = (dz+1)*9 + (dy+1)*3 + (dx+1) become [out, C, dz, dy, dx]."""
    import torch
    F = torch.nn.functional
    layers = [(torch.tensor(bf16_round(W).astype(np.float64), requires_grad=True),
               torch.tensor(b.astype(np.float64), requires_grad=True)) for W, b in net.layers(params)]
    g = torch.tensor(np.asarray(genomes, np.int64)).reshape(-1, 8, 8, 8)     # [n][z][y][x]
    x = F.one_hot(g, 5).permute(0, 4, 1, 2, 3).double()                      # [n, 5, z, y, x]
    gelu = lambda t: F.gelu(t, approximate="tanh")

    def conv(x, W, b, cin):
        w = W.reshape(W.shape[0], 27, cin).permute(0, 2, 1).reshape(W.shape[0], cin, 3, 3, 3)
        return F.conv3d(x, w, b, padding=1)
    (W1, b1), (W2, b2), (W3, b3), (H1, c1), (H2, c2), (H3, c3) = layers
    a = gelu(conv(x, W1, b1, 5))
    a = gelu(conv(a, W2, b2, 32))
    a = F.avg_pool3d(a, 2)
    a = gelu(conv(a, W3, b3, 32))
    f = a.mean(dim=(2, 3, 4))
    h = gelu(f @ H1.T + c1)
    h = gelu(h @ H2.T + c2)
    y = (h @ H3.T + c3)[:, 0]
    return y, layers


@pytest.mark.gpu
def test_cnn_forward_and_gradient_match_float64_reference():
    print("cnn fwd/bwd vs torch float64", flush=True)
    rng = np.random.default_rng(0)
    net = surrogate.CNN(seed=3)
    p = net.get_params()
    g = evolve.random_genomes(rng, 5, 8)
    t = rng.normal(size=5).astype(np.float32)
    y = net.forward(g)
    yr, layers = torch_reference(net, p, g)
    assert np.abs(y - yr.detach().numpy()).max() <= 1e-3 * max(1.0, np.abs(yr.detach().numpy()).max())
    loss, grad = net.loss_grad(g, t)
    import torch
    lr = 0.5 * ((yr - torch.tensor(t, dtype=torch.float64)) ** 2).mean()
    lr.backward()
    ref = np.concatenate([np.concatenate([W.grad.numpy().reshape(-1), b.grad.numpy()]) for W, b in layers])
    assert loss == pytest.approx(lr.item(), rel=1e-3)
    scale = np.abs(ref).max()
    assert np.abs(grad - ref).max() <= 2e-3 * scale
    # per layer: a transposed kernel or a dropped activation derivative
    # would decorrelate one block
    o = 0
    for (n_out, n_in) in net.LAYERS:
        blk = slice(o, o + n_out * n_in)
        assert np.corrcoef(grad[blk], ref[blk])[0, 1] > 0.9999
        o += n_out * n_in + n_out


def directional_pairs(g):
    """Count of (phase-0 actuator, phase-pi actuator) neighbours along +x minus
    the reverse order: a target only spatial arrangement explains."""
    G = g.reshape(-1, 8, 8, 8)
    a, b = G[:, :, :, :-1], G[:, :, :, 1:]
    return (((a == 3) & (b == 4)).sum((1, 2, 3)) - ((a == 4) & (b == 3)).sum((1, 2, 3))).astype(np.float64)


@pytest.mark.gpu
def test_cnn_learns_spatial_target_that_summary_features_miss():
    print("cnn training", flush=True)
    rng = np.random.default_rng(1)
    g = evolve.random_genomes(rng, 6000, 8)
    y = directional_pairs(g)
    s = surrogate.GenomeSurrogate(seed=2)
    s.fit(g[:5000], y[:5000], epochs=25, batch=256, lr=1e-3, seed=0)
    rho = surrogate.spearman(s.predict(g[5000:]), y[5000:])
    m = surrogate.Surrogate(seed=2)
    X = evolve.genome_features(g, 8)
    m.fit(X[:5000], y[:5000], epochs=15, batch=256, lr=1e-3, seed=0)
    rho_mlp = surrogate.spearman(m.predict(X[5000:]), y[5000:])
    print("rho cnn", rho, "rho summary-feature mlp", rho_mlp, flush=True)
    # measured on MI355X: 0.72 after 15 epochs against 0.08 for the MLP
    assert rho > 0.65 and rho > rho_mlp + 0.4
