"""Surrogate fitness model for voxel genomes.

Features (64 per robot): per height layer and base material, the share of
the layer's voxels holding the material and their front-to-back balance
(layered_features), computed from the robot as built (after Strip) or from
its genome.  The network is
libmercury_hip's MLP (mh_mlp_*): bf16 weights, fp32 accumulation, trained
on the GPU by hand-written forward/backward/Adam kernels.
"""
import ctypes

import numpy as np

from . import _lib as _libmod
from ._lib import check, f32p, i32p, load, ptr, u8p

N_FEATURES = 64


def layer_bins(voxels_per_axis, n_base):
    """Height bins of the layered features: one per voxel layer while
    2 * n_base * bins fits N_FEATURES, else layers grouped evenly."""
    return max(1, min(voxels_per_axis, N_FEATURES // (2 * max(n_base, 1))))


def layered_features(mat, voxels_per_axis, n_base):
    """The surrogate's 64 inputs from voxel materials mat[robot, voxel]
    (0 = empty, 1..n_base; voxel (x, y, z) at x + v (y + v z), y up).  Per
    height bin and base material: the share of the bin's voxels holding it
    and their front-to-back balance, the mean of x scaled to [-1, 1] over
    the bin's voxels (empty ones count 0).  Which materials touch the ground
    and how each layer's actuators sit front to back decide how far and in
    which direction a robot travels along x; the balance is odd under the
    x mirror, as the displacement is.  Zero-padded to N_FEATURES."""
    v = voxels_per_axis
    mat = np.asarray(mat).reshape(len(mat), v, v, v)  # [robot, z, y, x]
    bins = layer_bins(v, n_base)
    ybin = (np.arange(v) * bins) // v
    half = (v - 1) / 2.0
    xc = (np.arange(v, dtype=np.float64) - half) / (half if v > 1 else 1.0)
    per_bin = np.bincount(ybin, minlength=bins).astype(np.float64) * v * v
    out = np.zeros((len(mat), N_FEATURES), np.float32)
    f = 0
    for b in range(bins):
        sl = mat[:, :, ybin == b, :]
        for m in range(1, n_base + 1):
            hit = (sl == m)
            out[:, f] = hit.sum(axis=(1, 2, 3)) / per_bin[b]
            out[:, f + 1] = (hit * xc).sum(axis=(1, 2, 3)) / per_bin[b]
            f += 2
    return out


def genome_features(elements, voxels_per_axis, rows):
    """layered_features of built voxel robots (their voxels after Strip;
    every voxel holds one base material).  rows: the material set's base
    rows (k, dL0, omega, phi)."""
    nv = voxels_per_axis
    n = nv + 1
    idx = np.arange(n ** 3)
    gx, gy, gz = idx % n, (idx // n) % n, idx // (n * n)
    inside = (gx < nv) & (gy < nv) & (gz < nv)
    mat = np.zeros((len(elements), nv ** 3), np.int64)
    for r, e in enumerate(elements):
        enc = e.stripped_encoding[inside]
        nz = enc > 1
        mat[r, nz] = np.log2(enc[nz]).astype(np.int64)  # base index 1..nb (single-bit voxels)
    return layered_features(mat, nv, len(rows))


def spearman(a, b):
    """Spearman rank correlation (average ranks for ties)."""
    def rank(x):
        x = np.asarray(x, np.float64)
        order = np.argsort(x, kind="mergesort")
        r = np.empty(len(x), np.float64)
        r[order] = np.arange(len(x), dtype=np.float64)
        # average ranks over ties
        xs = x[order]
        i = 0
        while i < len(xs):
            j = i
            while j + 1 < len(xs) and xs[j + 1] == xs[i]:
                j += 1
            if j > i:
                r[order[i:j + 1]] = (i + j) / 2.0
            i = j + 1
        return r
    ra, rb = rank(a), rank(b)
    ra -= ra.mean()
    rb -= rb.mean()
    den = np.sqrt((ra ** 2).sum() * (rb ** 2).sum())
    return float((ra * rb).sum() / den) if den > 0 else 0.0


def mae(a, b):
    return float(np.mean(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64))))


class MLP:
    """mh_mlp handle: sizes[0] -> ... -> sizes[-1], GELU hidden layers."""

    def __init__(self, sizes, seed=0, device=0):
        self.sizes = list(sizes)
        sz = np.ascontiguousarray(self.sizes, np.int32)
        h = ctypes.c_void_p()
        check(load().mh_mlp_create(device, len(sizes) - 1, ptr(sz, i32p), int(seed), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        lib = getattr(_libmod, "_lib", None)
        if getattr(self, "_h", None) is not None and lib is not None:
            lib.mh_mlp_destroy(self._h)
            self._h = None

    def n_params(self):
        return load().mh_mlp_param_count(self._h)

    def get_params(self):
        out = np.zeros(self.n_params(), np.float32)
        check(load().mh_mlp_get_params(self._h, ptr(out, f32p)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(load().mh_mlp_set_params(self._h, ptr(p, f32p)))

    def layers(self, p=None):
        """Split a parameter vector into [(W, b), ...]."""
        p = self.get_params() if p is None else p
        out, o = [], 0
        for i, j in zip(self.sizes[:-1], self.sizes[1:]):
            W = p[o:o + i * j].reshape(j, i)
            o += i * j
            b = p[o:o + j]
            o += j
            out.append((W, b))
        return out

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), self.sizes[-1]), np.float32)
        check(load().mh_mlp_forward(self._h, ptr(x, f32p), len(x), ptr(y, f32p)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32).reshape(len(x), -1)
        loss = ctypes.c_float()
        g = np.zeros(self.n_params(), np.float32)
        check(load().mh_mlp_loss_grad(self._h, ptr(x, f32p), ptr(t, f32p), len(x),
                                      ctypes.byref(loss), ptr(g, f32p)))
        return float(loss.value), g

    def train(self, x, t, epochs=20, batch=512, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32).reshape(len(x), -1)
        loss = ctypes.c_float()
        check(load().mh_mlp_train(self._h, ptr(x, f32p), ptr(t, f32p), len(x), int(batch),
                                  int(epochs), float(lr), int(seed), ctypes.byref(loss)))
        return float(loss.value)


class Surrogate:
    """Standardized features -> MLP 64-256-256-1 -> fitness."""

    def __init__(self, hidden=(256, 256), seed=0, device=0):
        self.net = MLP([N_FEATURES, *hidden, 1], seed=seed, device=device)
        self.mu = None

    def fit(self, X, y, epochs=30, batch=512, lr=1e-3, seed=0, holdout=0.1, patience=3):
        """Adam on standardized inputs and targets.  With holdout > 0 a
        seeded share of the rows is kept out of training; after every epoch
        the rank correlation on it is measured, and the weights of the best
        epoch are kept (training stops after `patience` epochs without a
        better one).  Returns the last epoch's training loss."""
        X = np.asarray(X, np.float32)
        y = np.asarray(y, np.float32)
        self.mu = X.mean(0)
        self.sd = X.std(0) + 1e-6
        self.ymu = float(y.mean())
        self.ysd = float(y.std() + 1e-12)
        Xs, ys = (X - self.mu) / self.sd, (y - self.ymu) / self.ysd
        n_val = int(len(X) * holdout) if holdout > 0 and len(X) >= 200 else 0
        if n_val == 0:
            return self.net.train(Xs, ys, epochs=epochs, batch=batch, lr=lr, seed=seed)
        perm = np.random.default_rng(seed).permutation(len(X))
        val, tr = perm[:n_val], perm[n_val:]
        Xt, yt = np.ascontiguousarray(Xs[tr]), np.ascontiguousarray(ys[tr])
        best, best_p, since, loss = -np.inf, None, 0, 0.0
        for ep in range(epochs):
            loss = self.net.train(Xt, yt, epochs=1, batch=batch, lr=lr, seed=seed + ep)
            rho = spearman(self.net.forward(Xs[val])[:, 0], ys[val])
            if rho > best:
                best, best_p, since = rho, self.net.get_params(), 0
            else:
                since += 1
                if since >= patience:
                    break
        self.net.set_params(best_p)
        self.val_rho = best
        return loss

    def predict(self, X):
        X = np.asarray(X, np.float32)
        return self.net.forward((X - self.mu) / self.sd)[:, 0] * self.ysd + self.ymu


class CNN:
    """mh_cnn handle: the convolutional genome network (8^3 voxel materials
    -> 64 learned features -> MLP 64-256-256-1)."""

    LAYERS = [(32, 135), (32, 864), (64, 864), (256, 64), (256, 256), (1, 256)]

    def __init__(self, seed=0, device=0):
        h = ctypes.c_void_p()
        check(load().mh_cnn_create(device, int(seed), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        lib = getattr(_libmod, "_lib", None)
        if getattr(self, "_h", None) is not None and lib is not None:
            lib.mh_cnn_destroy(self._h)
            self._h = None

    def n_params(self):
        return load().mh_cnn_param_count(self._h)

    def get_params(self):
        out = np.zeros(self.n_params(), np.float32)
        check(load().mh_cnn_get_params(self._h, ptr(out, f32p)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        check(load().mh_cnn_set_params(self._h, ptr(p, f32p)))

    def layers(self, p=None):
        p = self.get_params() if p is None else p
        out, o = [], 0
        for n_out, n_in in self.LAYERS:
            W = p[o:o + n_out * n_in].reshape(n_out, n_in)
            o += n_out * n_in
            out.append((W, p[o:o + n_out]))
            o += n_out
        return out

    def forward(self, genomes):
        g = np.ascontiguousarray(genomes, np.uint8).reshape(-1, 512)
        y = np.zeros(len(g), np.float32)
        check(load().mh_cnn_forward(self._h, ptr(g, u8p), len(g), ptr(y, f32p)))
        return y

    def loss_grad(self, genomes, t):
        g = np.ascontiguousarray(genomes, np.uint8).reshape(-1, 512)
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        loss = ctypes.c_float()
        grad = np.zeros(self.n_params(), np.float32)
        check(load().mh_cnn_loss_grad(self._h, ptr(g, u8p), ptr(t, f32p), len(g), ctypes.byref(loss),
                                      ptr(grad, f32p)))
        return float(loss.value), grad

    def train(self, genomes, t, epochs=10, batch=256, lr=1e-3, seed=0):
        g = np.ascontiguousarray(genomes, np.uint8).reshape(-1, 512)
        t = np.ascontiguousarray(t, np.float32).reshape(-1)
        loss = ctypes.c_float()
        check(load().mh_cnn_train(self._h, ptr(g, u8p), ptr(t, f32p), len(g), int(batch), int(epochs),
                                  float(lr), int(seed), ctypes.byref(loss)))
        return float(loss.value)


class GenomeSurrogate:
    """Fitness of 8^3 voxel genomes predicted by the CNN (standardized
    targets); same fit/predict interface as Surrogate, on genomes instead
    of feature rows."""

    def __init__(self, seed=0, device=0):
        self.net = CNN(seed=seed, device=device)

    def fit(self, genomes, y, epochs=10, batch=256, lr=1e-3, seed=0):
        y = np.asarray(y, np.float64)
        self.ymu, self.ysd = float(y.mean()), float(y.std() + 1e-12)
        return self.net.train(genomes, (y - self.ymu) / self.ysd, epochs=epochs, batch=batch, lr=lr,
                              seed=seed)

    def predict(self, genomes):
        return self.net.forward(genomes).astype(np.float64) * self.ysd + self.ymu
