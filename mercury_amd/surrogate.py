"""Surrogate fitness model for voxel genomes.

Features (64 per robot) are computed from the robot as built (after Strip):
material fractions, per-material spatial moments, actuation-phase moments,
ground contact, front/back balance and spring statistics.  The network is
libmercury_hip's MLP (mh_mlp_*): bf16 weights, fp32 accumulation, trained
on the GPU by hand-written forward/backward/Adam kernels.
"""
import ctypes

import numpy as np

from . import _lib as _libmod
from ._lib import check, f32p, i32p, load, ptr, u8p

N_FEATURES = 64


def genome_features(elements, voxels_per_axis, rows):
    """64 features per voxel robot.  rows: the material set's base rows
    (k, dL0, omega, phi); actuated bases are those with dL0 > 0."""
    nv = voxels_per_axis
    n = nv + 1
    nb = len(rows)
    amp = np.array([r[1] for r in rows], np.float64)
    phase = np.array([r[3] for r in rows], np.float64)
    idx = np.arange(n ** 3)
    gx, gy, gz = idx % n, (idx // n) % n, idx // (n * n)
    inside = (gx < nv) & (gy < nv) & (gz < nv)
    cx, cy, cz = (gx[inside] / max(nv - 1, 1), gy[inside] / max(nv - 1, 1), gz[inside] / max(nv - 1, 1))
    coords = np.stack([cx, cy, cz], 1)
    out = np.zeros((len(elements), N_FEATURES), np.float32)
    for r, e in enumerate(elements):
        enc = e.stripped_encoding[inside]
        mat = np.zeros(len(enc), np.int64)
        nz = enc > 1
        mat[nz] = np.log2(enc[nz]).astype(np.int64)  # base index 1..nb (single-bit voxels)
        f = []
        counts = np.bincount(mat, minlength=nb + 1).astype(np.float64)
        f += list(counts / len(mat))                                   # 1 + nb
        solid = mat > 0
        for m in range(1, nb + 1):
            sel = mat == m
            c = coords[sel]
            f += list(c.mean(0)) if len(c) else [0.0] * 3               # 3 nb
            f += list(c.std(0)) if len(c) else [0.0] * 3                # 3 nb
        c = coords[solid]
        if len(c):
            f += list(c.mean(0)) + list(c.std(0)) + list(c.min(0)) + list(c.max(0))
        else:
            f += [0.0] * 12                                             # 12
        a = amp[mat - 1] * solid if nb else np.zeros(len(mat))
        ph = phase[np.maximum(mat - 1, 0)]
        wc, ws = a * np.cos(ph), a * np.sin(ph)
        cnt = max(solid.sum(), 1)
        for k in range(3):
            f += [float((wc * coords[:, k]).sum() / cnt), float((ws * coords[:, k]).sum() / cnt)]
        f += [float(wc.sum() / cnt), float(ws.sum() / cnt)]               # 8
        bottom = coords[:, 1] == 0
        for m in range(1, nb + 1):
            f.append(float(((mat == m) & bottom).sum()) / max(bottom.sum(), 1))
        f.append(float((solid & bottom).sum()) / max(bottom.sum(), 1))    # nb + 1
        front = coords[:, 0] >= 0.5
        for m in range(1, nb + 1):
            sel = mat == m
            f.append(float((sel & front).sum()) / max(sel.sum(), 1))      # nb
        active = e.spring_material != 0
        ln = e.mean_length[active]
        f += [float(active.mean()) if len(active) else 0.0,
              float(ln.mean()) if len(ln) else 0.0, float(ln.std()) if len(ln) else 0.0]  # 3
        act = [m for m in range(1, nb + 1) if amp[m - 1] > 0]
        cm = [coords[mat == m].mean(0) if (mat == m).any() else np.zeros(3) for m in act[:2]]
        f += list(cm[0] - cm[1]) if len(cm) == 2 else [0.0] * 3           # 3
        f = f[:N_FEATURES] + [0.0] * max(0, N_FEATURES - len(f))
        out[r] = np.asarray(f, np.float32)
    return out


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

    def fit(self, X, y, epochs=30, batch=512, lr=1e-3, seed=0):
        X = np.asarray(X, np.float32)
        y = np.asarray(y, np.float32)
        self.mu = X.mean(0)
        self.sd = X.std(0) + 1e-6
        self.ymu = float(y.mean())
        self.ysd = float(y.std() + 1e-12)
        return self.net.train((X - self.mu) / self.sd, (y - self.ymu) / self.ysd, epochs=epochs,
                              batch=batch, lr=lr, seed=seed)

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
