"""Python view of the oracle (CPU restatement of the Mercury step).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package (mercury_amd) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

u8p = ctypes.POINTER(ctypes.c_uint8)
u16p = ctypes.POINTER(ctypes.c_uint16)
u32p = ctypes.POINTER(ctypes.c_uint32)
f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)


class Material(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint32), ("k", ctypes.c_float), ("dL0", ctypes.c_float),
                ("omega", ctypes.c_float), ("phi", ctypes.c_float),
                ("encoding", ctypes.c_uint32)]


class Topology(ctypes.Structure):
    _fields_ = [("n_robots", ctypes.c_int), ("mass_off", u32p), ("spring_off", u32p),
                ("mass", f32p), ("pairs", u16p), ("lbar", f32p), ("mat", u8p)]


class MatTab(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("k", f32p), ("dL0", f32p), ("omega", f32p),
                ("phi", f32p)]


class Env(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_float), ("gravity", ctypes.c_float),
                ("ground_k", ctypes.c_float), ("mu_s", ctypes.c_float),
                ("mu_k", ctypes.c_float), ("drag", ctypes.c_float),
                ("damping", ctypes.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing ({LIB_PATH}); run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_reference_base.argtypes = [ctypes.POINTER(Material)]
        L.orc_composite_tables.argtypes = [ctypes.POINTER(Material), ctypes.c_int,
                                           ctypes.POINTER(Material), ctypes.POINTER(Material)]
        L.orc_encoded_idx.argtypes = [ctypes.c_uint32, ctypes.c_int]
        L.orc_composite_id.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
        L.orc_composite_id.restype = ctypes.c_uint32
        L.orc_voxel_build.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                      u32p, ctypes.POINTER(Material), ctypes.c_int, f32p, u32p,
                                      u16p, f32p, u32p, u32p, ctypes.c_int, u32p]
        L.orc_com.argtypes = [f32p, u8p, ctypes.c_int, f32p]
        L.orc_length.argtypes = [f32p, u8p, ctypes.c_int]
        L.orc_length.restype = ctypes.c_float
        L.orc_fitness.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.orc_fitness.restype = ctypes.c_float
        L.orc_sinf.argtypes = [ctypes.c_float]
        L.orc_sinf.restype = ctypes.c_float
        L.orc_xpbd_f32.argtypes = [ctypes.POINTER(Topology), ctypes.POINTER(MatTab),
                                   ctypes.c_float, ctypes.c_float, ctypes.c_int, f32p, f32p,
                                   f32p, ctypes.c_int]
        L.orc_xpbd_f64.argtypes = [ctypes.POINTER(Topology), ctypes.POINTER(MatTab),
                                   ctypes.c_float, ctypes.c_float, ctypes.c_int, f64p, f64p,
                                   f64p, ctypes.c_int]
        L.orc_force_f32.argtypes = [ctypes.POINTER(Topology), ctypes.POINTER(MatTab),
                                    ctypes.POINTER(Env), ctypes.c_float, ctypes.c_int, f32p,
                                    f32p, ctypes.c_int]
        L.orc_force_f64.argtypes = [ctypes.POINTER(Topology), ctypes.POINTER(MatTab),
                                    ctypes.POINTER(Env), ctypes.c_float, ctypes.c_int, f64p,
                                    f64p, ctypes.c_int]
        L.orc_step_times.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, f32p,
                                     ctypes.c_int]
        L.orc_knn.argtypes = [f32p, ctypes.c_int, ctypes.c_int, u16p, f32p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


# ------------------------------------------------------------------ materials
MAT_FIELDS = ("id", "k", "dL0", "omega", "phi", "encoding")


def _mats_to_np(arr):
    return {f: np.array([getattr(m, f) for m in arr]) for f in MAT_FIELDS}


def reference_base():
    out = (Material * 23)()
    n = lib().orc_reference_base(out)
    return [out[i] for i in range(n)]


def composite_tables(base):
    """base: list of Material (index 0 = air).  Returns (by_id, by_enc) dicts of arrays."""
    n = len(base)
    cc = (n - 1) * n // 2 + 1
    b = (Material * n)(*base)
    by_id, by_enc = (Material * cc)(), (Material * cc)()
    got = lib().orc_composite_tables(b, n, by_id, by_enc)
    assert got == cc
    return _mats_to_np(by_id), _mats_to_np(by_enc), by_enc


def make_base(rows):
    """rows: list of (k, dL0, omega, phi) for non-air bases; air prepended."""
    base = [Material(0, 0, 0, 0, 0, 1)]
    for i, (k, a, w, p) in enumerate(rows, start=1):
        base.append(Material(i, k, a, w, p, 1 << i))
    return base


def encoded_idx(enc, cc):
    return lib().orc_encoded_idx(enc, cc)


# ---------------------------------------------------------------------- voxel
def voxel_build(n, resolution, voxel_enc, by_enc_struct):
    cc = len(by_enc_struct)
    nv = n ** 3
    enc = np.ascontiguousarray(voxel_enc, dtype=np.uint32)
    maxs = 13 * nv
    pos = np.zeros(3 * nv, np.float32)
    mid = np.zeros(nv, np.uint32)
    pairs = np.zeros(2 * maxs, np.uint16)
    ln = np.zeros(maxs, np.float32)
    sid = np.zeros(maxs, np.uint32)
    senc = np.zeros(maxs, np.uint32)
    stripped = np.zeros(nv, np.uint32)
    ns = lib().orc_voxel_build(n, n, n, resolution, _p(enc, u32p), by_enc_struct, cc,
                               _p(pos, f32p), _p(mid, u32p), _p(pairs, u16p), _p(ln, f32p),
                               _p(sid, u32p), _p(senc, u32p), maxs, _p(stripped, u32p))
    assert ns >= 0
    return dict(mass_pos=pos.reshape(nv, 3), mass_id=mid, pairs=pairs[:2 * ns].reshape(ns, 2),
                length=ln[:ns], spring_id=sid[:ns], spring_enc=senc[:ns],
                stripped_enc=stripped)


def com(pos, is_air):
    out = np.zeros(3, np.float32)
    pos = np.ascontiguousarray(pos, np.float32)
    air = np.ascontiguousarray(is_air, np.uint8)
    lib().orc_com(_p(pos, f32p), _p(air, u8p), len(air), _p(out, f32p))
    return out


def length(pos, is_air):
    pos = np.ascontiguousarray(pos, np.float32)
    air = np.ascontiguousarray(is_air, np.uint8)
    return np.float32(lib().orc_length(_p(pos, f32p), _p(air, u8p), len(air)))


def fitness(com_x, base_x, length_):
    return np.float32(lib().orc_fitness(com_x, base_x, length_))


def sinf(x):
    return np.float32(lib().orc_sinf(float(x)))


def step_times(t0, dt, duration):
    n = lib().orc_step_times(t0, dt, duration, None, 0)
    out = np.zeros(max(n, 1), np.float32)
    lib().orc_step_times(t0, dt, duration, _p(out, f32p), n)
    return out[:n]


# ----------------------------------------------------------------- population
class Population:
    """Packed topology (SoA, robot-local pair indices) shared by all steps."""

    def __init__(self, mass_off, spring_off, mass, pairs, lbar, mat):
        self.mass_off = np.ascontiguousarray(mass_off, np.uint32)
        self.spring_off = np.ascontiguousarray(spring_off, np.uint32)
        self.mass = np.ascontiguousarray(mass, np.float32)
        self.pairs = np.ascontiguousarray(pairs, np.uint16).reshape(-1)
        self.lbar = np.ascontiguousarray(lbar, np.float32)
        self.mat = np.ascontiguousarray(mat, np.uint8)
        self.topo = Topology(len(self.mass_off) - 1, _p(self.mass_off, u32p),
                             _p(self.spring_off, u32p), _p(self.mass, f32p),
                             _p(self.pairs, u16p), _p(self.lbar, f32p), _p(self.mat, u8p))


class MaterialTable:
    def __init__(self, k, dL0, omega, phi):
        self.k = np.ascontiguousarray(k, np.float32)
        self.dL0 = np.ascontiguousarray(dL0, np.float32)
        self.omega = np.ascontiguousarray(omega, np.float32)
        self.phi = np.ascontiguousarray(phi, np.float32)
        self.tab = MatTab(len(self.k), _p(self.k, f32p), _p(self.dL0, f32p),
                          _p(self.omega, f32p), _p(self.phi, f32p))


def xpbd_f32(pop, mt, dt, t0, nsteps, pos, vel, stress=None, nthreads=1):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    vel = np.ascontiguousarray(vel, np.float32).copy()
    st = None if stress is None else np.ascontiguousarray(stress, np.float32).copy()
    lib().orc_xpbd_f32(ctypes.byref(pop.topo), ctypes.byref(mt.tab), dt, t0, nsteps,
                       _p(pos, f32p), _p(vel, f32p), _p(st, f32p), nthreads)
    return pos, vel, st


def xpbd_f64(pop, mt, dt, t0, nsteps, pos, vel, stress=None, nthreads=1):
    pos = np.ascontiguousarray(pos, np.float64).copy()
    vel = np.ascontiguousarray(vel, np.float64).copy()
    st = None if stress is None else np.ascontiguousarray(stress, np.float64).copy()
    lib().orc_xpbd_f64(ctypes.byref(pop.topo), ctypes.byref(mt.tab), dt, t0, nsteps,
                       _p(pos, f64p), _p(vel, f64p), _p(st, f64p), nthreads)
    return pos, vel, st


def make_env(dt, gravity=0.0, ground_k=0.0, mu_s=0.0, mu_k=0.0, drag=0.0, damping=1.0):
    return Env(dt, gravity, ground_k, mu_s, mu_k, drag, damping)


def force_f32(pop, mt, env, t0, nsteps, pos, vel, nthreads=1):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    vel = np.ascontiguousarray(vel, np.float32).copy()
    lib().orc_force_f32(ctypes.byref(pop.topo), ctypes.byref(mt.tab), ctypes.byref(env), t0,
                        nsteps, _p(pos, f32p), _p(vel, f32p), nthreads)
    return pos, vel


def force_f64(pop, mt, env, t0, nsteps, pos, vel, nthreads=1):
    pos = np.ascontiguousarray(pos, np.float64).copy()
    vel = np.ascontiguousarray(vel, np.float64).copy()
    lib().orc_force_f64(ctypes.byref(pop.topo), ctypes.byref(mt.tab), ctypes.byref(env), t0,
                        nsteps, _p(pos, f64p), _p(vel, f64p), nthreads)
    return pos, vel


def knn(pos, K):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    n = len(pos) // 3
    nbr = np.zeros(n * K, np.uint16)
    dist = np.zeros(n * K, np.float32)
    rc = lib().orc_knn(_p(pos, f32p), n, K, _p(nbr, u16p), _p(dist, f32p))
    assert rc == 0
    return nbr.reshape(n, K), dist.reshape(n, K)


# ------------------------------------------------------------------ surrogate
def _dpp(arrs):
    t = (f64p * len(arrs))(*[a.ctypes.data_as(f64p) for a in arrs])
    return t


def _mlp_sig(L):
    if getattr(L, "_mlp_sig", False):
        return
    dpp = ctypes.POINTER(f64p)
    L.orc_mlp_forward.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), dpp, dpp,
                                  ctypes.c_int, f64p, f64p, ctypes.c_int]
    L.orc_mlp_loss_grad.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), dpp, dpp,
                                    ctypes.c_int, f64p, f64p, dpp, dpp, ctypes.c_int]
    L.orc_mlp_loss_grad.restype = ctypes.c_double
    L.orc_adam.argtypes = [ctypes.c_int, f64p, f64p, f64p, f64p, ctypes.c_double,
                           ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    L._mlp_sig = True


def mlp_forward(W, b, x, bf16_operands=False):
    """W: list of [out x in] arrays, b: list of [out]; GELU hidden layers.
    bf16_operands rounds the activations entering each product to bf16 (via
    float32, nearest even), as the device's matrix-core GEMMs do."""
    L = lib()
    _mlp_sig(L)
    W = [np.ascontiguousarray(w, np.float64) for w in W]
    b = [np.ascontiguousarray(v, np.float64) for v in b]
    sizes = (ctypes.c_int * (len(W) + 1))(*([W[0].shape[1]] + [w.shape[0] for w in W]))
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros((x.shape[0], W[-1].shape[0]), np.float64)
    L.orc_mlp_forward(len(W), sizes, _dpp(W), _dpp(b), x.shape[0], x.ctypes.data_as(f64p),
                      y.ctypes.data_as(f64p), int(bool(bf16_operands)))
    return y


def mlp_loss_grad(W, b, x, target, bf16_operands=False):
    """Loss and gradients; bf16_operands also rounds the back-propagated
    gradients entering each product (bias gradients sum unrounded ones)."""
    L = lib()
    _mlp_sig(L)
    W = [np.ascontiguousarray(w, np.float64) for w in W]
    b = [np.ascontiguousarray(v, np.float64) for v in b]
    sizes = (ctypes.c_int * (len(W) + 1))(*([W[0].shape[1]] + [w.shape[0] for w in W]))
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(target, np.float64).reshape(x.shape[0], -1)
    gW = [np.zeros_like(w) for w in W]
    gb = [np.zeros_like(v) for v in b]
    loss = L.orc_mlp_loss_grad(len(W), sizes, _dpp(W), _dpp(b), x.shape[0],
                               x.ctypes.data_as(f64p), t.ctypes.data_as(f64p), _dpp(gW), _dpp(gb),
                               int(bool(bf16_operands)))
    return loss, gW, gb


def adam(p, g, m, v, lr, beta1, beta2, eps, t):
    L = lib()
    _mlp_sig(L)
    for a in (p, g, m, v):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    L.orc_adam(len(p), p.ctypes.data_as(f64p), g.ctypes.data_as(f64p), m.ctypes.data_as(f64p),
               v.ctypes.data_as(f64p), lr, beta1, beta2, eps, t)
