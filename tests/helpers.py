"""Shared builders for parity tests: one population, two views (product and
oracle), built independently (product: libmercury_hip, oracle: oracle/)."""
import numpy as np

import oracle
import mercury_amd as mh
from mercury_amd.simulator import pack

WATER = dict(gravity=0.0, ground_k=0.0, drag=1.0, damping=0.99)
LAND = dict(gravity=9.81, ground_k=1.0e5, mu_s=1.0, mu_k=0.8, drag=0.0, damping=0.99)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


class Case:
    """A population plus its oracle tables."""

    def __init__(self, elements, material_set, base_rows):
        self.elements = elements
        self.ms = material_set
        base = oracle.reference_base() if base_rows is None else oracle.make_base(base_rows)
        by_id, _, _ = oracle.composite_tables(base)
        self.mt = oracle.MaterialTable(by_id["k"], by_id["dL0"], by_id["omega"], by_id["phi"])
        self.a = pack(elements)
        a = self.a
        self.pop = oracle.Population(a["mass_offset"], a["spring_offset"], a["mass"], a["pairs"],
                                     a["mean_length"], a["spring_material"])

    @property
    def pos0(self):
        return self.a["pos"].copy()

    @property
    def vel0(self):
        return self.a["vel"].copy()

    def n_springs(self):
        return len(self.a["mean_length"])


def workload_case(n_robots=4, voxels=4, p_empty=0.2, seed=0, mass=0.001):
    ms = mh.workload_materials()
    els = mh.voxel_population(ms, n_robots, voxels_per_axis=voxels, p_empty=p_empty, seed=seed,
                              mass=mass)
    return Case(els, ms, mh.materials.WORKLOAD_ROWS)


def reference_case(n_robots=4, voxels=4, p_empty=0.2, seed=0):
    ms = mh.MaterialSet.reference()
    els = mh.voxel_population(ms, n_robots, voxels_per_axis=voxels, p_empty=p_empty, seed=seed,
                              mass=1.0, n_materials=22)
    return Case(els, ms, None)


def uneven_case(seed=3):
    """Robots of different lattice sizes in one population."""
    ms = mh.workload_materials()
    rng = np.random.default_rng(seed)
    els = []
    for v in (1, 2, 4, 8, 3, 6, 2, 5):
        enc = mh.robots.random_voxel_encodings(rng, v, 4, 0.25)
        els.append(mh.voxel_build(ms, v + 1, 10.0, enc, mass=0.001))
    return Case(els, ms, mh.materials.WORKLOAD_ROWS)


def mirror(case, solver, n_steps, dt, env_kw=None, t0=0.0, stress=False, f64=False,
           nthreads=4):
    """Run the oracle step on the case's initial state."""
    pos, vel = case.pos0, case.vel0
    if solver == mh.MH_SOLVER_XPBD_REF:
        st = np.zeros(case.n_springs(), np.float64 if f64 else np.float32) if stress else None
        fn = oracle.xpbd_f64 if f64 else oracle.xpbd_f32
        p, v, s = fn(case.pop, case.mt, dt, t0, n_steps, pos, vel, st, nthreads=nthreads)
        return p.reshape(-1, 3), v.reshape(-1, 3), s
    env = oracle.make_env(dt, **(env_kw or {}))
    fn = oracle.force_f64 if f64 else oracle.force_f32
    p, v = fn(case.pop, case.mt, env, t0, n_steps, pos, vel, nthreads=nthreads)
    return p.reshape(-1, 3), v.reshape(-1, 3), None


def gpu_sim(case, solver, dt, env_kw=None, **cfg_kw):
    kw = dict(env_kw or {})
    cfg = mh.SimConfig(solver=solver, time_step=dt, **kw, **cfg_kw)
    sim = mh.Simulator(case.ms, cfg)
    sim.SetElements(case.elements)
    return sim


def rel_err(x32, x64, mass_offset):
    """Per-robot max |x32 - x64| / max |x64|."""
    out = []
    for r in range(len(mass_offset) - 1):
        a = np.asarray(x32[mass_offset[r]:mass_offset[r + 1]], np.float64)
        b = np.asarray(x64[mass_offset[r]:mass_offset[r + 1]], np.float64)
        if len(b):
            out.append(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
    return np.array(out)


def adam_first_step_f32(p, g, lr, b1=0.9, b2=0.999, eps=1e-8):
    """The device's first Adam step (t = 1, zero moments) restated in float32,
    one rounding per operation in the kernel's order: m = b1 0 + (1 - b1) g,
    v = b2 0 + ((1 - b2) g) g, p - (lr (m / bc1)) / (sqrt(v / bc2) + eps)
    with bc = 1 - b^1."""
    f = np.float32
    p, g = np.asarray(p, f), np.asarray(g, f)
    b1, b2, eps, lr = f(b1), f(b2), f(eps), f(lr)
    bc1, bc2 = f(1) - b1, f(1) - b2
    m = b1 * f(0) + (f(1) - b1) * g
    v = b2 * f(0) + ((f(1) - b2) * g) * g
    return p - (lr * (m / bc1)) / (np.sqrt(v / bc2) + eps)
