"""Simulator: the reference's host API over libmercury_hip.

Method names and meaning follow common/simulator/Simulator.h:19-149:
Initialize, SetElements / SetElement, Simulate, Collect, Reset, getDeltaT,
setTimeStep, getTotalTime.  Errors raise MhError instead of exit().
Additions: fitness/com (device-side scoring), autotune, info, timing.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (Config, Population, SimInfo, check, f32p, load, ptr, u8p, u16p, u32p)
from .materials import MaterialSet
from .robots import Element


@dataclass
class SimConfig:
    """Config::Simulator (config.h:16-21) plus the force-solver environment."""
    solver: int = _lib.MH_SOLVER_XPBD_REF
    time_step: float = 1e-4
    gravity: float = 9.81
    ground_k: float = 1.0e5
    mu_s: float = 1.0
    mu_k: float = 0.8
    drag: float = 0.0
    damping: float = 0.99
    device: int = 0
    kernel: int = _lib.MH_KERNEL_AUTO
    block_size: int = 0
    robots_per_block: int = 0
    substeps_per_launch: int = 0
    use_graph: int = 1

    def to_c(self):
        c = Config()
        load().mh_config_default(ctypes.byref(c))
        c.solver, c.dt, c.gravity, c.ground_k = self.solver, self.time_step, self.gravity, self.ground_k
        c.mu_s, c.mu_k, c.drag, c.damping = self.mu_s, self.mu_k, self.drag, self.damping
        c.device, c.kernel, c.block_size = self.device, self.kernel, self.block_size
        c.robots_per_block, c.substeps_per_launch = self.robots_per_block, self.substeps_per_launch
        c.use_graph = self.use_graph
        return c


@dataclass
class ElementTracker:
    """Simulator.h:11-17: which slice of the packed population is one robot."""
    ID: int
    mass_begin: int
    mass_end: int
    spring_begin: int
    spring_end: int


def pack(elements):
    """Concatenate elements into the packed SoA population arrays."""
    nm = np.array([e.n_masses for e in elements], np.int64)
    ns = np.array([e.n_springs for e in elements], np.int64)
    mo = np.zeros(len(elements) + 1, np.uint32)
    so = np.zeros(len(elements) + 1, np.uint32)
    mo[1:] = np.cumsum(nm)
    so[1:] = np.cumsum(ns)
    cat = lambda xs, dt, shape: (np.concatenate(xs).astype(dt, copy=False) if xs
                                 else np.zeros(shape, dt))
    arrays = dict(
        mass_offset=mo, spring_offset=so,
        pos=np.ascontiguousarray(cat([e.pos for e in elements], np.float32, (0, 3))),
        vel=np.ascontiguousarray(cat([e.vel if e.vel is not None else np.zeros((e.n_masses, 3), np.float32)
                                      for e in elements], np.float32, (0, 3))),
        mass=np.ascontiguousarray(cat([e.mass if e.mass is not None else np.ones(e.n_masses, np.float32)
                                       for e in elements], np.float32, (0,))),
        mass_material=np.ascontiguousarray(cat([e.mass_material for e in elements], np.uint8, (0,))),
        pairs=np.ascontiguousarray(cat([e.pairs for e in elements], np.uint16, (0, 2))),
        mean_length=np.ascontiguousarray(cat([e.mean_length for e in elements], np.float32, (0,))),
        spring_material=np.ascontiguousarray(cat([e.spring_material for e in elements], np.uint8, (0,))),
    )
    return arrays


class Simulator:
    def __init__(self, materials=None, config=None):
        self._h = None
        self.materials = materials if materials is not None else MaterialSet.reference()
        self.Initialize(config or SimConfig())

    # ------------------------------------------------------------ lifecycle
    def Initialize(self, config=None):
        """Simulator::Initialize (Simulator.cu:103-109)."""
        self.close()
        self.config = config or SimConfig()
        h = ctypes.c_void_p()
        check(load().mh_sim_create(ctypes.byref(self.config.to_c()), self.materials.handle,
                                   ctypes.byref(h)))
        self._h = h
        self._elements = []
        self._n_robots = 0
        self._arrays = None

    def close(self):
        if self._h is not None and _lib._lib is not None:
            _lib._lib.mh_sim_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    # ------------------------------------------------------------- elements
    def SetElement(self, element):
        return self.SetElements([element])[0]

    def SetElements(self, elements):
        """Simulator::SetElements (Simulator.cu:206-367): copy to the device."""
        self.set_population(pack(elements))
        self._elements = list(elements)
        mo, so = self._arrays["mass_offset"], self._arrays["spring_offset"]
        return [ElementTracker(i, int(mo[i]), int(mo[i + 1]), int(so[i]), int(so[i + 1]))
                for i in range(len(elements))]

    def set_population(self, a):
        """Upload a packed population (the dict pack() or
        robots.voxel_build_batch() returns) without Element objects."""
        n = len(a["mass_offset"]) - 1
        a = {k: np.ascontiguousarray(v) for k, v in a.items()}
        pop = Population(n, ptr(a["mass_offset"], u32p), ptr(a["spring_offset"], u32p),
                         ptr(a["pos"], f32p), ptr(a["vel"], f32p), ptr(a["mass"], f32p),
                         ptr(a["mass_material"], u8p), ptr(a["pairs"], u16p),
                         ptr(a["mean_length"], f32p), ptr(a["spring_material"], u8p))
        check(load().mh_sim_set_population(self._h, ctypes.byref(pop)))
        self._arrays = a
        self._n_robots = n
        self._elements = []

    def set_voxel_population(self, voxel_encodings, n, resolution, mass=1.0):
        """Voxel robots of one n^3 lattice built on the device from their
        encodings (n_robots x n^3; mh_sim_set_voxel_population): the same
        results as voxel_build_batch followed by set_population."""
        enc = np.ascontiguousarray(voxel_encodings, np.uint32).reshape(-1, n ** 3)
        check(load().mh_sim_set_voxel_population(self._h, len(enc), n, n, n, float(resolution), float(mass),
                                                 ptr(enc, u32p)))
        self._arrays = None
        self._n_robots = len(enc)
        self._elements = []

    def voxel_materials(self):
        """(mass ids, spring ids, stripped encodings) of a device-built voxel
        population, robot-major."""
        i = self.info()
        mid = np.zeros(i.n_masses, np.uint8)
        sid = np.zeros(i.n_springs, np.uint8)
        st = np.zeros(i.n_masses, np.uint32)
        check(load().mh_sim_voxel_materials(self._h, ptr(mid, u8p), ptr(sid, u8p), ptr(st, u32p)))
        return mid, sid, st

    # ------------------------------------------------------------- stepping
    def Simulate(self, sim_duration, trackStresses=False):
        """Simulator::Simulate (Simulator.cu:369-458): float time loop."""
        check(load().mh_sim_simulate(self._h, float(sim_duration), int(bool(trackStresses))))

    def Step(self, n_steps, trackStresses=False):
        check(load().mh_sim_step(self._h, int(n_steps), int(bool(trackStresses))))

    def synchronize(self):
        check(load().mh_sim_synchronize(self._h))

    def Collect(self, trackers=None):
        """Simulator::Collect (Simulator.cu:537-595): elements with the
        simulated positions/velocities and (air-zero) spring stresses."""
        pos, vel, st = self.collect_arrays()
        trackers = trackers if trackers is not None else range(len(self._elements))
        out = []
        for t in trackers:
            i = t.ID if isinstance(t, ElementTracker) else int(t)
            mo, so = self._arrays["mass_offset"], self._arrays["spring_offset"]
            e = self._elements[i].copy()
            e.pos = pos[mo[i]:mo[i + 1]].copy()
            e.vel = vel[mo[i]:mo[i + 1]].copy()
            e.stress = st[so[i]:so[i + 1]].copy()
            out.append(e)
        return out

    def collect_arrays(self):
        info = self.info()
        pos = np.zeros((info.n_masses, 3), np.float32)
        vel = np.zeros((info.n_masses, 3), np.float32)
        st = np.zeros(info.n_springs, np.float32)
        check(load().mh_sim_collect(self._h, ptr(pos, f32p), ptr(vel, f32p), ptr(st, f32p)))
        return pos, vel, st

    # ------------------------------------------------------------- scoring
    def fitness(self, mode=_lib.MH_FITNESS_REFERENCE):
        out = np.zeros(max(self._n_robots, 1), np.float32)
        check(load().mh_sim_fitness(self._h, int(mode), ptr(out, f32p)))
        return out[:self._n_robots]

    def com(self):
        out = np.zeros((max(self._n_robots, 1), 3), np.float32)
        check(load().mh_sim_com(self._h, ptr(out, f32p)))
        return out[:self._n_robots]

    # --------------------------------------------------------------- time
    def Reset(self):
        check(load().mh_sim_reset(self._h))

    def getDeltaT(self):
        return float(load().mh_sim_get_time_step(self._h))

    def setTimeStep(self, dt):
        check(load().mh_sim_set_time_step(self._h, float(dt)))

    def getTotalTime(self):
        return float(load().mh_sim_total_time(self._h))

    def setTotalTime(self, t):
        """Continue the actuation clock at t (no reference counterpart: the
        population upload restarts it at 0)."""
        check(load().mh_sim_set_total_time(self._h, float(t)))

    def set_stress(self, stress):
        """Set every spring's accumulated stress (original spring order; no
        reference counterpart: the population upload zeroes it)."""
        st = np.ascontiguousarray(stress, np.float32)
        if len(st) != self.info().n_springs:
            raise ValueError("one stress value per spring")
        check(load().mh_sim_set_stress(self._h, ptr(st, f32p)))

    # -------------------------------------------------------------- tuning
    def info(self):
        i = SimInfo()
        check(load().mh_sim_get_info(self._h, ctypes.byref(i)))
        return i

    def autotune(self, n_steps=200):
        check(load().mh_sim_autotune(self._h, int(n_steps)))
        return self.info()

    def set_timing(self, on=True):
        check(load().mh_sim_set_timing(self._h, int(bool(on))))

    def last_kernel_ms(self):
        ms = ctypes.c_float()
        n = ctypes.c_int()
        check(load().mh_sim_last_kernel_ms(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return float(ms.value), int(n.value)
