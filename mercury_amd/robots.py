"""Robot elements and seeded populations.

An Element mirrors the reference's Element (common/simulator/element.h:9-30)
as arrays: masses (pos, vel, mass, material id) and springs (pairs, mean
length, material id).  Voxel robots are built by libmercury_hip's
mh_voxel_build (VoxelRobot::Build semantics).
"""
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, f32p, load, ptr, u8p, u16p, u32p
import ctypes


@dataclass
class Element:
    pos: np.ndarray            # (n, 3) float32
    mass_material: np.ndarray  # (n,) uint8 composite id (0 = air)
    pairs: np.ndarray          # (s, 2) uint16 robot-local
    mean_length: np.ndarray    # (s,) float32
    spring_material: np.ndarray  # (s,) uint8 composite id
    vel: np.ndarray = None     # (n, 3) float32, zeros if None
    mass: np.ndarray = None    # (n,) float32, ones if None
    rest_length: np.ndarray = None
    stress: np.ndarray = None

    @property
    def n_masses(self):
        return len(self.pos)

    @property
    def n_springs(self):
        return len(self.pairs)

    def copy(self):
        c = Element(self.pos.copy(), self.mass_material.copy(), self.pairs.copy(),
                    self.mean_length.copy(), self.spring_material.copy())
        c.vel = None if self.vel is None else self.vel.copy()
        c.mass = None if self.mass is None else self.mass.copy()
        c.rest_length = None if self.rest_length is None else self.rest_length.copy()
        return c


def voxel_build(material_set, n, resolution, voxel_encoding, mass=1.0):
    """VoxelRobot::Build for an n^3 lattice; voxel_encoding has n^3 entries."""
    L = load()
    nv = n ** 3
    enc = np.ascontiguousarray(voxel_encoding, np.uint32)
    if enc.shape != (nv,):
        raise ValueError("voxel_encoding must have n**3 entries")
    maxs = L.mh_voxel_max_springs(n, n, n)
    pos = np.zeros((nv, 3), np.float32)
    mid = np.zeros(nv, np.uint8)
    pairs = np.zeros((maxs, 2), np.uint16)
    ln = np.zeros(maxs, np.float32)
    sid = np.zeros(maxs, np.uint8)
    senc = np.zeros(maxs, np.uint32)
    stripped = np.zeros(nv, np.uint32)
    ns = ctypes.c_int()
    check(L.mh_voxel_build(material_set.handle, n, n, n, float(resolution), ptr(enc, u32p),
                           ptr(pos, f32p), ptr(mid, u8p), ptr(pairs, u16p), ptr(ln, f32p),
                           ptr(sid, u8p), ptr(senc, u32p), maxs, ctypes.byref(ns),
                           ptr(stripped, u32p)))
    s = ns.value
    e = Element(pos, mid, pairs[:s].copy(), ln[:s].copy(), sid[:s].copy())
    e.rest_length = ln[:s].copy()
    e.mass = np.full(nv, mass, np.float32)
    e.spring_encoding = senc[:s].copy()
    e.stripped_encoding = stripped
    return e


def rest_lengths(pos, pairs):
    """SoftBody::Reset mean lengths from the current proto positions."""
    pos = np.ascontiguousarray(pos, np.float32)
    pairs = np.ascontiguousarray(pairs, np.uint16)
    out = np.zeros(len(pairs), np.float32)
    check(load().mh_rest_lengths(ptr(pos, f32p), ptr(pairs, u16p), len(pairs), ptr(out, f32p)))
    return out


def baseline(pos, mass_material):
    pos = np.ascontiguousarray(pos, np.float32)
    mm = np.ascontiguousarray(mass_material, np.uint8)
    com = np.zeros(3, np.float32)
    ln = ctypes.c_float()
    check(load().mh_baseline(ptr(pos, f32p), ptr(mm, u8p), len(mm), ptr(com, f32p),
                             ctypes.byref(ln)))
    return com, np.float32(ln.value)


def last_layer_air(n, enc):
    idx = np.arange(n ** 3)
    x, y, z = idx % n, (idx // n) % n, idx // (n * n)
    enc = np.array(enc, dtype=np.uint32, copy=True)
    enc[(x == n - 1) | (y == n - 1) | (z == n - 1)] = 1
    return enc


def random_voxel_encodings(rng, n_voxels_per_axis, n_materials, p_empty):
    """Per-voxel material draw for an (n_voxels_per_axis)^3 voxel grid on an
    (n+1)^3 lattice: 'empty' (air) with probability p_empty, otherwise one of
    the n_materials non-air bases uniformly.  Lattice points on the last layer
    carry no voxel (air), as VoxelRobot::Initialize (VoxelRobot.cpp:375-378)."""
    n = n_voxels_per_axis + 1
    mats = rng.integers(1, n_materials + 1, n ** 3)
    if p_empty > 0:
        mats[rng.random(n ** 3) < p_empty] = 0
    return last_layer_air(n, (np.uint32(1) << mats.astype(np.uint32)).astype(np.uint32))


def voxel_population(material_set, n_robots, voxels_per_axis=8, p_empty=0.2, seed=0,
                     spacing=0.1, mass=0.001, n_materials=4):
    """Seeded voxel population: n_robots robots on a (voxels_per_axis+1)^3
    lattice with the given spacing (m) and point mass (kg)."""
    rng = np.random.default_rng(seed)
    n = voxels_per_axis + 1
    res = 1.0 / spacing
    out = []
    for _ in range(n_robots):
        enc = random_voxel_encodings(rng, voxels_per_axis, n_materials, p_empty)
        out.append(voxel_build(material_set, n, res, enc, mass=mass))
    return out
