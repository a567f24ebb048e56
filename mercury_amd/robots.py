"""Robot elements and seeded populations.

An Element mirrors the reference's Element (common/simulator/element.h:9-30)
as arrays: masses (pos, vel, mass, material id) and springs (pairs, mean
length, material id).  Voxel robots are built by libmercury_hip's
mh_voxel_build (VoxelRobot::Build semantics).
"""
from dataclasses import dataclass, field

import numpy as np

import ctypes

from ._lib import check, f32p, i32p, load, ptr, u8p, u16p, u32p


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


def voxel_template(n, resolution):
    """What every n^3 voxel robot shares (mh_voxel_template): mass positions
    (n^3, 3), springs (S, 2) in DFS order, their lengths, and per spring the
    4 lattice points (0xffff = none) whose voxels decide its material."""
    L = load()
    nv = n ** 3
    maxs = L.mh_voxel_max_springs(n, n, n)
    pos = np.zeros((nv, 3), np.float32)
    pairs = np.zeros((maxs, 2), np.uint16)
    ln = np.zeros(maxs, np.float32)
    vox = np.zeros((maxs, 4), np.uint16)
    ns = ctypes.c_int()
    check(L.mh_voxel_template(n, n, n, float(resolution), ptr(pos, f32p), ptr(pairs, u16p), ptr(ln, f32p),
                              ptr(vox, u16p), maxs, ctypes.byref(ns)))
    s = ns.value
    return pos, pairs[:s].copy(), ln[:s].copy(), vox[:s].copy()


def voxel_build_batch(material_set, n, resolution, voxel_encodings, mass=1.0, threads=0):
    """VoxelRobot::BatchBuild: many n^3 lattices built by native threads into
    one concatenated population (dict of arrays, as Simulator.set_population
    takes them)."""
    L = load()
    nv = n ** 3
    enc = np.ascontiguousarray(voxel_encodings, np.uint32).reshape(-1, nv)
    R = len(enc)
    cap = max(1, R * L.mh_voxel_max_springs(n, n, n))
    pos = np.zeros((R * nv, 3), np.float32)
    mid = np.zeros(R * nv, np.uint8)
    so = np.zeros(R + 1, np.uint32)
    pairs = np.zeros((cap, 2), np.uint16)
    ln = np.zeros(cap, np.float32)
    sid = np.zeros(cap, np.uint8)
    check(L.mh_voxel_build_batch(material_set.handle, R, n, n, n, float(resolution), ptr(enc, u32p),
                                 int(threads), ptr(pos, f32p), ptr(mid, u8p), ptr(so, u32p),
                                 ptr(pairs, u16p), ptr(ln, f32p), ptr(sid, u8p), cap))
    S = int(so[-1])
    return dict(mass_offset=(np.arange(R + 1, dtype=np.uint32) * nv), spring_offset=so, pos=pos,
                vel=np.zeros_like(pos), mass=np.full(R * nv, mass, np.float32), mass_material=mid,
                pairs=pairs[:S], mean_length=ln[:S], spring_material=sid[:S])


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


def nn_materials(weights, sizes, pos):
    """Material index (0-based output) per mass from a bias-free ReLU MLP on
    its position (NNRobot::forward, NNRobot.h:41-92)."""
    pos = np.ascontiguousarray(pos, np.float32)
    w = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float32).reshape(-1) for x in weights]))
    sz = np.ascontiguousarray(sizes, np.int32)
    out = np.zeros(len(pos), np.uint8)
    check(load().mh_nn_materials(len(sizes) - 1, ptr(sz, i32p), ptr(w, f32p), ptr(pos, f32p),
                                 len(pos), ptr(out, u8p)))
    return out


def knn_batch(mass_offset, pos, K, device=0):
    """KNN rows (neighbour index, distance) for every mass, on the GPU."""
    mo = np.ascontiguousarray(mass_offset, np.uint32)
    pos = np.ascontiguousarray(pos, np.float32)
    M = int(mo[-1])
    nbr = np.zeros((M, K), np.uint16)
    dist = np.zeros((M, K), np.float32)
    check(load().mh_knn_batch(device, len(mo) - 1, ptr(mo, u32p), ptr(pos, f32p), K,
                              ptr(nbr, u16p), ptr(dist, f32p)))
    return nbr, dist


def knn_springs(material_set, nbr, dist, mass_encoding, min_length=1e-3):
    n, K = nbr.shape
    enc = np.ascontiguousarray(mass_encoding, np.uint32)
    maxs = n * K
    pairs = np.zeros((maxs, 2), np.uint16)
    ln = np.zeros(maxs, np.float32)
    mid = np.zeros(maxs, np.uint8)
    ns = ctypes.c_int()
    check(load().mh_knn_springs(material_set.handle, n, K, ptr(np.ascontiguousarray(nbr), u16p),
                                ptr(np.ascontiguousarray(dist), f32p), ptr(enc, u32p),
                                float(min_length), ptr(pairs, u16p), ptr(ln, f32p), ptr(mid, u8p),
                                maxs, ctypes.byref(ns)))
    s = ns.value
    return pairs[:s].copy(), ln[:s].copy(), mid[:s].copy()


def shift_xy(pos):
    """ShiftX then ShiftY (VoxelRobot.cpp:30-56, NNRobot.cpp:29-55): max x
    to 0, then min y to 0, by float additions."""
    pos = np.array(pos, np.float32, copy=True)
    if len(pos):
        pos[:, 0] = (pos[:, 0] + np.float32(-pos[:, 0].max())).astype(np.float32)
        pos[:, 1] = (pos[:, 1] + np.float32(-pos[:, 1].min())).astype(np.float32)
    return pos


def nn_population(material_set, n_robots, mass_range=(200, 2000), cube=0.5, K=25,
                  hidden=(25, 25), n_materials=4, seed=0, mass=0.01, device=0):
    """Seeded NN-robot population: per robot a mass count uniform in
    mass_range, positions uniform in a cube of side `cube` (m), materials
    from a random-init bias-free MLP 3 -> hidden -> n_materials on the
    position, springs to the K nearest neighbours (computed on the GPU)."""
    rng = np.random.default_rng(seed)
    sizes = [3, *hidden, n_materials]
    robots = []
    counts = rng.integers(mass_range[0], mass_range[1] + 1, n_robots)
    all_pos, mats = [], []
    for r in range(n_robots):
        m = int(counts[r])
        pos = (rng.random((m, 3)) * cube).astype(np.float32)
        weights = [rng.uniform(-1, 1, (o, i)).astype(np.float32)
                   for i, o in zip(sizes[:-1], sizes[1:])]
        # the network sees positions relative to the cube centre, like the
        # reference's origin-centred input cloud (NNRobot.h:115-135)
        idx = nn_materials(weights, sizes, (pos - np.float32(cube / 2)).astype(np.float32))
        all_pos.append(pos)
        mats.append(idx)
    mo = np.zeros(n_robots + 1, np.uint32)
    mo[1:] = np.cumsum(counts)
    nbr, dist = knn_batch(mo, np.concatenate(all_pos), K, device)
    for r in range(n_robots):
        a, b = int(mo[r]), int(mo[r + 1])
        base_id = mats[r].astype(np.uint32) + 1          # base material ids 1..n
        enc = (np.uint32(1) << base_id).astype(np.uint32)
        pairs, ln, sid = knn_springs(material_set, nbr[a:b], dist[a:b], enc)
        e = Element(shift_xy(all_pos[r]), base_id.astype(np.uint8), pairs, ln, sid)
        e.rest_length = ln.copy()
        e.mass = np.full(b - a, mass, np.float32)
        e.mass_encoding = enc
        robots.append(e)
    return robots
