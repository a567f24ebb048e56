"""Devo: stress-sorted spring replacement between simulation cycles.

Follows Simulator::Devo (common/simulator/Simulator.cu:640-668) and its
kernel replaceSprings (common/simulator/devo_kernel.cu:81-165), run on the
host between two uploads of the population:

  per robot, the springs are ranked by accumulated stress, largest first
  (a stable descending sort, as cub's segmented radix sort pairs,
  Simulator.cu:597-628); the top `replaced` springs get a new pair of
  distinct random masses (randomIntegerPairKernel, devo_kernel.cu:55-72),
  the material decode(enc_a | enc_b) of the two masses' encodings, and the
  rest length |x_a - x_b| / (1 + dL0 sin(omega t + phi)) from the masses'
  current positions and the new material at the current time.

Deliberate differences, each a defect of the reference:
  - the reference indexes element e's sorted springs from rank e * replaced
    (sortedSpringId = e * springsPerElement + i with i the global counter,
    devo_kernel.cu:113), so later robots replace lower-stress springs; here
    every robot replaces its own top `replaced`;
  - the reference updates the spring's material encoding but not the
    material id the step kernel reads (devo_kernel.cu:161-162 vs
    sim_kernel.cu:219), so the old material acts until the next upload;
    here the new material acts at once;
  - the random pairs come from numpy's PCG64 seeded by (seed, cycle, robot)
    instead of curand, so runs are reproducible on any device.
The stress ranking uses the stress accumulated since the evaluation's
first upload: the evaluator carries it over each re-upload
(Simulator.set_stress), so a replaced spring keeps its slot's stress, as in
the reference, whose stress buffer only SetElements clears
(Simulator.cu:355).
Lengths |x_a - x_b| are evaluated as SoftBody::Reset evaluates them
(mh_rest_lengths, Eigen's norm order); the reference's devo kernel uses
CUDA's norm3df and sinf, which are not reproduced (host float32 sin here).
"""
import numpy as np

from .robots import rest_lengths

_F = np.float32


def random_pairs(rng, n, n_masses):
    """n pairs of distinct mass indices in [0, n_masses)."""
    a = rng.integers(0, n_masses, n)
    b = rng.integers(0, n_masses, n)
    same = a == b
    while same.any():
        b[same] = rng.integers(0, n_masses, int(same.sum()))
        same = a == b
    return np.stack([a, b], 1).astype(np.uint16)


class MaterialLookup:
    """Composite ids and (k, dL0, omega, phi) of a material set, cached."""

    def __init__(self, material_set):
        self.ms = material_set
        self._enc_of_id = {}
        self._id_of_enc = {}
        self._row = {}

    def encoding(self, cid):
        cid = int(cid)
        if cid not in self._enc_of_id:
            self._enc_of_id[cid] = self.ms.by_id(cid)[1]
        return self._enc_of_id[cid]

    def decode(self, enc):
        enc = int(enc)
        if enc not in self._id_of_enc:
            self._id_of_enc[enc] = self.ms.decode(enc)[0]
        return self._id_of_enc[enc]

    def row(self, cid):
        cid = int(cid)
        if cid not in self._row:
            self._row[cid] = self.ms.by_id(cid)[0]
        return self._row[cid]


def devo_robot(el, stress, pos, t, replaced, rng, lookup):
    """One robot's replacement.  el: the robot (its springs are replaced in a
    copy); stress: per spring; pos: current positions (n, 3); t: simulated
    time.  Returns (new element, indices of the replaced springs)."""
    out = el.copy()
    ns, nm = el.n_springs, el.n_masses
    r = min(int(replaced), ns)
    if r <= 0 or nm < 2:
        return out, np.zeros(0, np.int64)
    order = np.argsort(-np.asarray(stress, np.float64), kind="stable")[:r]
    pairs = random_pairs(rng, r, nm)
    menc = np.array([lookup.encoding(c) for c in el.mass_material], np.uint32)
    enc = menc[pairs[:, 0]] | menc[pairs[:, 1]]
    sid = np.array([lookup.decode(e) for e in enc], np.uint8)
    d = rest_lengths(pos, pairs)
    rows = np.array([lookup.row(c) for c in sid], _F).reshape(-1, 4)
    rel = rows[:, 1] * np.sin(rows[:, 2] * _F(t) + rows[:, 3]).astype(_F)
    lbar = (d / (_F(1) + rel)).astype(_F)
    out.pairs = np.array(el.pairs, np.uint16, copy=True)
    out.pairs[order] = pairs
    out.spring_material = np.array(el.spring_material, np.uint8, copy=True)
    out.spring_material[order] = sid
    out.mean_length = np.array(el.mean_length, _F, copy=True)
    out.mean_length[order] = lbar
    rest = el.mean_length if el.rest_length is None else el.rest_length
    out.rest_length = np.array(rest, _F, copy=True)
    out.rest_length[order] = lbar
    return out, order


def rest_from_proto(el, proto_pos):
    """SoftBody::Reset (SoftBody.cpp:68-80): every spring's mean and rest
    length become the distance of its masses' proto positions."""
    out = el.copy()
    p = np.array(proto_pos, _F)
    L = rest_lengths(p, el.pairs) if el.n_springs else np.zeros(0, _F)
    out.pos = p
    out.vel = None
    out.mean_length = L
    out.rest_length = L.copy()
    return out
