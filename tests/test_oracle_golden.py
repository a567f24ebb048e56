"""Pin the oracle against golden vectors written by the compiled reference.

Fixtures come from tests/golden/make_golden.py, which runs the reference's own
material.h / VoxelRobot.cpp / SoftBody.cpp (built by oracle/ref/Makefile).
"""
import glob
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def ref_tables():
    base = oracle.reference_base()
    by_id, by_enc, by_enc_struct = oracle.composite_tables(base)
    return base, by_id, by_enc, by_enc_struct


def test_material_tables_match_reference(ref_tables):
    g = np.load(os.path.join(GOLDEN, "materials_reference.npz"))
    base, by_id, by_enc, _ = ref_tables
    gid = g["by_id"]
    assert len(gid) == 254 == len(by_id["k"])
    assert np.array_equal(by_id["id"], gid["id"])
    for f in ("k", "dL0", "omega", "phi"):
        assert np.array_equal(bits(by_id[f]), bits(gid[f])), f
    assert np.array_equal(by_id["encoding"], gid["enc"])
    gb = g["base"]
    for i, m in enumerate(base):
        assert (m.id, m.encoding) == (int(gb["id"][i]), int(gb["enc"][i]))
        for f in ("k", "dL0", "omega", "phi"):
            assert bits(getattr(m, f)) == bits(gb[f][i]), (i, f)


def test_decode_matches_reference(ref_tables):
    g = np.load(os.path.join(GOLDEN, "materials_reference.npz"))
    _, _, by_enc, _ = ref_tables
    for e, idx, mid in zip(g["dec_input"], g["dec_idx"], g["dec_id"]):
        got = oracle.encoded_idx(int(e), 254)
        assert got == idx, hex(int(e))
        assert by_enc["id"][got] == mid
    # TestMatEncoding known answers (all four comparisons were equal)
    assert list(g["known_answers"]) == [0, 0, 0, 0]


def test_known_answers_mat_encoding(ref_tables):
    """common/tests/test_simulator.cpp:65-81 restated on the oracle tables."""
    base, by_id, by_enc, _ = ref_tables

    def dec(e):
        i = oracle.encoded_idx(e, 254)
        return tuple(by_enc[f][i] for f in ("k", "dL0", "omega", "phi"))

    def mat(m):
        return (m.k, m.dL0, m.omega, m.phi)

    air, add0, bone = base[0], base[1], base[22]
    assert dec(air.encoding) == mat(air)
    assert dec(bone.encoding) == mat(bone)
    assert dec(add0.encoding) == mat(add0)
    assert dec(air.encoding | bone.encoding) == mat(air)


VOXEL_FILES = sorted(glob.glob(os.path.join(GOLDEN, "voxel_*.npz")))


@pytest.mark.parametrize("path", VOXEL_FILES, ids=[os.path.basename(p) for p in VOXEL_FILES])
def test_voxel_build_matches_reference(path, ref_tables):
    g = np.load(path)
    _, _, _, by_enc_struct = ref_tables
    n = int(g["n"])
    out = oracle.voxel_build(n, float(g["res"]), g["voxel_enc"], by_enc_struct)
    assert np.array_equal(bits(out["mass_pos"]), bits(g["mass_pos"]))
    assert np.array_equal(out["mass_id"], g["mass_id"])
    assert np.array_equal(out["pairs"], g["pairs"])
    assert np.array_equal(bits(out["length"]), bits(g["rest_len"]))
    assert np.array_equal(bits(out["length"]), bits(g["mean_len"]))
    assert np.array_equal(out["spring_id"], g["spring_id"])
    assert np.array_equal(out["spring_enc"], g["spring_enc"])


@pytest.mark.parametrize("path", VOXEL_FILES, ids=[os.path.basename(p) for p in VOXEL_FILES])
def test_fitness_matches_reference(path):
    g = np.load(path)
    air = (g["mass_id"] == 0).astype(np.uint8)
    pos0 = g["mass_pos"]
    base = oracle.com(pos0, air)
    assert np.array_equal(bits(base), bits(g["base_com"]))
    assert bits(oracle.length(pos0, air)) == bits(g["length"])
    for k in range(len(g["pos_sets"])):
        c = oracle.com(g["pos_sets"][k], air)
        assert np.array_equal(bits(c), bits(g["com"][k]))
        f = oracle.fitness(c[0], base[0], oracle.length(pos0, air))
        assert bits(f) == bits(g["fitness"][k])


@pytest.mark.parametrize("path", VOXEL_FILES, ids=[os.path.basename(p) for p in VOXEL_FILES])
def test_reset_mean_length_formula(path):
    """SoftBody::Reset recomputes mean lengths from the shifted positions with
    Eigen's Vector3f::norm: sqrtf(dx*dx + (dy*dy + dz*dz))."""
    g = np.load(path)
    p = g["mass_pos"].astype(np.float32)
    a, b = g["pairs"][:, 0], g["pairs"][:, 1]
    d = p[a] - p[b]
    yz = d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    L = np.sqrt(d[:, 0] * d[:, 0] + yz)
    assert np.array_equal(bits(L), bits(g["reset_mean_len"]))


def test_spring_count_invariants(ref_tables):
    _, _, _, by_enc_struct = ref_tables
    for n, want in ((5, 1036), (9, 7448), (12, 18788)):
        out = oracle.voxel_build(n, 1.0, np.full(n ** 3, 1 << 22, np.uint32), by_enc_struct)
        assert len(out["pairs"]) == want
        d = out["mass_pos"][out["pairs"][:, 1]] - out["mass_pos"][out["pairs"][:, 0]]
        l1 = np.abs(np.rint(d)).sum(1)
        assert (l1 == 1).sum() == 3 * (n - 1) * n * n
        assert (l1 == 2).sum() == 6 * (n - 1) ** 2 * n
        assert (l1 == 3).sum() == 4 * (n - 1) ** 3
