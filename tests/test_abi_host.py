"""C-ABI host functions (no GPU): exports, materials, voxel topology, baseline.

The product library is checked against the golden vectors of the compiled
reference (tests/golden) and against the oracle.
"""
import glob
import os
import re

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, REPO

mh = pytest.importorskip("mercury_amd")


def bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def header_functions():
    names = []
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**(mh_[a-z0-9_]+)\s*\(",
                            txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    import ctypes
    lib = mh.load()
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_material_tables():
    g = np.load(os.path.join(GOLDEN, "materials_reference.npz"))
    ms = mh.MaterialSet.reference()
    assert ms.count() == 254
    gid = g["by_id"]
    for i in range(254):
        kdwp, enc = ms.by_id(i)
        assert np.array_equal(bits(kdwp), bits([gid["k"][i], gid["dL0"][i], gid["omega"][i],
                                                gid["phi"][i]])), i
        assert enc == gid["enc"][i]
    for e, idx, mid in zip(g["dec_input"], g["dec_idx"], g["dec_id"]):
        got_id, got_idx = ms.decode(int(e))
        assert (got_id, got_idx) == (mid, idx), hex(int(e))


def test_custom_material_set_matches_oracle():
    rows = mh.materials.WORKLOAD_ROWS
    ms = mh.workload_materials()
    assert ms.count() == 11
    by_id, by_enc, _ = oracle.composite_tables(oracle.make_base(rows))
    for i in range(11):
        kdwp, enc = ms.by_id(i)
        want = [by_id[f][i] for f in ("k", "dL0", "omega", "phi")]
        assert np.array_equal(bits(kdwp), bits(want)), i
        assert enc == by_id["encoding"][i]
    for e in range(1 << 5):
        assert ms.decode(e)[1] == oracle.encoded_idx(e, 11)


VOXEL_FILES = sorted(glob.glob(os.path.join(GOLDEN, "voxel_*.npz")))


@pytest.mark.parametrize("path", VOXEL_FILES, ids=[os.path.basename(p) for p in VOXEL_FILES])
def test_voxel_build_matches_reference(path):
    g = np.load(path)
    ms = mh.MaterialSet.reference()
    e = mh.voxel_build(ms, int(g["n"]), float(g["res"]), g["voxel_enc"])
    assert np.array_equal(bits(e.pos), bits(g["mass_pos"]))
    assert np.array_equal(e.mass_material, g["mass_id"])
    assert np.array_equal(e.pairs, g["pairs"])
    assert np.array_equal(bits(e.mean_length), bits(g["mean_len"]))
    assert np.array_equal(e.spring_material, g["spring_id"])
    assert np.array_equal(e.spring_encoding, g["spring_enc"])
    reset = mh.robots.rest_lengths(e.pos, e.pairs)
    assert np.array_equal(bits(reset), bits(g["reset_mean_len"]))
    com, ln = mh.robots.baseline(e.pos, e.mass_material)
    assert np.array_equal(bits(com), bits(g["base_com"]))
    assert bits(ln) == bits(g["length"])


def test_voxel_build_workload_set_matches_oracle():
    ms = mh.workload_materials()
    _, _, by_enc_struct = oracle.composite_tables(oracle.make_base(mh.materials.WORKLOAD_ROWS))
    rng = np.random.default_rng(7)
    for trial in range(6):
        n = int(rng.integers(3, 10))
        enc = mh.robots.random_voxel_encodings(rng, n - 1, 4, p_empty=float(rng.uniform(0, 0.7)))
        e = mh.voxel_build(ms, n, 10.0, enc)
        o = oracle.voxel_build(n, 10.0, enc, by_enc_struct)
        assert np.array_equal(bits(e.pos), bits(o["mass_pos"]))
        assert np.array_equal(e.mass_material, o["mass_id"])
        assert np.array_equal(e.pairs, o["pairs"])
        assert np.array_equal(bits(e.mean_length), bits(o["length"]))
        assert np.array_equal(e.spring_material, o["spring_id"])
        assert np.array_equal(e.stripped_encoding, o["stripped_enc"])


def test_voxel_build_rejects_bad_input():
    ms = mh.workload_materials()
    with pytest.raises(mh.MhError):
        mh.voxel_build(ms, 3, 10.0, np.full(27, 1 << 9, np.uint32))  # bit outside the set
    with pytest.raises(mh.MhError):
        mh.voxel_build(ms, 3, 0.0, np.full(27, 2, np.uint32))       # resolution 0


def test_voxel_build_batch_matches_single_builds():
    """VoxelRobot::BatchBuild: the threaded batch gives the per-robot builds,
    concatenated, for any thread count."""
    ms = mh.workload_materials()
    rng = np.random.default_rng(5)
    encs = np.stack([mh.robots.random_voxel_encodings(rng, 4, 4, 0.3) for _ in range(37)])
    encs[3] = 1  # an all-air robot: no springs
    ref = mh.pack([mh.robots.voxel_build(ms, 5, 10.0, e, mass=0.002) for e in encs])
    for threads in (1, 3, 0):
        got = mh.robots.voxel_build_batch(ms, 5, 10.0, encs, mass=0.002, threads=threads)
        for k in ref:
            assert np.array_equal(ref[k], got[k]), (threads, k)


@pytest.mark.parametrize("reference_set", [False, True])
def test_voxel_template_gives_every_build(reference_set):
    """The shape template (positions, DFS springs, lengths) is what
    mh_voxel_build gives any robot of the lattice, and decoding the OR of
    the template's voxels' non-air bits reproduces every spring material
    after Strip, for base and two-material voxel encodings."""
    from mercury_amd.robots import voxel_template
    ms = mh.MaterialSet.reference() if reference_set else mh.workload_materials()
    nb = 22 if reference_set else 4
    rng = np.random.default_rng(8)
    for n in (2, 4, 6):
        pos, pairs, ln, vox = voxel_template(n, 10.0)
        for trial in range(6):
            mats = rng.integers(1, nb + 1, (n ** 3, 2))
            enc = (np.uint32(1) << mats[:, 0].astype(np.uint32))
            if trial % 2:
                enc |= np.uint32(1) << mats[:, 1].astype(np.uint32)
            enc[rng.random(n ** 3) < (0.2 if trial < 4 else 0.6)] = 1
            enc = mh.robots.last_layer_air(n, enc)
            el = mh.voxel_build(ms, n, 10.0, enc)
            assert np.array_equal(el.pos.view(np.uint32), pos.view(np.uint32))
            assert np.array_equal(el.pairs, pairs) and np.array_equal(el.mean_length.view(np.uint32), ln.view(np.uint32))
            # non-air bits of each voxel's decoded material after Strip
            nab = np.array([0 if e == 1 else int(ms.by_id(int(ms.decode(int(e))[0]))[1]) for e in el.stripped_encoding],
                           np.uint32)
            ored = np.zeros(len(pairs), np.uint32)
            for k in range(4):
                ok = vox[:, k] != 0xffff
                ored[ok] |= nab[vox[ok, k]]
            want = np.array([ms.decode(int(e))[0] for e in ored], np.uint8)
            assert np.array_equal(want, el.spring_material)
