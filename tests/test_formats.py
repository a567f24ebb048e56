"""Robot text formats (SoftBody / VoxelRobot Encode and Decode): texts
written by the compiled reference's own encoders and decoders
(tests/golden/formats_reference.npz), known answers written out from the
reference's stream semantics, exact round trips, and byte equality of the
C++ view (include/mercury_hip_io.hpp) and the Python view
(mercury_amd.formats).  CPU only."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import mercury_amd as mh
from mercury_amd import evolve, formats
from mercury_amd.materials import MaterialSet
from mercury_amd.robots import Element
from conftest import REPO


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if np.asarray(a).dtype == np.float32 else np.uint8)


def test_softbody_known_answer():
    # two masses 0.1 m apart (1 g each), one spring; float32 values printed
    # with 9 significant digits; the empty face and cell lists leave their
    # keys on one line (SoftBody.cpp:147-172)
    el = Element(np.array([[0, 0, 0], [0.1, 0, 0]], np.float32), np.array([22, 21], np.uint8),
                 np.array([[0, 1]], np.uint16), np.array([0.1], np.float32), np.array([233], np.uint8),
                 mass=np.array([0.001, 0.001], np.float32))
    want = ("type=NNRobot\n"
            "masses=0,0,0,0,0.00100000005,22;1,0.100000001,0,0,0.00100000005,21\n"
            "springs=0,1,0.100000001,0.100000001,233\n"
            "faces=cells=boundary_count=0\n")
    assert formats.encode_softbody(el) == want
    d, ids, faces, cells, bc = formats.decode_softbody(want)
    assert list(ids) == [0, 1] and faces == [] and cells == [] and bc == 0
    assert np.array_equal(bits(d.pos), bits(el.pos)) and np.array_equal(d.pairs, el.pairs)
    # faces, cells and a boundary count
    s = formats.encode_softbody(el, faces=[(0, 1, 1)], cells=[(0, 1, 1, 0, 0.5, 6)], boundary_count=2)
    assert s.endswith("faces=0,1,1\ncells=0,1,1,0,0.5,6\nboundary_count=2\n")
    _, _, faces, cells, bc = formats.decode_softbody(s)
    assert faces == [(0, 1, 1)] and cells[0][:4] == (0, 1, 1, 0) and cells[0][5] == 6 and bc == 2


def test_softbody_round_trip_is_exact():
    ms = mh.workload_materials()
    rng = np.random.default_rng(3)
    g = evolve.random_genomes(rng, 1, 4)
    enc = evolve.genome_to_lattice(g, 4)[0]
    el = mh.voxel_build(ms, 5, 10.0, enc, mass=0.001)
    text = formats.encode_softbody(el)
    d, ids, _, _, _ = formats.decode_softbody(text)
    assert np.array_equal(ids, np.arange(el.n_masses))
    for a, b in ((d.pos, el.pos), (d.mean_length, el.mean_length), (d.rest_length, el.rest_length),
                 (d.mass, el.mass)):
        assert np.array_equal(bits(np.asarray(a, np.float32)), bits(np.asarray(b, np.float32)))
    assert np.array_equal(d.pairs, el.pairs)
    assert np.array_equal(d.mass_material, el.mass_material)
    assert np.array_equal(d.spring_material, el.spring_material)
    assert formats.encode_softbody(d) == text


def test_voxel_robot_known_answer():
    ms = MaterialSet.reference()
    table = formats.reference_voxel_materials(ms)
    omega = "%f" % float(np.float32(table[22][2]))
    mats = np.zeros(8, np.uint8)
    mats[0] = 22  # bone at the one material point of a 2^3 lattice
    text = formats.encode_voxel_robot(mats, (2, 2, 2), 10.0, table)
    lines = text.split("\n")
    assert lines[:5] == ["type=VoxelRobot", "0.200000", "0.200000", "0.200000", "10.000000"]
    # Voxel::to_string: id, indices, center, base, then Material::to_string
    assert lines[5] == ("v 0,0,0,0,0.050000,0.050000,0.050000,0.000000,0.000000,0.000000,"
                        f"100.000000, 0.000000, {omega}, 0.000000, 0.941176, 0.819608, 0.384314, 0.000000")
    assert lines[6].startswith("v 1,1,0,0,0.150000,0.050000,0.050000,0.100000,0.000000,0.000000,0.000000, ")
    assert len(lines) == 5 + 8 + 1 and lines[-1] == ""
    back, shape, res, _ = formats.decode_voxel_robot(text, table)
    assert shape == (2, 2, 2) and res == 10.0 and np.array_equal(back, mats)


def test_voxel_robot_round_trip_workload_genomes():
    ms = mh.workload_materials()
    table = [(0.0,) * 8] + [tuple(float(x) for x in r) + (0.0,) * 4 for r in ms.base_rows]
    rng = np.random.default_rng(5)
    enc = evolve.genome_to_lattice(evolve.random_genomes(rng, 3, 8), 8)
    for e in enc:
        mats = np.log2(e).astype(np.uint8)
        text = formats.encode_voxel_robot(mats, (9, 9, 9), 10.0, table)
        back, shape, _, _ = formats.decode_voxel_robot(text, table)
        assert shape == (9, 9, 9) and np.array_equal(back, mats)
        assert np.array_equal(formats.voxel_encodings(back), e)
        assert formats.encode_voxel_robot(back, shape, 10.0, table) == text


def test_decoders_reject_malformed_files():
    with pytest.raises(ValueError):
        formats.decode_softbody("type=VoxelRobot\n")
    with pytest.raises(ValueError):
        formats.decode_softbody("type=NNRobot\nmasses=0,1,2\n")
    with pytest.raises(ValueError):
        formats.decode_voxel_robot("type=VoxelRobot\n0.1\n0.1\n0.1\n10\n", [(0.0,) * 8])
    assert formats.solution_file_name(3, 0.25) == "solution_3_fitness_0.250000"


def test_cpp_and_python_views_write_the_same_bytes():
    lib_dir = os.path.join(REPO, "mercury_amd")
    n = 5
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "io_demo")
        subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I" + os.path.join(REPO, "include"),
                        os.path.join(REPO, "tests", "cpp", "io_demo.cpp"), "-o", exe, "-L" + lib_dir,
                        "-lmercury_hip", "-Wl,-rpath," + lib_dir], check=True, capture_output=True, text=True)
        a, b = os.path.join(tmp, "soft.txt"), os.path.join(tmp, "vox.txt")
        r = subprocess.run([exe, a, b, str(n)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        soft_cpp, vox_cpp = open(a).read(), open(b).read()
    ms = MaterialSet.reference()
    mats = np.zeros(n ** 3, np.uint8)
    for i in range(n ** 3):
        x, y, z = i % n, (i // n) % n, i // (n * n)
        mats[i] = 0 if n - 1 in (x, y, z) else (7 * i) % 23
    el = mh.voxel_build(ms, n, 10.0, formats.voxel_encodings(mats), mass=0.001)
    assert formats.encode_softbody(el) == soft_cpp
    size = np.float32(n) / np.float32(10.0)
    assert formats.encode_voxel_robot(mats, (n, n, n), 10.0, formats.reference_voxel_materials(ms),
                                      sizes=(size, size, size)) == vox_cpp


def _golden_formats():
    return np.load(os.path.join(REPO, "tests", "golden", "formats_reference.npz"))


def _text(z, key):
    return z[key].tobytes().decode()


@pytest.mark.parametrize("name", ["n4_base", "n5_random", "n9_p20_0"])
def test_encoders_write_the_reference_encoders_bytes(name):
    """Against texts written by the compiled reference (VoxelRobot::Encode,
    SoftBody::Encode of the built robot; tests/golden/make_golden.py):
    both encoders write the same bytes, and both decoders read them back."""
    z = _golden_formats()
    n, size, res = int(z[f"{name}_n"]), np.float32(z[f"{name}_size"]), float(z[f"{name}_res"])
    enc = z[f"{name}_enc"]
    soft, vox = _text(z, f"{name}_soft"), _text(z, f"{name}_voxel")
    ms = MaterialSet.reference()
    el = mh.voxel_build(ms, n, res, enc, mass=1.0)
    assert formats.encode_softbody(el) == soft
    mats = np.log2(enc).astype(np.uint8)  # one base material (or air) per lattice point
    table = formats.reference_voxel_materials(ms)
    assert formats.encode_voxel_robot(mats, (n, n, n), res, table, sizes=(size, size, size)) == vox
    d, ids, faces, cells, bc = formats.decode_softbody(soft)
    assert formats.encode_softbody(d) == soft and (faces, cells, bc) == ([], [], 0)
    back, shape, r, _ = formats.decode_voxel_robot(vox, table)
    assert shape == (n, n, n) and r == res and np.array_equal(back, mats)
    # the reference's own Decode -> Encode: VoxelRobot::Decode appends the
    # file's voxels to the default robot's instead of replacing them, so its
    # text ends with them; SoftBody::Decode cannot read a file whose face
    # list is empty (the key "cells=" follows "faces=" on the same line)
    rt = _text(z, f"{name}_voxel_roundtrip").split("\n")
    lines = vox.split("\n")
    assert rt[:5] == lines[:5] and rt[-(n ** 3) - 1:] == lines[5:]
    assert _text(z, f"{name}_soft_roundtrip").startswith("exception: stoi")


def test_reference_decoder_reads_faces_and_cells_we_write():
    """A soft body with faces, cells and a boundary count, written by this
    encoder, read by the reference's SoftBody::Decode and written again by
    its Encode: the same bytes."""
    z = _golden_formats()
    text = _text(z, "faces_input")
    el = Element(np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [0, 0, 0.1]], np.float32),
                 np.array([22, 21, 1, 0], np.uint8),
                 np.array([[0, 1], [1, 2], [2, 3], [0, 3]], np.uint16),
                 np.array([0.1, 0.141421363, 0.141421363, 0.1], np.float32),
                 np.array([233, 22, 5, 0], np.uint8),
                 mass=np.array([0.001, 0.002, 0.001, 0.0005], np.float32))
    assert formats.encode_softbody(el, faces=[(0, 1, 2), (1, 2, 3)],
                                   cells=[(0, 1, 2, 3, 0.000166666665, 4)], boundary_count=3) == text
    assert _text(z, "faces_reference") == text
