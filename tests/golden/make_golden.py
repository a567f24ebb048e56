"""Write the golden fixtures under tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_dump (built by oracle/ref/Makefile from the reference's
own material.h, VoxelRobot.cpp and SoftBody.cpp) on seeded inputs and stores
inputs + outputs as compressed .npz files (no pickles).  Re-run after
changing the cases:  python tests/golden/make_golden.py
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_DUMP = os.path.join(REPO, "oracle", "_ref", "ref_dump")

MAT_DT = np.dtype([("id", "<u4"), ("k", "<f4"), ("dL0", "<f4"), ("omega", "<f4"),
                   ("phi", "<f4"), ("enc", "<u4")])


def run(mode, payload):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(payload)
        subprocess.run([REF_DUMP, mode, fin, fout], check=True)
        with open(fout, "rb") as f:
            return f.read()


class Reader:
    def __init__(self, b):
        self.b, self.o = b, 0

    def arr(self, dtype, n):
        dt = np.dtype(dtype)
        a = np.frombuffer(self.b, dtype=dt, count=n, offset=self.o).copy()
        self.o += dt.itemsize * n
        return a

    def i32(self):
        return int(self.arr("<i4", 1)[0])


def materials():
    rng = np.random.default_rng(1234)
    encs = [0]
    encs += [1 << i for i in range(23)]
    encs += [(1 << i) | (1 << j) for i in range(23) for j in range(i + 1, 23)]
    for _ in range(200):
        k = int(rng.integers(3, 8))
        bits = rng.choice(23, size=k, replace=False)
        encs.append(int(sum(1 << int(b) for b in bits)))
    encs.append((1 << 23) - 1)
    encs = np.array(encs, dtype="<u4")
    payload = struct.pack("<i", len(encs)) + encs.tobytes()
    r = Reader(run("materials", payload))
    cc = r.i32()
    by_id = r.arr(MAT_DT, cc)
    base = r.arr(MAT_DT, 23)
    n = r.i32()
    rec = r.arr(np.dtype([("input", "<u4"), ("idx", "<i4")] + MAT_DT.descr), n)
    known = r.arr("<i4", 4)
    np.savez_compressed(os.path.join(HERE, "materials_reference.npz"),
                        by_id=by_id, base=base, dec_input=rec["input"], dec_idx=rec["idx"],
                        dec_id=rec["id"], dec_k=rec["k"], dec_dL0=rec["dL0"],
                        dec_omega=rec["omega"], dec_phi=rec["phi"], dec_encoding=rec["enc"],
                        known_answers=known)
    return cc


def voxel_case(name, n, size, res, enc, rng, nsets=2):
    nvox = n ** 3
    assert enc.shape == (nvox,)
    payload = struct.pack("<ffi", size, res, nvox) + enc.astype("<u4").tobytes()
    # position sets for fitness: identity and a perturbation; written after
    # the build so we need the masses first -> two passes
    first = Reader(run("voxel", payload + struct.pack("<i", 0)))
    nm = first.i32()
    m0 = first.arr(np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("id", "<u4"),
                             ("enc", "<u4")]), nm)
    pos0 = np.stack([m0["x"], m0["y"], m0["z"]], axis=1)
    sets = [pos0.copy()]
    for _ in range(nsets - 1):
        d = rng.normal(0.0, 0.02, size=pos0.shape).astype(np.float32)
        d[:, 0] += np.float32(rng.uniform(-0.05, 0.3))
        sets.append((pos0 + d).astype(np.float32))
    sets = np.stack(sets).astype("<f4")
    payload2 = payload + struct.pack("<i", len(sets)) + sets.tobytes()
    r = Reader(run("voxel", payload2))
    nm = r.i32()
    masses = r.arr(np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("id", "<u4"),
                             ("enc", "<u4")]), nm)
    ns = r.i32()
    springs = r.arr(np.dtype([("m0", "<u2"), ("m1", "<u2"), ("rest", "<f4"), ("mean", "<f4"),
                              ("id", "<u4"), ("enc", "<u4")]), ns)
    base_com = r.arr("<f4", 3)
    length = r.arr("<f4", 1)[0]
    reset_mean = r.arr("<f4", ns)
    reset_base_com = r.arr("<f4", 3)
    k = r.i32()
    fit = r.arr(np.dtype([("cx", "<f4"), ("cy", "<f4"), ("cz", "<f4"), ("fit", "<f4")]), k)
    np.savez_compressed(
        os.path.join(HERE, f"voxel_{name}.npz"),
        n=np.int32(n), size=np.float32(size), res=np.float32(res), voxel_enc=enc.astype("<u4"),
        mass_pos=np.stack([masses["x"], masses["y"], masses["z"]], 1),
        mass_id=masses["id"], mass_enc=masses["enc"],
        pairs=np.stack([springs["m0"], springs["m1"]], 1), rest_len=springs["rest"],
        mean_len=springs["mean"], spring_id=springs["id"], spring_enc=springs["enc"],
        base_com=base_com, length=np.float32(length), reset_mean_len=reset_mean,
        reset_base_com=reset_base_com, pos_sets=sets,
        com=np.stack([fit["cx"], fit["cy"], fit["cz"]], 1), fitness=fit["fit"])
    return nm, ns


def format_cases(cases):
    """VoxelRobot::Encode and SoftBody::Encode texts of built robots (base
    material voxels), and the reference's own Decode -> Encode round trip of
    each, into formats_reference.npz (texts as uint8 arrays)."""
    out = {}
    for name, n, size, res, enc in cases:
        payload = struct.pack("<ffi", size, res, n ** 3) + enc.astype("<u4").tobytes()
        r = Reader(run("encode", payload))
        texts = []
        for _ in range(4):
            k = r.i32()
            texts.append(r.arr("u1", k))
        vox, soft, vox_rt, soft_rt = texts
        out[f"{name}_n"] = np.int32(n)
        out[f"{name}_size"] = np.float32(size)
        out[f"{name}_res"] = np.float32(res)
        out[f"{name}_enc"] = enc.astype("<u4")
        out[f"{name}_voxel"] = vox
        out[f"{name}_soft"] = soft
        out[f"{name}_voxel_roundtrip"] = vox_rt
        out[f"{name}_soft_roundtrip"] = soft_rt
        print("format", name, len(vox), len(soft))
    out["names"] = np.array([c[0] for c in cases])
    # a soft body with faces and cells, written by this project's encoder,
    # read by the reference's SoftBody::Decode and written again
    sys.path.insert(0, REPO)
    from mercury_amd import formats
    from mercury_amd.robots import Element
    el = Element(np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [0, 0, 0.1]], np.float32),
                 np.array([22, 21, 1, 0], np.uint8),
                 np.array([[0, 1], [1, 2], [2, 3], [0, 3]], np.uint16),
                 np.array([0.1, 0.141421363, 0.141421363, 0.1], np.float32),
                 np.array([233, 22, 5, 0], np.uint8),
                 mass=np.array([0.001, 0.002, 0.001, 0.0005], np.float32))
    text = formats.encode_softbody(el, faces=[(0, 1, 2), (1, 2, 3)],
                                   cells=[(0, 1, 2, 3, 0.000166666665, 4)], boundary_count=3)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.txt"), os.path.join(td, "out.txt")
        with open(fin, "w") as f:
            f.write(text)
        subprocess.run([REF_DUMP, "decode", "soft", fin, fout], check=True)
        with open(fout, "rb") as f:
            out["faces_input"] = np.frombuffer(text.encode(), np.uint8)
            out["faces_reference"] = np.frombuffer(f.read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "formats_reference.npz"), **out)


def last_layer_air(n, enc):
    idx = np.arange(n ** 3)
    x, y, z = idx % n, (idx // n) % n, idx // (n * n)
    enc = enc.copy()
    enc[(x == n - 1) | (y == n - 1) | (z == n - 1)] = 1
    return enc


def main():
    if not os.path.exists(REF_DUMP):
        sys.exit("oracle/_ref/ref_dump missing: run make -C oracle/ref")
    print("composites", materials())
    rng = np.random.default_rng(20240611)
    cases = []
    # default 12^3 robot (VoxelRobot::Initialize: bone, last layer air)
    n = 12
    enc = last_layer_air(n, np.full(n ** 3, 1 << 22, dtype=np.uint32))
    cases.append(("default12", n, 12.0, 1.0, enc))
    # n=5, spacing 0.1: uniform over the 23 base materials
    n = 5
    enc = last_layer_air(n, (1 << rng.integers(0, 23, n ** 3)).astype(np.uint32))
    cases.append(("n5_random", n, 0.5, 10.0, enc))
    # n=9, spacing 0.1: 20% air, rest uniform over 22 non-air materials
    for s in range(3):
        n = 9
        mats = rng.integers(1, 23, n ** 3)
        mats[rng.random(n ** 3) < 0.2] = 0
        enc = last_layer_air(n, (1 << mats).astype(np.uint32))
        cases.append((f"n9_p20_{s}", n, 0.9, 10.0, enc))
    # n=9 sparse (60% air): several fragments, exercises Strip
    n = 9
    mats = rng.integers(1, 23, n ** 3)
    mats[rng.random(n ** 3) < 0.6] = 0
    cases.append(("n9_sparse", n, 0.9, 10.0, last_layer_air(n, (1 << mats).astype(np.uint32))))
    # n=7, composite voxel materials (two-bit encodings, as BuildFromCircles)
    n = 7
    a = rng.integers(1, 23, n ** 3)
    b = rng.integers(1, 23, n ** 3)
    enc = ((1 << a) | (1 << b)).astype(np.uint32)
    enc[rng.random(n ** 3) < 0.15] = 0
    cases.append(("n7_composite", n, 0.7, 10.0, last_layer_air(n, enc)))
    # n=6, no last-layer rule (materials everywhere), spacing 1
    n = 6
    cases.append(("n6_full", n, 6.0, 1.0, (1 << rng.integers(0, 23, n ** 3)).astype(np.uint32)))
    # all air
    n = 4
    cases.append(("n4_air", n, 4.0, 1.0, np.ones(n ** 3, dtype=np.uint32)))
    for c in cases:
        nm, ns = voxel_case(*c, rng=rng)
        print(c[0], nm, ns)
    # robot text formats of base-material robots: 3^3 over the 23 base
    # materials, and one n=5 and one n=9 case from above
    frng = np.random.default_rng(77)
    n = 4
    fcases = [("n4_base", n, 0.4, 10.0, last_layer_air(n, (1 << frng.integers(0, 23, n ** 3)).astype(np.uint32)))]
    fcases += [c for c in cases if c[0] in ("n5_random", "n9_p20_0")]
    format_cases(fcases)


if __name__ == "__main__":
    main()
