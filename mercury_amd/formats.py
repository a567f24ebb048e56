"""Text formats of the reference's robots (the Python twin of
include/mercury_hip_io.hpp; both write the same bytes).

  encode_softbody / decode_softbody      SoftBody::Encode / Decode
      (common/evolvables/SoftBody.cpp:147-309; item formats from
      common/simulator/{mass,spring,face,cell}.h)
  encode_voxel_robot / decode_voxel_robot  VoxelRobot::Encode / Decode
      (common/evolvables/VoxelRobot.cpp:504-610; voxel lines as
      Voxel::to_string, VoxelRobot.h:51-63)
  solution_file_name                     Optimizer::WriteSolutions
      (optimize/optimizer/optimizer_impl.h:35-48)

Soft body: "type=NNRobot\\n", then "masses=", "springs=", "faces=", "cells="
each followed by its items joined with ';' and a newline after the last
item (none for an empty list, as in the reference), then
"boundary_count=<n>\\n"; numbers with 9 significant digits (%.9g).  Voxel
robot: "type=VoxelRobot\\n", the three sizes and the resolution on lines of
their own (%f), then one "v ..." line per lattice point with its id,
indices, center, base and base material (k, dL0, omega, phi, colour), all
floats as %f.
"""
import numpy as np

from .robots import Element

# display colours of the reference's base materials (material.h:111-133):
# adductor muscles, abductor muscles, tissue, bone
_COLOURS = {"adductor": (106, 211, 250), "abductor": (89, 217, 78), "tissue": (217, 102, 78),
            "bone": (240, 209, 98)}


def _g9(v):
    return "%.9g" % float(v)


def _f6(v):
    return "%f" % float(v)


def _join(items, out):
    out.append(";".join(items))
    if items:
        out.append("\n")


def encode_softbody(el, ids=None, faces=(), cells=(), boundary_count=0):
    """SoftBody::Encode of an Element (masses at their proto positions el.pos).
    ids: mass ids (default 0..n-1); faces: (m0, m1, m2) rows; cells:
    (m0, m1, m2, m3, mean_volume, material_encoding) rows."""
    n = el.n_masses
    ids = np.arange(n) if ids is None else np.asarray(ids)
    mass = np.ones(n, np.float32) if el.mass is None else np.asarray(el.mass, np.float32)
    pos = np.asarray(el.pos, np.float32)
    rest = el.mean_length if el.rest_length is None else el.rest_length
    out = ["type=NNRobot\n", "masses="]
    _join([f"{int(ids[i])},{_g9(pos[i, 0])},{_g9(pos[i, 1])},{_g9(pos[i, 2])},{_g9(mass[i])},"
           f"{int(el.mass_material[i])}" for i in range(n)], out)
    out.append("springs=")
    _join([f"{int(el.pairs[k, 0])},{int(el.pairs[k, 1])},{_g9(np.float32(rest[k]))},"
           f"{_g9(np.float32(el.mean_length[k]))},{int(el.spring_material[k])}"
           for k in range(el.n_springs)], out)
    out.append("faces=")
    _join([f"{int(f[0])},{int(f[1])},{int(f[2])}" for f in faces], out)
    out.append("cells=")
    _join([f"{int(c[0])},{int(c[1])},{int(c[2])},{int(c[3])},{_g9(np.float32(c[4]))},"
           f"{int(c[5]) & 0xFFFF}" for c in cells], out)
    out.append(f"boundary_count={int(boundary_count)}\n")
    return "".join(out)


def _fields(text):
    """key=value fields in file order; an empty list leaves the next key on
    the same line."""
    keys = ("springs=", "faces=", "cells=", "boundary_count=")
    out = []
    for line in text.split("\n"):
        rest = line
        while rest and "=" in rest:
            key, value = rest.split("=", 1)
            tail = ""
            if value.startswith(keys):
                tail, value = value, ""
            out.append((key, value))
            rest = tail
    return out


def decode_softbody(text):
    """SoftBody::Decode of a file's text.  Returns (Element, ids, faces,
    cells, boundary_count); the element's materials are composite ids."""
    ids, pos, mass, mmat = [], [], [], []
    pairs, rest, mean, smat = [], [], [], []
    faces, cells, boundary = [], [], 0
    for key, value in _fields(text):
        if key == "type":
            if value != "NNRobot":
                raise ValueError("not a soft body file")
            continue
        if not value:
            continue
        items = value.split(";")
        if key == "masses":
            for it in items:
                f = it.split(",")
                if len(f) != 6:
                    raise ValueError(f"bad mass {it!r}")
                ids.append(int(f[0]))
                pos.append([np.float32(float(x)) for x in f[1:4]])
                mass.append(np.float32(float(f[4])))
                mmat.append(int(f[5]))
        elif key == "springs":
            for it in items:
                f = it.split(",")
                if len(f) != 5:
                    raise ValueError(f"bad spring {it!r}")
                pairs.append((int(f[0]), int(f[1])))
                rest.append(np.float32(float(f[2])))
                mean.append(np.float32(float(f[3])))
                smat.append(int(f[4]))
        elif key == "faces":
            faces = [tuple(int(x) for x in it.split(",")) for it in items]
        elif key == "cells":
            for it in items:
                f = it.split(",")
                cells.append((int(f[0]), int(f[1]), int(f[2]), int(f[3]), np.float32(float(f[4])),
                              int(f[5])))
        elif key == "boundary_count":
            boundary = int(value)
    el = Element(np.array(pos, np.float32).reshape(-1, 3), np.array(mmat, np.uint8),
                 np.array(pairs, np.uint16).reshape(-1, 2), np.array(mean, np.float32),
                 np.array(smat, np.uint8), mass=np.array(mass, np.float32))
    el.rest_length = np.array(rest, np.float32)
    return el, np.array(ids, np.int64), faces, cells, boundary


def reference_voxel_materials(material_set):
    """(k, dL0, omega, phi, r, g, b, a) of the reference's base materials by
    base id (0 = air), as a voxel line carries them."""
    rows = [(0.0,) * 8]
    for i in range(1, 23):
        k, dl, om, ph = (float(x) for x in material_set.by_id(i)[0])
        c = _COLOURS["adductor" if i <= 10 else "abductor" if i <= 20 else
                     "tissue" if i == 21 else "bone"]
        rows.append((k, dl, om, ph) + tuple(float(np.float32(np.float32(x) / np.float32(255.0)))
                                            for x in c) + (0.0,))
    return rows


def encode_voxel_robot(materials, shape, resolution, table, sizes=None):
    """VoxelRobot::Encode.  materials: base id per lattice point (0 = air),
    point i at (i % nx, i // nx % ny, i // nx // ny); shape (nx, ny, nz);
    table: rows (k, dL0, omega, phi, r, g, b, a) by base id; sizes default
    to shape / resolution."""
    nx, ny, nz = shape
    res = np.float32(resolution)
    if sizes is None:
        sizes = tuple(np.float32(s) / res for s in shape)
    mats = np.asarray(materials)
    if mats.shape != (nx * ny * nz,):
        raise ValueError("one material per lattice point expected")
    inv = np.float32(1.0) / res
    half = inv * np.float32(0.5)
    out = ["type=VoxelRobot\n"] + [_f6(np.float32(s)) + "\n" for s in sizes] + [_f6(res) + "\n"]
    for i in range(nx * ny * nz):
        ix, iy, iz = i % nx, i // nx % ny, i // nx // ny % nz
        b = [inv * np.float32(ix), inv * np.float32(iy), inv * np.float32(iz)]
        c = [x + half for x in b]
        m = table[int(mats[i])]
        out.append(f"v {i},{ix},{iy},{iz}," + ",".join(_f6(x) for x in c + b) + "," +
                   ", ".join(_f6(np.float32(x)) for x in m) + "\n")
    return "".join(out)


def decode_voxel_robot(text, table):
    """VoxelRobot::Decode of a file's text.  Returns (materials, shape,
    resolution, sizes) with each voxel's material matched to the table by
    (k, dL0, omega, phi) at the 6 decimals the format keeps."""
    lines = text.split("\n")
    if len(lines) < 5 or lines[0] != "type=VoxelRobot":
        raise ValueError("not a voxel robot file")
    sizes = tuple(np.float32(float(x)) for x in lines[1:4])
    res = np.float32(float(lines[4]))
    shape = tuple(int(round(float(s * res))) for s in sizes)
    n = shape[0] * shape[1] * shape[2]

    def key(vals):
        return tuple(int(np.rint(float(np.float32(v)) * 1e6)) for v in vals[:4])
    lookup = {}
    for t, row in enumerate(table):
        lookup.setdefault(key(row), t)
    mats = np.zeros(n, np.uint8)
    seen = 0
    for line in lines[5:]:
        if not line:
            continue
        if line[0] != "v":
            raise ValueError("bad voxel line")
        f = line[1:].split(",")
        if len(f) != 18:
            raise ValueError(f"bad voxel line {line!r}")
        vid = int(f[0])
        if not 0 <= vid < n:
            raise ValueError("voxel id out of range")
        k = key([float(x) for x in f[10:14]])
        if k not in lookup:
            raise ValueError("voxel material not in the table")
        mats[vid] = lookup[k]
        seen += 1
    if seen != n:
        raise ValueError("voxel count does not match the sizes")
    return mats, shape, float(res), sizes


def voxel_encodings(materials):
    """Bit encodings (1 << base id; air = 1) of per-point base ids, the input
    of voxel_build / mh_voxel_build."""
    return (np.uint32(1) << np.asarray(materials, np.uint32)).astype(np.uint32)


def solution_file_name(i, fitness):
    """Optimizer::WriteSolutions' file name of solution i."""
    return f"solution_{int(i)}_fitness_{float(np.float32(fitness)):f}"
