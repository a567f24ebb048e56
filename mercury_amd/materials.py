"""Material sets: the reference's 22-material table or a custom base list.

Mirrors the reference's materials class (common/simulator/material.h:109-322):
id_lookup, decode, encodedCompositeIdx.  The tables are built by
libmercury_hip (mh_materials_*).
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import check, load


class MaterialSet:
    """Owns an mh_material_set handle."""

    def __init__(self, handle, base_rows):
        self._h = handle
        self.base_rows = base_rows  # (k, dL0, omega, phi) of each non-air base

    @classmethod
    def reference(cls):
        h = ctypes.c_void_p()
        check(load().mh_materials_reference(ctypes.byref(h)))
        s = cls(h, None)
        s.base_rows = [tuple(s.by_id(i)[0]) for i in range(1, 23)]
        return s

    @classmethod
    def custom(cls, rows):
        """rows: list of (k, dL0, omega, phi); base i gets id i+1, bit i+1."""
        arr = (_lib.BaseMaterial * len(rows))(*[_lib.BaseMaterial(*r) for r in rows])
        h = ctypes.c_void_p()
        check(load().mh_materials_create(len(rows), arr, ctypes.byref(h)))
        return cls(h, list(rows))

    @property
    def handle(self):
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) and _lib._lib is not None:
            _lib._lib.mh_materials_destroy(self._h)
            self._h = None

    def count(self):
        n = load().mh_materials_count(self._h)
        if n < 0:
            check(n)
        return n

    def by_id(self, i):
        """id_lookup(i) -> ((k, dL0, omega, phi), encoding)."""
        out = (ctypes.c_float * 4)()
        enc = ctypes.c_uint32()
        check(load().mh_materials_by_id(self._h, int(i), out, ctypes.byref(enc)))
        return np.array(out[:], np.float32), int(enc.value)

    def decode(self, encoding):
        """decode(encoding) -> (composite id, encoded index)."""
        i = ctypes.c_uint32()
        idx = ctypes.c_int32()
        check(load().mh_materials_decode(self._h, int(encoding), ctypes.byref(i),
                                         ctypes.byref(idx)))
        return int(i.value), int(idx.value)

    def table(self):
        """Arrays (k, dL0, omega, phi) indexed by composite id."""
        n = self.count()
        rows = np.stack([self.by_id(i)[0] for i in range(n)])
        return {"k": rows[:, 0].copy(), "dL0": rows[:, 1].copy(),
                "omega": rows[:, 2].copy(), "phi": rows[:, 3].copy()}


# Materials of the voxel workloads (stiff, soft, two actuated phases at 2 Hz
# with 14 % amplitude).  The actuated stiffness is not given by the workload;
# it is set between soft and stiff, like the reference's muscles (k=50) sit
# between tissue (40) and bone (100).
STIFF_K, SOFT_K, MUSCLE_K = 1.0e4, 2.0e3, 5.0e3
ACT_AMPLITUDE, ACT_OMEGA = 0.14, 2.0 * 2.0 * math.pi
WORKLOAD_ROWS = [
    (STIFF_K, 0.0, ACT_OMEGA, 0.0),               # id 1: stiff
    (SOFT_K, 0.0, ACT_OMEGA, 0.0),                # id 2: soft
    (MUSCLE_K, ACT_AMPLITUDE, ACT_OMEGA, 0.0),    # id 3: actuated, phase 0
    (MUSCLE_K, ACT_AMPLITUDE, ACT_OMEGA, math.pi),  # id 4: actuated, phase pi
]


def workload_materials():
    return MaterialSet.custom(WORKLOAD_ROWS)
