"""ctypes binding of libmercury_hip.so (the C-ABI in include/mercury_hip.h).

The library is the only compute path: if it is missing or fails to load, every
call raises.  There is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmercury_hip.so")

c_int, c_float, c_uint32, c_uint64 = ctypes.c_int, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint64
vp = ctypes.c_void_p
u8p = ctypes.POINTER(ctypes.c_uint8)
u16p = ctypes.POINTER(ctypes.c_uint16)
u32p = ctypes.POINTER(ctypes.c_uint32)
i32p = ctypes.POINTER(ctypes.c_int32)
f32p = ctypes.POINTER(ctypes.c_float)

MH_SOLVER_XPBD_REF, MH_SOLVER_FORCE = 0, 1
MH_KERNEL_AUTO, MH_KERNEL_LDS, MH_KERNEL_STREAM, MH_KERNEL_MASS, MH_KERNEL_MASS_GATHER = 0, 1, 2, 4, 8
MH_KERNEL_MASS_STREAM = 16
MH_FITNESS_REFERENCE, MH_FITNESS_DISPLACEMENT = 0, 1


class MhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mercury_hip error {code}: {msg}")
        self.code = code


class BaseMaterial(ctypes.Structure):
    _fields_ = [("k", c_float), ("dL0", c_float), ("omega", c_float), ("phi", c_float)]


class Config(ctypes.Structure):
    _fields_ = [("solver", c_int), ("dt", c_float), ("gravity", c_float),
                ("ground_k", c_float), ("mu_s", c_float), ("mu_k", c_float),
                ("drag", c_float), ("damping", c_float), ("device", c_int),
                ("kernel", c_int), ("block_size", c_int), ("robots_per_block", c_int),
                ("substeps_per_launch", c_int), ("use_graph", c_int)]


class Population(ctypes.Structure):
    _fields_ = [("n_robots", ctypes.c_int32), ("mass_offset", u32p), ("spring_offset", u32p),
                ("pos", f32p), ("vel", f32p), ("mass", f32p), ("mass_material_id", u8p),
                ("spring_pairs", u16p), ("spring_mean_length", f32p),
                ("spring_material_id", u8p)]


class SimInfo(ctypes.Structure):
    _fields_ = [(n, c_int) for n in (
        "n_robots", "n_masses", "n_springs", "n_active_springs", "kernel", "block_size",
        "n_blocks", "substeps_per_launch", "lds_bytes", "max_masses_per_block",
        "max_springs_per_block", "n_lds_robots", "n_stream_robots", "stream_block_size",
        "robots_per_block", "n_stream_springs", "stream_chunk", "graph_builds",
        "stream_record_bytes")]


# name -> (restype, argtypes)
_SIGS = {
    "mh_last_error": (ctypes.c_char_p, []),
    "mh_version": (c_int, []),
    "mh_materials_reference": (c_int, [ctypes.POINTER(vp)]),
    "mh_materials_create": (c_int, [c_int, ctypes.POINTER(BaseMaterial), ctypes.POINTER(vp)]),
    "mh_materials_destroy": (c_int, [vp]),
    "mh_materials_count": (c_int, [vp]),
    "mh_materials_by_id": (c_int, [vp, c_int, f32p, u32p]),
    "mh_materials_decode": (c_int, [vp, c_uint32, u32p, i32p]),
    "mh_voxel_max_springs": (c_int, [c_int, c_int, c_int]),
    "mh_voxel_build": (c_int, [vp, c_int, c_int, c_int, c_float, u32p, f32p, u8p, u16p, f32p,
                               u8p, u32p, c_int, ctypes.POINTER(c_int), u32p]),
    "mh_voxel_template": (c_int, [c_int, c_int, c_int, c_float, f32p, u16p, f32p, u16p, c_int, i32p]),
    "mh_voxel_build_batch": (c_int, [vp, c_int, c_int, c_int, c_int, c_float, u32p, c_int, f32p,
                                     u8p, u32p, u16p, f32p, u8p, ctypes.c_int64]),
    "mh_rest_lengths": (c_int, [f32p, u16p, c_int, f32p]),
    "mh_baseline": (c_int, [f32p, u8p, c_int, f32p, f32p]),
    "mh_config_default": (None, [ctypes.POINTER(Config)]),
    "mh_sim_create": (c_int, [ctypes.POINTER(Config), vp, ctypes.POINTER(vp)]),
    "mh_sim_destroy": (c_int, [vp]),
    "mh_sim_set_population": (c_int, [vp, ctypes.POINTER(Population)]),
    "mh_sim_simulate": (c_int, [vp, c_float, c_int]),
    "mh_sim_step": (c_int, [vp, c_int, c_int]),
    "mh_sim_synchronize": (c_int, [vp]),
    "mh_sim_collect": (c_int, [vp, f32p, f32p, f32p]),
    "mh_sim_fitness": (c_int, [vp, c_int, f32p]),
    "mh_sim_com": (c_int, [vp, f32p]),
    "mh_sim_reset": (c_int, [vp]),
    "mh_sim_total_time": (c_float, [vp]),
    "mh_sim_set_total_time": (c_int, [vp, c_float]),
    "mh_sim_set_voxel_population": (c_int, [vp, c_int, c_int, c_int, c_int, c_float, c_float, u32p]),
    "mh_sim_voxel_materials": (c_int, [vp, u8p, u8p, u32p]),
    "mh_sim_set_stress": (c_int, [vp, f32p]),
    "mh_sim_get_time_step": (c_float, [vp]),
    "mh_sim_set_time_step": (c_int, [vp, c_float]),
    "mh_sim_get_info": (c_int, [vp, ctypes.POINTER(SimInfo)]),
    "mh_sim_autotune": (c_int, [vp, c_int]),
    "mh_sim_device_state": (c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp),
                                    ctypes.POINTER(vp)]),
    "mh_sim_set_timing": (c_int, [vp, c_int]),
    "mh_sim_last_kernel_ms": (c_int, [vp, f32p, ctypes.POINTER(c_int)]),
    "mh_knn_batch": (c_int, [c_int, c_int, u32p, f32p, c_int, u16p, f32p]),
    "mh_knn_springs": (c_int, [vp, c_int, c_int, u16p, f32p, u32p, c_float, u16p, f32p, u8p,
                               c_int, ctypes.POINTER(c_int)]),
    "mh_nn_materials": (c_int, [c_int, i32p, f32p, f32p, c_int, u8p]),
    "mh_mlp_create": (c_int, [c_int, c_int, i32p, c_uint64, ctypes.POINTER(vp)]),
    "mh_mlp_destroy": (c_int, [vp]),
    "mh_mlp_param_count": (c_int, [vp]),
    "mh_mlp_get_params": (c_int, [vp, f32p]),
    "mh_mlp_set_params": (c_int, [vp, f32p]),
    "mh_mlp_forward": (c_int, [vp, f32p, c_int, f32p]),
    "mh_mlp_forward_noalloc": (c_int, [vp, vp, c_int, vp, vp]),
    "mh_mlp_loss_grad": (c_int, [vp, f32p, f32p, c_int, f32p, f32p]),
    "mh_mlp_train": (c_int, [vp, f32p, f32p, c_int, c_int, c_int, c_float, c_uint64, f32p]),
    "mh_cnn_create": (c_int, [c_int, ctypes.c_uint64, ctypes.POINTER(vp)]),
    "mh_cnn_destroy": (c_int, [vp]),
    "mh_cnn_param_count": (c_int, [vp]),
    "mh_cnn_get_params": (c_int, [vp, f32p]),
    "mh_cnn_set_params": (c_int, [vp, f32p]),
    "mh_cnn_forward": (c_int, [vp, u8p, c_int, f32p]),
    "mh_cnn_loss_grad": (c_int, [vp, u8p, f32p, c_int, f32p, f32p]),
    "mh_cnn_train": (c_int, [vp, u8p, f32p, c_int, c_int, c_int, c_float, ctypes.c_uint64, f32p]),
    "mh_calibrate_lds": (c_int, [c_int, f32p]),
    "mh_calibrate_copy": (c_int, [c_int, ctypes.c_size_t, c_int, f32p]),
    "mh_calibrate_valu": (c_int, [c_int, c_int, f32p]),
    "mh_check_arith": (c_int, [c_int, c_int, ctypes.c_uint32, ctypes.c_uint32,
                               ctypes.POINTER(ctypes.c_uint64)]),
}

_lib = None


def load():
    """Load libmercury_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built: run python -m mercury_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = load().mh_last_error()
        raise MhError(rc, msg.decode() if msg else "")
    return rc


def ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)
