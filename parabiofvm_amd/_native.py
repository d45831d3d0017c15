"""ctypes binding of libparabiofvm_hip.so (C ABI declared in include/parabiofvm_hip.h).

The library is built in-tree (parabiofvm_amd/libparabiofvm_hip.so, see
parabiofvm_amd/csrc/Makefile).  Loading fails loudly when it is missing: there
is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# PARABIOFVM_HIP_LIB selects another build of the same library (e.g. one compiled with other tuning flags).
LIB_PATH = os.environ.get("PARABIOFVM_HIP_LIB") or os.path.join(HERE, "libparabiofvm_hip.so")

BFVM_SWEEP_PARTITIONED = 0
BFVM_SWEEP_EXACT = 1
BFVM_SWEEP_STAGED = 2
BFVM_STORAGE_F32 = 0
BFVM_STORAGE_BF16 = 1
BFVM_STORAGE_F16 = 2
REDUCTION = {"atomic": 0, "sorted": 1}
STORAGE = {"f32": BFVM_STORAGE_F32, "bf16": BFVM_STORAGE_BF16, "f16": BFVM_STORAGE_F16}


class BfvmError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


_F = ctypes.POINTER(ctypes.c_float)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_I32 = ctypes.POINTER(ctypes.c_int32)


class MicroenvironmentDesc(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int32),
        ("bounding_box_mins", ctypes.c_int32 * 3),
        ("bounding_box_maxs", ctypes.c_int32 * 3),
        ("voxel_shape", ctypes.c_int32 * 3),
        ("substrates_count", ctypes.c_int32),
        ("diffusion_time_step", ctypes.c_float),
        ("diffusion_coefficients", _F),
        ("decay_rates", _F),
        ("dirichlet_min_boundary_values", _F * 3),
        ("dirichlet_max_boundary_values", _F * 3),
        ("dirichlet_min_boundary_conditions", _U8 * 3),
        ("dirichlet_max_boundary_conditions", _U8 * 3),
        ("dirichlet_interior_voxels_count", ctypes.c_int32),
        ("dirichlet_interior_voxels", _I32),
        ("dirichlet_interior_values", _F),
        ("dirichlet_interior_conditions", _U8),
        ("compute_internalized_substrates", ctypes.c_int32),
    ]


class AgentsDesc(ctypes.Structure):
    _fields_ = [
        ("agents_count", ctypes.c_int32),
        ("secretion_rates", _F),
        ("saturation_densities", _F),
        ("uptake_rates", _F),
        ("net_export_rates", _F),
        ("internalized_substrates", _F),
        ("fraction_released_at_death", _F),
        ("volumes", _F),
        ("positions", _F),
    ]


class DeviceBuffers(ctypes.Structure):
    _fields_ = [
        ("densities", ctypes.c_void_p),
        ("gradients", ctypes.c_void_p),
        ("internalized_substrates", ctypes.c_void_p),
        ("voxel_count", ctypes.c_int64),
        ("grid_shape", ctypes.c_int32 * 3),
        ("substrates_count", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
        ("storage", ctypes.c_int32),
        ("density_scales", ctypes.c_void_p),
    ]


# name -> (restype, argtypes); every exported symbol of include/parabiofvm_hip.h
PROTOTYPES = {
    "bfvm_last_error": (ctypes.c_char_p, []),
    "bfvm_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "bfvm_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_initialize": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MicroenvironmentDesc)]),
    "bfvm_set_storage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "bfvm_set_secretion_reduction": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "bfvm_bulk_store_rates": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "bfvm_bulk_solve": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_set_sweep_algorithm": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "bfvm_store_densities": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "bfvm_store_densities_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "bfvm_store_agents": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(AgentsDesc)]),
    "bfvm_load_densities": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "bfvm_load_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "bfvm_load_internalized": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "bfvm_diffusion_solve": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_dirichlet_solve": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_gradient_solve": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_cell_simulate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "bfvm_cell_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "bfvm_timestep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "bfvm_wait_for_all": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_get_device_buffers": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DeviceBuffers)]),
    "bfvm_sweep_partition": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I32, _I32, _I32]),
    "bfvm_set_sweep_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "bfvm_set_gradient_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "bfvm_set_sweep_workgroups": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "bfvm_set_gradient_stacked": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "bfvm_set_gradient_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "bfvm_autotune": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_batch_autotune": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_tuning_save": (ctypes.c_int, [ctypes.c_char_p]),
    "bfvm_tuning_load": (ctypes.c_int, [ctypes.c_char_p]),
    "bfvm_exact_coefficients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p]),
    "bfvm_voxel_positions": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "bfvm_voxel_positions_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                                    ctypes.c_void_p]),
    "bfvm_benchmark": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _F]),
    "bfvm_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "bfvm_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_batch_timestep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "bfvm_batch_wait": (ctypes.c_int, [ctypes.c_void_p]),
    "bfvm_batch_benchmark": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _F]),
}

KERNELS = {"sweep_x": 0, "sweep_y": 1, "sweep_z": 2, "gradient": 3, "dirichlet_interior": 4, "dirichlet_faces": 5,
           "copy": 6,
           "secretion": 10, "secretion_recompute": 11, "diffusion": 12, "gradient_phase": 13, "timestep": 14}

_lib = None


def load():
    """Load the HIP library; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not found: build it with __graft_entry__.build() "
                          f"(make -C parabiofvm_amd/csrc)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(fn, code):
    if code != 0:
        msg = load().bfvm_last_error()
        raise BfvmError(fn, code, msg.decode() if msg else "")


def tuning_save(path):
    """Write the process's launch-shape tuning cache to a text file."""
    check("bfvm_tuning_save", load().bfvm_tuning_save(os.fsencode(path)))


def tuning_load(path):
    """Merge a tuning cache file written by tuning_save into the process's cache."""
    check("bfvm_tuning_load", load().bfvm_tuning_load(os.fsencode(path)))
