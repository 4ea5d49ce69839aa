"""ctypes mirror of include/mpm_engine.h and the loader of the engine library.

The structures here are plain data; they are shared by every Python caller of
the C ABI.  Loading the engine fails loudly when the HIP library is missing:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PACKAGE_DIR = Path(__file__).resolve().parent
LIB_PATH = PACKAGE_DIR / "libsquishy_volumes_amd.so"

# ParticleFlags (rust/crates/file_frame/src/particles.rs:23-33) + extensions.
IS_SOLID = 1 << 0
IS_FLUID = 1 << 1
USE_VISCOSITY = 1 << 2
USE_SAND_ALPHA = 1 << 3
HAS_GOAL = 1 << 4
TOMBSTONED = 1 << 5
FAILED = 1 << 6
FIXED_COROTATED = 1 << 8
SNOW = 1 << 9

STATUS_TABLE_TRIES_EXCEEDED = 1
STATUS_TABLE_ENTRY_MISSING = 2
STATUS_INDIRECT_LIMIT_EXCEEDED = 4
STATUS_PARTICLE_CLOSE_TO_INVERTED = 8
STATUS_ENERGY_ERROR = 16
STATUS_OUT_OF_GRID = 32
STATUS_SVD_FP64 = 64

OPT_GRID_GRAVITY = 1 << 0
OPT_STICKY_PLANE = 1 << 1
OPT_BOX_WALLS = 1 << 2
OPT_SPHERE = 1 << 3

OK = 0
ERR_ENERGY = 1
ERR_CLOSE_TO_INVERTED = 2
ERR_INVALID_ARGUMENT = -1
ERR_HIP = -2
ERR_NO_DEVICE = -3
ERR_OUT_OF_MEMORY = -4
ERR_GRID_TOO_LARGE = -5
ERR_CANCELED = -6
ERR_WRONG_FRAME = -7
ERR_ZERO_TIME_STEP = -8
ERR_OUT_OF_GRID = -9
ERR_CAPACITY = -10

# Substeps one mpm_step_to call runs at most (MPM_MAX_SUBSTEPS_PER_CALL).
MAX_SUBSTEPS_PER_CALL = 1 << 20

# mpm_particle_parameters as a numpy record (36 bytes, no padding).
PARAMS_DTYPE = np.dtype(
    [
        ("mass", "<f4"),
        ("initial_volume", "<f4"),
        ("viscosity_dynamic", "<f4"),
        ("viscosity_bulk", "<f4"),
        ("mu", "<f4"),
        ("lambda", "<f4"),
        ("sand_alpha", "<f4"),
        ("bulk_modulus", "<f4"),
        ("exponent", "<i4"),
    ]
)


class MpmConsts(C.Structure):
    _fields_ = [
        ("grid_node_size", C.c_float),
        ("time_step", C.c_float),
        ("domain_min", C.c_float * 3),
        ("domain_max", C.c_float * 3),
        ("frames_per_second", C.c_uint32),
        ("options", C.c_uint32),
        ("plane_z", C.c_float),
        ("wall_min", C.c_float * 3),
        ("wall_max", C.c_float * 3),
        ("snow_compression", C.c_float),
        ("snow_stretch", C.c_float),
        ("snow_hardening", C.c_float),
        ("device", C.c_int32),
        ("sphere_center", C.c_float * 3),
        ("sphere_radius", C.c_float),
        ("scene_period", C.c_float * 3),
        ("scene_repeat", C.c_uint32 * 3),
    ]


class MpmParticleParameters(C.Structure):
    _fields_ = [(name, C.c_float) for name in PARAMS_DTYPE.names[:-1]] + [("exponent", C.c_int32)]


assert C.sizeof(MpmParticleParameters) == PARAMS_DTYPE.itemsize


class MpmParticlesView(C.Structure):
    _fields_ = [
        ("count", C.c_uint64),
        ("time", C.c_double),
        ("flags", C.c_void_p),
        ("parameters", C.c_void_p),
        ("elastic_energies", C.c_void_p),
        ("collider_bits", C.c_void_p),
        ("positions", C.c_void_p),
        ("position_gradients", C.c_void_p),
        ("velocities", C.c_void_p),
        ("velocity_gradients", C.c_void_p),
        ("initial_positions", C.c_void_p),
        ("plastic_jp", C.c_void_p),
    ]


class MpmFrameInput(C.Structure):
    _fields_ = [
        ("frame", C.c_uint64),
        ("gravity_a", C.c_float * 3),
        ("gravity_b", C.c_float * 3),
        ("has_b", C.c_int32),
        ("particle_flags_a", C.c_void_p),
        ("particle_flags_b", C.c_void_p),
        ("goal_positions_a", C.c_void_p),
        ("goal_positions_b", C.c_void_p),
        ("vertex_positions_a", C.c_void_p),
        ("vertex_positions_b", C.c_void_p),
        ("triangle_frictions_a", C.c_void_p),
        ("triangle_frictions_b", C.c_void_p),
        ("triangle_dampings_a", C.c_void_p),
        ("triangle_dampings_b", C.c_void_p),
    ]


class MpmColliderMesh(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint32),
        ("num_triangles", C.c_uint32),
        ("triangles", C.c_void_p),
        ("triangle_colliders", C.c_void_p),
    ]


class MpmStatus(C.Structure):
    _fields_ = [
        ("device_status", C.c_uint32),
        ("result", C.c_int32),
        ("substeps", C.c_uint64),
        ("time", C.c_double),
        ("active_blocks", C.c_uint64),
        ("active_nodes", C.c_uint64),
        ("svd_fp64_fallbacks", C.c_uint64),
        ("graph_launches", C.c_uint64),
        ("grid_blocks", C.c_uint64),
        ("window_redos", C.c_uint64),
    ]


class MpmGridView(C.Structure):
    _fields_ = [
        ("capacity", C.c_uint64),
        ("count", C.c_uint64),
        ("node_ids", C.c_void_p),
        ("collider_bits", C.c_void_p),
        ("masses", C.c_void_p),
        ("velocities", C.c_void_p),
    ]


class MpmTuning(C.Structure):
    _fields_ = [
        ("resort_interval", C.c_int32),
        ("persistent_blocks", C.c_int32),
        ("graph", C.c_int32),
        ("particles_per_lane", C.c_int32),
        ("grid_update_threads", C.c_int32),
        ("learned_fused", C.c_int32),
        ("grid_window", C.c_int32),
    ]


class MpmClosure(C.Structure):
    _fields_ = [
        ("params", C.c_void_p),
        ("hull_lo", C.c_float * 6),
        ("hull_hi", C.c_float * 6),
        ("judge_stride", C.c_int32),
    ]


class MpmTimeStepLimits(C.Structure):
    _fields_ = [
        ("by_velocity", C.c_float),
        ("by_deformation", C.c_float),
        ("by_sound", C.c_float),
        ("by_isolated", C.c_float),
    ]


class MpmClosureStats(C.Structure):
    _fields_ = [
        ("evaluated", C.c_uint64),
        ("fallbacks", C.c_uint64),
        ("judged", C.c_uint64),
        ("err2", C.c_double),
        ("ref2", C.c_double),
    ]


# Every exported symbol, with its ctypes signature.
SIGNATURES = {
    "mpm_create": (C.c_int, [C.POINTER(MpmConsts), C.POINTER(MpmParticlesView), C.POINTER(MpmTuning), C.POINTER(C.c_void_p)]),
    "mpm_step_to": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(MpmFrameInput), C.c_void_p, C.POINTER(MpmStatus)]),
    "mpm_read_particles": (C.c_int, [C.c_void_p, C.POINTER(MpmParticlesView)]),
    "mpm_read_grid": (C.c_int, [C.c_void_p, C.POINTER(MpmGridView)]),
    "mpm_destroy": (None, [C.c_void_p]),
    "mpm_error_string": (C.c_char_p, [C.c_int]),
    "mpm_time": (C.c_double, [C.c_void_p]),
    "mpm_set_profiling": (C.c_int, [C.c_void_p, C.c_int32]),
    "mpm_kernel_times": (
        C.c_int,
        [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_int32)],
    ),
    "mpm_stream": (C.c_void_p, [C.c_void_p]),
    "mpm_device_count": (C.c_int, []),
    "mpm_dev_radix_sort_temp_bytes": (C.c_size_t, [C.c_uint64]),
    "mpm_dev_radix_sort": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32, C.c_void_p, C.c_void_p],
    ),
    "mpm_dev_radix_sort_noalloc_max_keys": (C.c_uint64, []),
    "mpm_dev_radix_sort_noalloc": (
        C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32, C.c_void_p],
    ),
    "mpm_dev_scan_temp_bytes": (C.c_size_t, [C.c_uint64]),
    "mpm_dev_exclusive_scan": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    "mpm_dev_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "mpm_dev_stress": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p],
    ),
    "mpm_dev_svd": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "mpm_mlp_param_count": (C.c_int32, []),
    "mpm_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "mpm_mlp_grad": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p],
    ),
    "mpm_mlp_adam": (
        C.c_int,
        [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_void_p],
    ),
    "mpm_set_closure": (C.c_int, [C.c_void_p, C.POINTER(MpmClosure)]),
    "mpm_read_closure_stats": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(MpmClosureStats)]),
    "mpm_closure_pairs": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64,
                                    C.POINTER(C.c_uint64)]),
    "mpm_run_stage": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(MpmFrameInput)]),
    "mpm_read_time_step_limits": (C.c_int, [C.c_void_p, C.POINTER(MpmTimeStepLimits)]),
    "mpm_set_time_step": (C.c_int, [C.c_void_p, C.c_float]),
    "mpm_set_adaptive_time_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_float]),
    "mpm_debug_cancel_after": (C.c_int, [C.c_void_p, C.c_int64]),
    "mpm_set_collider_mesh": (C.c_int, [C.c_void_p, C.POINTER(MpmColliderMesh)]),
    "mpm_frame_file_size": (C.c_uint64, [C.POINTER(MpmParticlesView), C.POINTER(MpmGridView)]),
    "mpm_frame_encode": (C.c_int, [C.POINTER(MpmParticlesView), C.POINTER(MpmGridView), C.c_void_p, C.c_uint64,
                                   C.POINTER(C.c_uint64)]),
    "mpm_frame_decode_counts": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "mpm_frame_decode": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(MpmParticlesView), C.POINTER(MpmGridView)]),
    "mpm_read_grid_momentum": (C.c_int, [C.c_void_p, C.POINTER(MpmGridView)]),
}

_LIB = None


class MissingEngineLibrary(RuntimeError):
    pass


def load_library(path: Path | str | None = None) -> C.CDLL:
    """Load the HIP engine library; raises MissingEngineLibrary when absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise MissingEngineLibrary(
            f"{p} is missing: build the HIP engine first (python -m squishy_volumes_amd.build); "
            "the engine has no CPU fallback"
        )
    # One HIP runtime per process: torch's build bundles its own
    # libamdhip64.so.7 under the same soname as ROCm's.  Whichever is loaded
    # first serves both, and torch's GPU initialisation fails on any runtime
    # but its own, so torch (when installed) is loaded before the engine.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def ptr(a) -> C.c_void_p | None:
    """Address of a numpy array or torch tensor (None for None)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return C.c_void_p(a.ctypes.data)
    return C.c_void_p(int(a.data_ptr()))
