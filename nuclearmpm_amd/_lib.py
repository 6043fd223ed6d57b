"""ctypes binding of libnclr_mpm.so (the HIP engine's C ABI, include/nclr_mpm.h).

The library is built in-tree by __graft_entry__.build() (or `make -C
nuclearmpm_amd/csrc`).  Loading fails loudly when it is missing: there is no
CPU fallback in this package.
"""
from __future__ import annotations

import ctypes
import os

try:  # load torch's HIP runtime first so the engine shares it (same SONAME libamdhip64.so.7)
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is optional for the plain engine
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnclr_mpm.so")

MPM_OK = 0
MPM_ERR_INVALID = 1
MPM_ERR_OUT_OF_RANGE = 2
MPM_ERR_HIP = 3
MPM_ERR_UNSUPPORTED = 4


class mpm_options(ctypes.Structure):
    _fields_ = [
        ("identity_mode", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("timing", ctypes.c_int32),
        ("p2g_chunk", ctypes.c_int32),
        ("g2p_threads", ctypes.c_int32),
        ("graph", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
        ("fast_svd", ctypes.c_int32),
        ("compact", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 7),
    ]


class mpm_scene(ctypes.Structure):
    _fields_ = [
        ("material", ctypes.c_int32),
        ("res", ctypes.c_int32),
        ("dt", ctypes.c_float),
        ("E", ctypes.c_float),
        ("nu", ctypes.c_float),
        ("gravity", ctypes.c_float),
        ("n", ctypes.c_int64),
    ]


class mpm_cube(ctypes.Structure):
    _fields_ = [
        ("res", ctypes.c_int32 * 3),
        ("lo", ctypes.c_float * 3),
        ("hi", ctypes.c_float * 3),
        ("v", ctypes.c_float * 3),
    ]


class EngineError(RuntimeError):
    """Error reported by the HIP engine."""


class EngineMissing(ImportError):
    """libnclr_mpm.so is not built."""


_lib = None

_c_float_p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineMissing(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C nuclearmpm_amd/csrc")
    L = ctypes.CDLL(LIB_PATH)
    i32, i64, f32 = ctypes.c_int, ctypes.c_int64, ctypes.c_float
    H = ctypes.c_void_p
    L.mpm_default_options.argtypes = [ctypes.POINTER(mpm_options)]
    L.mpm_default_options.restype = None
    L.mpm_create.argtypes = [i32, i32, i32, f32, f32, f32, f32, i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                             ctypes.POINTER(mpm_options), ctypes.POINTER(H)]
    L.mpm_create.restype = i32
    L.mpm_create_batch.argtypes = [i32, i32, ctypes.POINTER(mpm_scene), _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                   ctypes.POINTER(mpm_options), ctypes.POINTER(H)]
    L.mpm_create_batch.restype = i32
    L.mpm_get_grid_scene.argtypes = [H, i32, _vp, _vp]
    L.mpm_get_grid_scene.restype = i32
    L.mpm_get_scene_params.argtypes = [H, i32, _c_float_p, _c_float_p, _c_float_p, _c_float_p]
    L.mpm_get_scene_params.restype = i32
    L.mpm_num_scenes.argtypes = [H]
    L.mpm_num_scenes.restype = i32
    L.mpm_set_tuning.argtypes = [H, i32, i32]
    L.mpm_set_tuning.restype = i32
    L.mpm_get_tuning.argtypes = [H, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.mpm_get_tuning.restype = i32
    L.mpm_set_compact.argtypes = [H, i32]
    L.mpm_set_compact.restype = i32
    L.mpm_get_compact.argtypes = [H, ctypes.POINTER(i32)]
    L.mpm_get_compact.restype = i32
    L.mpm_autotune.argtypes = [H, i32]
    L.mpm_autotune.restype = i32
    L.mpm_active_nodes_scene.argtypes = [H, i32, ctypes.POINTER(i64)]
    L.mpm_active_nodes_scene.restype = i32
    L.mpm_advance.argtypes = [H, i32]
    L.mpm_advance.restype = i32
    L.mpm_synchronize.argtypes = [H]
    L.mpm_synchronize.restype = i32
    L.mpm_scene_status.argtypes = [H, _vp, _vp]
    L.mpm_scene_status.restype = i32
    L.mpm_get_particles.argtypes = [H, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mpm_get_particles.restype = i32
    L.mpm_set_particles.argtypes = [H, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.mpm_set_particles.restype = i32
    L.mpm_get_grid.argtypes = [H, _vp, _vp]
    L.mpm_get_grid.restype = i32
    L.mpm_get_params.argtypes = [H, _c_float_p, _c_float_p, _c_float_p, _c_float_p]
    L.mpm_get_params.restype = i32
    L.mpm_num_particles.argtypes = [H]
    L.mpm_num_particles.restype = i64
    L.mpm_num_nodes.argtypes = [H]
    L.mpm_num_nodes.restype = i64
    L.mpm_steps_taken.argtypes = [H]
    L.mpm_steps_taken.restype = i64
    L.mpm_kernel_time.argtypes = [H, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]
    L.mpm_kernel_time.restype = i32
    L.mpm_reset_kernel_times.argtypes = [H]
    L.mpm_reset_kernel_times.restype = i32
    L.mpm_set_timing.argtypes = [H, i32]
    L.mpm_set_timing.restype = i32
    L.mpm_active_nodes.argtypes = [H, ctypes.POINTER(i64)]
    L.mpm_active_nodes.restype = i32
    L.mpm_sort_keys.argtypes = [_vp, i64, i32, _vp, i32]
    L.mpm_sort_keys.restype = i32
    L.mpm_sort_keys_noalloc.argtypes = [_vp, i64, i32, _vp, _vp, i32]
    L.mpm_sort_keys_noalloc.restype = i32
    L.mpm_svd.argtypes = [i32, i64, _vp, _vp, _vp, _vp, i32, i32]
    L.mpm_svd.restype = i32
    L.mpm_grad_channels.argtypes = [i32]
    L.mpm_grad_channels.restype = i32
    L.mpm_state_bytes.argtypes = [H]
    L.mpm_state_bytes.restype = i64
    for name in ("mpm_save_state", "mpm_load_state", "mpm_get_state_device", "mpm_set_state_device"):
        getattr(L, name).argtypes = [H, _vp]
        getattr(L, name).restype = i32
    for name in ("mpm_grad_from_particle_order", "mpm_grad_to_particle_order"):
        getattr(L, name).argtypes = [H, _vp, _vp]
        getattr(L, name).restype = i32
    L.mpm_backward_step.argtypes = [H, _vp, _vp, _vp]
    L.mpm_backward_step.restype = i32
    L.mpm_backward_segment.argtypes = [H, i32, _vp, _vp, _vp]
    L.mpm_backward_segment.restype = i32
    L.mpm_tape_bytes.argtypes = [H, i32]
    L.mpm_tape_bytes.restype = i64
    L.mpm_record_bytes.argtypes = [H, i32, i32]
    L.mpm_record_bytes.restype = i64
    L.mpm_advance_recorded.argtypes = [H, i32, _vp, i32]
    L.mpm_advance_recorded.restype = i32
    L.mpm_backward_recorded.argtypes = [H, i32, _vp, i32, _vp, _vp, _vp]
    L.mpm_backward_recorded.restype = i32
    L.mpm_set_scene_elasticity.argtypes = [H, i32, f32, f32]
    L.mpm_set_scene_elasticity.restype = i32
    u32x4 = ctypes.POINTER(ctypes.c_uint32)
    L.mpm_gen_cubes.argtypes = [i32, i32, ctypes.POINTER(mpm_cube), i32, i64, _vp, i32]
    L.mpm_gen_cubes.restype = i32
    L.mpm_rand_advance.argtypes = [u32x4, ctypes.c_uint64]
    L.mpm_rand_advance.restype = i32
    L.mpm_gen_rand.argtypes = [i64, u32x4, _vp, i32]
    L.mpm_gen_rand.restype = i32
    L.mpm_gen_sample.argtypes = [i32, i32, _vp, i32, f32, u32x4, i32, _vp, i32]
    L.mpm_gen_sample.restype = i32
    L.mpm_gyroid.argtypes = [i64, _vp, f32, f32, _vp, i32]
    L.mpm_gyroid.restype = i32
    L.mpm_gen_gyroid.argtypes = [ctypes.POINTER(mpm_cube), f32, f32, f32, i32, _vp, i64, ctypes.POINTER(i64), i32]
    L.mpm_gen_gyroid.restype = i32
    L.mpm_project_points.argtypes = [i64, _vp, i32, _vp, i32]
    L.mpm_project_points.restype = i32
    L.mpm_destroy.argtypes = [H]
    L.mpm_destroy.restype = i32
    L.mpm_last_error.argtypes = []
    L.mpm_last_error.restype = ctypes.c_char_p
    L.mpm_version.argtypes = []
    L.mpm_version.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc: int) -> None:
    """Raise the Python counterpart of an MPM_ERR_* code (IndexError for out_of_range,
    like the reference's std::out_of_range from std::vector::at)."""
    if rc == MPM_OK:
        return
    msg = lib().mpm_last_error().decode()
    if rc == MPM_ERR_OUT_OF_RANGE:
        raise IndexError(msg)
    if rc == MPM_ERR_INVALID:
        raise ValueError(msg)
    if rc == MPM_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise EngineError(msg)


def sort_keys(keys, bits: int, device: int = 0):
    """The engine's stable GPU counting sort: permutation ordering `keys` (uint32) by their low
    `bits` bits, ties in input order."""
    import numpy as np

    k = np.ascontiguousarray(keys, dtype=np.uint32)
    perm = np.empty(k.shape[0], np.uint32)
    check(lib().mpm_sort_keys(ptr(k), k.shape[0], int(bits), ptr(perm), int(device)))
    return perm


def ptr(a):
    """Address of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)
