"""Loader of the package's native libraries (built in-tree by build.py).

libxfdtd_hip.so   -- the gfx950 engine behind include/xfdtd_hip.h
libxfdtd_host.so  -- scene set-up + the SimulationHD drop-in (include/xfdtd_host.h)

There is no fallback: a missing library raises ImportError with the path it
looked for, so a GPU run can never silently run anything but the HIP path.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(_HERE, "libxfdtd_hip.so")
HOST_LIB = os.path.join(_HERE, "libxfdtd_host.so")

_hip = None
_host = None


class XFError(RuntimeError):
    """A nonzero status from the native ABI, with its last-error message."""


c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_f64 = ctypes.c_double
c_vp = ctypes.c_void_p
c_dp = ctypes.POINTER(ctypes.c_double)


class XfPlan(ctypes.Structure):
    """xf_plan of include/xfdtd_hip.h."""

    _fields_ = [
        ("precision", c_i32),
        ("grid_dim", c_i32 * 3),
        ("block_dim", c_i32 * 3),
        ("use_graph", c_i32),
        ("timing", c_i32),
        ("tile_y", c_i32),
        ("device", c_i32),
        ("chunks_x", c_i32),
        ("tiling_flags", c_i32),
        ("coeff_mode", c_i32),
        ("watchdog", c_i32),
    ]


COEFF_AUTO, COEFF_TABLE, COEFF_DENSE = 0, 1, 2


TILE_XCD_REMAP = 1
TILE_NT = 4
TILE_PML_SPLIT = 8
TILE_K32 = 16
TILE_PML_STREAMS = 32
TILE_PML_ZTAIL = 64
TILE_TFSF_ZTAIL = 128
TILE_SET = 256


class XfTiming(ctypes.Structure):
    """xf_timing of include/xfdtd_hip.h."""

    _fields_ = [
        ("step_ms", c_f64),
        ("yee_h_ms", c_f64),
        ("yee_e_ms", c_f64),
        ("tfsf_ms", c_f64),
        ("nf2ff_ms", c_f64),
        ("ade_ms", c_f64),
        ("yee_h_launches", c_i64),
        ("yee_e_launches", c_i64),
        ("steps", c_i64),
        ("bytes_per_cell_step", c_f64),
        ("run_ms", c_f64),
    ]


class XfTuning(ctypes.Structure):
    """xf_tuning of include/xfdtd_hip.h."""

    _fields_ = [
        ("tile_y", c_i32),
        ("planes", c_i32),
        ("blocks", c_i32),
        ("nontemporal", c_i32),
        ("xcd_remap", c_i32),
        ("pml_split", c_i32),
        ("k32", c_i32),
        ("step_ms", c_f64),
    ]


class XfhMaterial(ctypes.Structure):
    """xfh_material of include/xfdtd_host.h."""

    _fields_ = [
        ("kind", c_i32),
        ("eps_r", c_f64),
        ("mu_r", c_f64),
        ("sigma_e", c_f64),
        ("sigma_m", c_f64),
        ("n_pole", c_i32),
        ("p", c_dp),
    ]


def _load(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python -m xfdtd_cuda_amd.build` first")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def hip() -> ctypes.CDLL:
    """The engine library with its prototypes declared."""
    global _hip
    if _hip is None:
        lib = _load(HIP_LIB)
        lib.xf_last_error.restype = ctypes.c_char_p
        lib.xf_create.argtypes = [c_vp, ctypes.POINTER(XfPlan), ctypes.POINTER(c_vp)]
        lib.xf_upload.argtypes = [c_vp]
        lib.xf_run.argtypes = [c_vp, c_i64]
        lib.xf_download.argtypes = [c_vp]
        lib.xf_destroy.argtypes = [c_vp]
        lib.xf_current_step.argtypes = [c_vp, ctypes.POINTER(c_i64)]
        lib.xf_get_timing.argtypes = [c_vp, ctypes.POINTER(XfTiming)]
        lib.xf_reset_timing.argtypes = [c_vp]
        lib.xf_set_timing.argtypes = [c_vp, c_i32]
        lib.xf_describe.argtypes = [c_vp, ctypes.c_char_p, ctypes.c_size_t]
        lib.xf_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        lib.xf_stream_bench.argtypes = [c_i64, c_i32, c_i32, c_i32, c_i32, ctypes.POINTER(c_f64)]
        lib.xf_stream_bench.restype = ctypes.c_int
        lib.xf_set_tiling.argtypes = [c_vp, c_i32, c_i32, c_i32]
        lib.xf_set_tiling.restype = ctypes.c_int
        pt = ctypes.POINTER(XfTuning)
        batch_sigs = {
            "xf_autotune": [c_vp, c_i32, pt],
            "xf_batch_create": [ctypes.POINTER(c_vp), c_i32, ctypes.POINTER(XfPlan),
                                ctypes.POINTER(c_vp)],
            "xf_batch_upload": [c_vp],
            "xf_batch_run": [c_vp, c_i64],
            "xf_batch_download": [c_vp],
            "xf_batch_destroy": [c_vp],
            "xf_batch_set_tuning": [c_vp, pt],
            "xf_batch_get_tuning": [c_vp, pt],
            "xf_batch_autotune": [c_vp, c_i32, pt],
            "xf_batch_get_timing": [c_vp, ctypes.POINTER(XfTiming)],
            "xf_batch_reset_timing": [c_vp],
            "xf_batch_set_timing": [c_vp, c_i32],
            "xf_batch_describe": [c_vp, ctypes.c_char_p, ctypes.c_size_t],
        }
        for name, args in batch_sigs.items():
            getattr(lib, name).argtypes = args
            getattr(lib, name).restype = ctypes.c_int
        for fn in ("xf_create", "xf_upload", "xf_run", "xf_download", "xf_destroy",
                   "xf_current_step", "xf_get_timing", "xf_reset_timing", "xf_set_timing",
                   "xf_describe",
                   "xf_device_count"):
            getattr(lib, fn).restype = ctypes.c_int
        _hip = lib
    return _hip


def host() -> ctypes.CDLL:
    """The host scene library with its prototypes declared."""
    global _host
    if _host is None:
        hip()  # resolves the engine dependency first
        lib = _load(HOST_LIB)
        lib.xfh_last_error.restype = ctypes.c_char_p
        p3d = ctypes.POINTER(c_f64)
        p3i = ctypes.POINTER(c_i64)
        pm = ctypes.POINTER(XfhMaterial)
        sigs = {
            "xfh_sim_create": [c_f64, c_f64, c_f64, c_f64, ctypes.POINTER(c_vp)],
            "xfh_sim_destroy": [c_vp],
            "xfh_sim_add_cube": [c_vp, ctypes.c_char_p, p3d, p3d, pm],
            "xfh_sim_add_sphere": [c_vp, ctypes.c_char_p, p3d, c_f64, pm],
            "xfh_sim_add_cylinder": [c_vp, ctypes.c_char_p, p3d, c_f64, c_f64, pm],
            "xfh_sim_add_pml": [c_vp, c_i32, c_i64, c_i64, c_f64, c_f64, c_f64, c_f64],
            "xfh_sim_add_tfsf3d": [c_vp, p3i, c_f64, c_f64, c_f64, c_i32, c_f64, c_f64, c_f64],
            "xfh_sim_add_tfsf2d": [c_vp, p3i, c_f64, c_i32, c_f64, c_f64, c_f64],
            "xfh_sim_add_nf2ff_fd": [c_vp, p3i, p3d, c_i32],
            "xfh_sim_set_coefficient_layout": [c_vp, c_i32],
            "xfh_sim_set_cell_sizes": [c_vp, c_i32, c_dp, c_i64],
            "xfh_sim_init": [c_vp, c_i64],
            "xfh_sim_scene": [c_vp, ctypes.POINTER(c_vp)],
            "xfh_sim_info": [c_vp, p3i, ctypes.POINTER(c_f64), ctypes.POINTER(c_i32)],
            "xfh_scene_array": [c_vp, ctypes.c_char_p, ctypes.POINTER(c_dp), p3i,
                                ctypes.POINTER(c_i32)],
            "xfh_sim_run_hd": [c_vp, c_i64, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32), c_i32,
                               c_i32],
            "xfh_sim_far_field": [c_vp, c_i32, c_i32, p3d, c_i64, c_f64, c_f64, c_f64, p3d, p3d,
                                  p3d],
            "xfh_sim_waveform": [c_vp, p3d, p3d, c_i64],
            "xfh_material_permittivity": [pm, c_f64, p3d, p3d],
            "xfh_sim_add_point_source": [c_vp, c_i32, p3i, c_i32, c_f64, c_f64, c_f64, c_f64],
            "xfh_sim_add_learned_boundary": [c_vp, c_i32, c_i32, c_i64,
                                             ctypes.POINTER(ctypes.c_uint16), c_i32, c_i64,
                                             c_i64],
            "xfh_sim_lb_samples": [c_vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_float)),
                                   p3i, p3i],
            "xfh_sim_add_movie_monitor": [c_vp, c_i32, p3d, p3d, c_i64, ctypes.c_char_p],
            "xfh_sim_output_movies": [c_vp, ctypes.POINTER(c_i64)],
            "xfh_sim_add_nf2ff_td": [c_vp, p3i, c_f64, c_f64],
            "xfh_sim_far_field_td": [c_vp, c_i32, ctypes.POINTER(c_i64), p3d, p3d, p3d],
            "xfh_sim_output_far_field": [c_vp, c_i32, p3d, c_i64, p3d, c_i64, ctypes.c_char_p,
                                         ctypes.c_char_p],
            "xfh_sim_output_far_field_td": [c_vp, c_i32, ctypes.c_char_p],
            "xfh_sim_output_waveform": [c_vp, ctypes.c_char_p],
        }
        for name, args in sigs.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _host = lib
    return _host


def check_hip(status: int, what: str) -> None:
    if status != 0:
        msg = hip().xf_last_error().decode(errors="replace")
        raise XFError(f"{what} failed (status {status}): {msg}")


def check_host(status: int, what: str) -> None:
    if status != 0:
        msg = host().xfh_last_error().decode(errors="replace")
        raise XFError(f"{what} failed (status {status}): {msg}")


def device_count() -> int:
    n = ctypes.c_int(0)
    hip().xf_device_count(ctypes.byref(n))
    return int(n.value)


def stream_bench(nbytes: int, reps: int = 10, device: int = 0, variant: int = 0,
                 blocks: int = 0) -> float:
    """Measured device copy bandwidth in GB/s (read + write bytes)."""
    g = ctypes.c_double(0.0)
    check_hip(hip().xf_stream_bench(nbytes, reps, device, variant, blocks, ctypes.byref(g)),
              "xf_stream_bench")
    return g.value
