"""CPU oracle of the FDTD time step (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
this module; the product (xfdtd_cuda_amd) never does.  It wraps
oracle/libxf_oracle.so (built from xf_oracle.c by oracle/Makefile), which
advances an xf_scene in place on the host with the reference's algorithm.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "libxf_oracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError(f"{LIB} is not built (make -C oracle)")
        l = ctypes.CDLL(LIB)
        l.xfo_run.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        l.xfo_run.restype = ctypes.c_int
        l.xfo_last_error.restype = ctypes.c_char_p
        l.xfo_num_threads.restype = ctypes.c_int
        l.xfo_set_num_threads.argtypes = [ctypes.c_int]
        l.xfo_yee_component_triple_loop.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        l.xfo_yee_component_triple_loop.restype = ctypes.c_int
        l.xfo_yee_component.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        l.xfo_yee_component.restype = ctypes.c_int
        l.xfo_lb_predict.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_void_p]
        l.xfo_lb_predict.restype = ctypes.c_int
        l.xfo_lb_mfma16.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int64, ctypes.c_void_p]
        l.xfo_lb_mfma16.restype = ctypes.c_int
        _lib = l
    return _lib


def run(scene_ptr: int, n_steps: int) -> None:
    """Advance the scene n_steps on the CPU (in place)."""
    st = lib().xfo_run(ctypes.c_void_p(scene_ptr), n_steps)
    if st != 0:
        raise RuntimeError("oracle: " + lib().xfo_last_error().decode())


_native_lib = None


def native_lib() -> ctypes.CDLL:
    """The oracle built with -march=native for this host (make -C oracle
    native, run here): the CPU baseline's timing build.  Same C source and
    evaluation order as lib()."""
    global _native_lib
    if _native_lib is None:
        subprocess.run(["make", "-C", _HERE, "-s", "-B", "native"], check=True)
        l = ctypes.CDLL(os.path.join(_HERE, "libxf_oracle_native.so"))
        l.xfo_run.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        l.xfo_run.restype = ctypes.c_int
        l.xfo_last_error.restype = ctypes.c_char_p
        l.xfo_num_threads.restype = ctypes.c_int
        l.xfo_set_num_threads.argtypes = [ctypes.c_int]
        _native_lib = l
    return _native_lib


def time_steps(scene_ptr: int, n_steps: int, threads: int, repeats: int = 3):
    """Wall time of n_steps oracle steps on the native build with `threads`
    OpenMP threads, `repeats` windows one after another (the scene advances);
    returns the list of seconds per window."""
    import time
    l = native_lib()
    keep = l.xfo_num_threads()
    l.xfo_set_num_threads(threads)
    out = []
    try:
        for _ in range(repeats):
            t0 = time.perf_counter()
            st = l.xfo_run(ctypes.c_void_p(scene_ptr), n_steps)
            out.append(time.perf_counter() - t0)
            if st != 0:
                raise RuntimeError("oracle: " + l.xfo_last_error().decode())
    finally:
        l.xfo_set_num_threads(keep)
    return out


def set_threads(n: int) -> None:
    lib().xfo_set_num_threads(n)


def threads() -> int:
    return lib().xfo_num_threads()


def yee_component_triple_loop(scene_ptr: int, is_e: bool, axis: int) -> None:
    lib().xfo_yee_component_triple_loop(ctypes.c_void_p(scene_ptr), int(is_e), axis)


def yee_component(scene_ptr: int, is_e: bool, axis: int) -> None:
    lib().xfo_yee_component(ctypes.c_void_p(scene_ptr), int(is_e), axis)


def lb_predict(weights_u16, mur_k: float, x) -> np.ndarray:
    """The learned layer's column predictor (C restatement) on n x 12 inputs."""
    w = np.ascontiguousarray(weights_u16, dtype=np.uint16)
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape[0], np.float32)
    lib().xfo_lb_predict(w.ctypes.data, mur_k, x.ctypes.data, x.shape[0], out.ctypes.data)
    return out


def lb_mfma16(c, a, b) -> np.ndarray:
    """The predictor's 16-term product model: c + sum_k a[:, k] b[:, k] as two
    exact groups of 8 (k-slots 0..7, then 8..15), one fp32 rounding each."""
    c = np.ascontiguousarray(c, dtype=np.float32)
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    out = np.empty(c.shape[0], np.float32)
    lib().xfo_lb_mfma16(c.ctypes.data, a.ctypes.data, b.ctypes.data, c.shape[0], out.ctypes.data)
    return out
