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


def _batch_create(sims, **plan):
    p = _native.XfPlan()
    for k, v in plan.items():
        setattr(p, k, v)
    arr = (ctypes.c_void_p * max(1, len(sims)))(*[s.scene_ptr() for s in sims])
    h = ctypes.c_void_p()
    st = _native.hip().xf_batch_create(arr, len(sims), ctypes.byref(p), ctypes.byref(h))
    return st, h


def test_batch_create_validates_without_a_device():
    sims = [scenes.dielectric_sphere(n=n, pml=3, steps=10, tfsf_inset=6, nf2ff_inset=4)
            for n in (12, 15)]
    st, h = _batch_create(sims)
    assert st == 0, _msg()
    if xf.device_count() == 0:
        assert _native.hip().xf_batch_upload(h) == 2 and "device" in _msg()
        assert _native.hip().xf_batch_run(h, 1) == 3
    _native.hip().xf_batch_destroy(h)


def test_batch_rejects_2d_ade_and_empty():
    s3 = scenes.dielectric_sphere(n=12, pml=3, steps=10, tfsf_inset=6, nf2ff_inset=4)
    s2 = scenes.cylinder_2d(n=40, pml=4, steps=10, tfsf_inset=8, nf2ff_inset=6)
    st, _ = _batch_create([s3, s2])
    assert st == 4 and "3-D" in _msg()
    sd = xf.Simulation(3e-3, 3e-3, 3e-3, 0.9)
    sd.add_cube("domain", (0, 0, 0), (0.03,) * 3, xf.Material.air())
    sd.add_sphere("d", (0.015,) * 3, 0.006, xf.Material.drude(1.0, [1e10], [1e9]))
    sd.init(10)
    st, _ = _batch_create([s3, sd])
    assert st == 4 and "ADE" in _msg()
    st, _ = _batch_create([])
    assert st == 1
    st, _ = _batch_create([s3], precision=9)
    assert st == 1 and "precision" in _msg()


def test_batch_and_tuning_null_arguments():
    lib = _native.hip()
    t = _native.XfTuning()
    assert lib.xf_batch_run(None, 1) == 1
    assert lib.xf_batch_set_tuning(None, ctypes.byref(t)) == 1
    assert lib.xf_batch_autotune(None, 1, ctypes.byref(t)) == 1
    assert lib.xf_autotune(None, 1, ctypes.byref(t)) == 1


def test_set_tiling_rejects_more_than_four_rows():
    """The Yee kernels are built for at most 256 threads per block (64 k-lanes
    x 4 rows); 5..8 rows would launch 320..512 threads and fail at launch.
    A tiling set before the upload is kept in the plan (xf_describe shows it;
    upload re-applies the plan)."""
    s = scenes.dielectric_sphere(n=16, pml=4, steps=20, tfsf_inset=6, nf2ff_inset=5)
    st, h = _create(s)
    assert st == 0, _msg()
    lib = _native.hip()
    try:
        for rows in (5, 6, 7, 8, -1):
            assert lib.xf_set_tiling(h, rows, 0, 0) == 1, rows
            assert "tile_y" in _msg()
        assert lib.xf_set_tiling(h, 3, 0, 0) == 0, _msg()
        buf = ctypes.create_string_buffer(512)
        assert lib.xf_describe(h, buf, 512) == 0
        assert "tile_y=3" in buf.value.decode()
    finally:
        lib.xf_destroy(h)


def test_packed_size_overflow_returns_a_status():
    """A grid whose padded field arrays would not fit any device (here
    (2^20 + 1)^3 values of 8 bytes, which also overflows int64 bytes) comes
    back as XF_ERR_NOMEM from xf_create, not as an exception or a crash."""
    s = scenes.dielectric_sphere(n=16, pml=4, steps=20, tfsf_inset=6, nf2ff_inset=5)
    ptr = s.scene_ptr()
    dims = [ctypes.c_int64.from_address(ptr + off) for off in (8, 16, 24)]  # xf_scene nx, ny, nz
    keep = [d.value for d in dims]
    assert keep == [24, 24, 24]
    try:
        for d in dims:
            d.value = 1 << 20
        st, _ = _create(s)
        assert st == 5 and "packed grid too large" in _msg(), (st, _msg())
    finally:
        for d, v in zip(dims, keep):
            d.value = v
    st, h = _create(s)
    assert st == 0, _msg()
    _native.hip().xf_destroy(h)


def test_dispersive_cell_inside_cpml_rejected():
    """The engine applies the CPML E corrections inside the Yee kernels,
    before the ADE kernel, while the reference adds them after the ADE update
    (domain_hd.cu:35-87): a dispersive medium reaching into a CPML slab is
    refused with XF_ERR_UNSUPPORTED instead of being stepped in the wrong
    order.  The same medium kept clear of the slabs is accepted."""
    dl = 3e-3
    for reach, expect in ((0.03, 4), (0.006, 0)):
        s = xf.Simulation(dl, dl, dl, 0.9)
        s.add_cube("domain", (0, 0, 0), (0.03,) * 3, xf.Material.air())
        s.add_pml_all(3)
        c = 0.015
        s.add_cube("drude", (c - reach / 2,) * 3, (reach,) * 3, xf.Material.drude(1.0, [1e10], [1e9]))
        s.init(10)
        st, h = _create(s)
        assert st == expect, (reach, st, _msg())
        if st == 0:
            _native.hip().xf_destroy(h)
        else:
            assert "inside a CPML slab" in _msg()


def test_dispersive_material_without_parameters_rejected():
    """xfh_material kinds 2-5 read eps_inf and the poles from p: a null p or a
    pole count below 1 is an XF_ERR_INVALID, not a null dereference."""
    lib = _native.host()
    s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.9)
    s.add_cube("domain", (0, 0, 0), (0.03,) * 3, xf.Material.air())
    centre = (ctypes.c_double * 3)(0.015, 0.015, 0.015)
    for kind, n_pole, p in ((2, 1, None), (3, 1, None), (4, 0, (ctypes.c_double * 4)(1, 1, 1, 1)),
                            (5, -1, (ctypes.c_double * 6)(1, 1, 1, 1, 1, 1))):
        m = _native.XfhMaterial(kind, 1.0, 1.0, 0.0, 0.0, n_pole, None if p is None else
                                ctypes.cast(p, _native.c_dp))
        st = lib.xfh_sim_add_sphere(s._h, b"bad", centre, ctypes.c_double(0.006), ctypes.byref(m))
        assert st != 0, kind
        assert b"dispersive material" in lib.xfh_last_error(), lib.xfh_last_error()
