"""Python view of the scene model and the SimulationHD drop-in.

The names follow the reference's drivers (test/dielectric_sphere_scatter.cu,
test/benchmark.cu, test/cylinder_scatter_2d.cu): a Simulation is built from
objects, PML boundaries, a TFSF source and NF2FF boxes, then run through
SimulationHD(sim).setGridDim/setBlockDim/run(n) (include/xfdtd_cuda/
simulation/simulation_hd.cuh:37-88).  Everything below calls the C ABIs of
include/xfdtd_host.h and include/xfdtd_hip.h.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Iterable, Sequence

import numpy as np

from . import _native
from ._native import XfhMaterial, XfPlan, XfTiming, XfTuning, c_dp, c_i32, c_i64, c_vp

DIRECTIONS = {"XN": 0, "XP": 1, "YN": 2, "YP": 3, "ZN": 4, "ZP": 5}
FACES = ("XN", "XP", "YN", "YP", "ZN", "ZP")
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2, "bf16c": 3, "fp16": 4}

EPSILON_0 = 8.854187817e-12
MU_0 = 4.0 * np.pi * 1e-7
C_0 = 1.0 / np.sqrt(EPSILON_0 * MU_0)
ETA_0 = np.sqrt(MU_0 / EPSILON_0)


@dataclass
class Material:
    """xfh_material: kind 0 plain, 1 PEC, 2 Drude, 3 Debye, 4 Lorentz, 5 M-Lorentz."""

    kind: int = 0
    eps_r: float = 1.0
    mu_r: float = 1.0
    sigma_e: float = 0.0
    sigma_m: float = 0.0
    poles: list = field(default_factory=list)  # flat pole numbers (see xfdtd_host.h)
    n_pole: int = 0

    @staticmethod
    def air() -> "Material":
        return Material()

    @staticmethod
    def pec() -> "Material":
        return Material(kind=1)

    @staticmethod
    def dielectric(eps_r: float, mu_r: float = 1.0, sigma_e: float = 0.0,
                   sigma_m: float = 0.0) -> "Material":
        return Material(0, eps_r, mu_r, sigma_e, sigma_m)

    @staticmethod
    def drude(eps_inf: float, omega_p: Sequence[float], gamma: Sequence[float]) -> "Material":
        p = [eps_inf]
        for w, g in zip(omega_p, gamma):
            p += [w, g]
        return Material(kind=2, eps_r=eps_inf, poles=p, n_pole=len(omega_p))

    @staticmethod
    def debye(eps_inf: float, delta_eps: Sequence[float], tau: Sequence[float]) -> "Material":
        p = [eps_inf]
        for d, t in zip(delta_eps, tau):
            p += [d, t]
        return Material(kind=3, eps_r=eps_inf, poles=p, n_pole=len(delta_eps))

    @staticmethod
    def lorentz(eps_inf: float, delta_eps: Sequence[float], omega_0: Sequence[float],
                delta: Sequence[float]) -> "Material":
        p = [eps_inf]
        for d, w, g in zip(delta_eps, omega_0, delta):
            p += [d, w, g]
        return Material(kind=4, eps_r=eps_inf, poles=p, n_pole=len(delta_eps))

    @staticmethod
    def mlorentz(eps_inf: float, poles: Iterable[Sequence[float]]) -> "Material":
        """poles: (a0, a1, b0, b1, b2) per pole, chi = (a0 + a1 jw)/(b0 + b1 jw + b2 (jw)^2)."""
        p = [eps_inf]
        n = 0
        for q in poles:
            p += list(q)
            n += 1
        return Material(kind=5, eps_r=eps_inf, poles=p, n_pole=n)

    def _c(self):
        arr = (ctypes.c_double * max(1, len(self.poles)))(*self.poles) if self.poles else None
        m = XfhMaterial(self.kind, self.eps_r, self.mu_r, self.sigma_e, self.sigma_m,
                        self.n_pole, ctypes.cast(arr, c_dp) if arr is not None else c_dp())
        return m, arr

    def permittivity(self, omega: float) -> complex:
        m, keep = self._c()
        re, im = ctypes.c_double(), ctypes.c_double()
        _native.check_host(_native.host().xfh_material_permittivity(
            ctypes.byref(m), omega, ctypes.byref(re), ctypes.byref(im)), "permittivity")
        del keep
        return complex(re.value, im.value)


@dataclass
class Waveform:
    """Waveform::gaussian(tau, t0) / cosineModulatedGaussian / sine."""

    kind: int = 0
    tau: float = 0.0
    t0: float = 0.0
    freq: float = 0.0

    @staticmethod
    def gaussian(tau: float, t0: float) -> "Waveform":
        return Waveform(0, tau, t0, 0.0)

    @staticmethod
    def cosine_modulated_gaussian(tau: float, t0: float, freq: float) -> "Waveform":
        return Waveform(1, tau, t0, freq)

    @staticmethod
    def sine(freq: float) -> "Waveform":
        return Waveform(2, 0.0, 0.0, freq)

    def __call__(self, t):
        t = np.asarray(t, dtype=np.float64)
        if self.kind == 0:
            return np.exp(-((t - self.t0) / self.tau) ** 2)
        if self.kind == 1:
            return np.exp(-((t - self.t0) / self.tau) ** 2) * np.cos(
                2 * np.pi * self.freq * (t - self.t0))
        return np.sin(2 * np.pi * self.freq * t)


def _face_mask(faces) -> int:
    m = 0
    for f in faces:
        m |= 1 << FACES.index(f)
    return m


def _d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def _i3(v):
    v = list(v) + [0] * (3 - len(v))
    return (ctypes.c_int64 * 3)(*[int(x) for x in v])


class Simulation:
    """xfdtd::Simulation{dx, dy, dz, cfl} (include/xfdtd_cuda_amd/xfdtd_core.h)."""

    def __init__(self, dx: float, dy: float, dz: float, cfl: float):
        self._lib = _native.host()
        h = c_vp()
        _native.check_host(self._lib.xfh_sim_create(dx, dy, dz, cfl, ctypes.byref(h)),
                           "Simulation")
        self._h = h
        self.time_step = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.xfh_sim_destroy(h)
            self._h = None

    # ---- objects
    def add_cube(self, name: str, origin, size, material: Material | None = None):
        m, keep = (material or Material.air())._c()
        _native.check_host(self._lib.xfh_sim_add_cube(self._h, name.encode(), _d3(origin),
                                                      _d3(size), ctypes.byref(m)), "addObject")
        del keep
        return self

    def add_sphere(self, name: str, center, radius: float, material: Material):
        m, keep = material._c()
        _native.check_host(self._lib.xfh_sim_add_sphere(self._h, name.encode(), _d3(center),
                                                        radius, ctypes.byref(m)), "addObject")
        del keep
        return self

    def add_cylinder(self, name: str, center, radius: float, height: float,
                     material: Material):
        m, keep = material._c()
        _native.check_host(self._lib.xfh_sim_add_cylinder(
            self._h, name.encode(), _d3(center), radius, height, ctypes.byref(m)), "addObject")
        del keep
        return self

    # ---- boundaries, sources, monitors
    def add_pml(self, thickness: int, direction: str, order: int = 3, sigma_ratio: float = 1.0,
                alpha_min: float = 0.0, alpha_max: float = 0.05, kappa_max: float = 1.0):
        _native.check_host(self._lib.xfh_sim_add_pml(
            self._h, DIRECTIONS[direction], thickness, order, sigma_ratio, alpha_min, alpha_max,
            kappa_max), "addBoundary")
        return self

    def add_pml_all(self, thickness: int, faces: Iterable[str] = FACES, **kw):
        for f in faces:
            self.add_pml(thickness, f, **kw)
        return self

    def add_tfsf3d(self, start, theta: float, phi: float, psi: float, waveform: Waveform):
        _native.check_host(self._lib.xfh_sim_add_tfsf3d(
            self._h, _i3(start), theta, phi, psi, waveform.kind, waveform.tau, waveform.t0,
            waveform.freq), "addWaveformSource")
        self.waveform = waveform
        return self

    def add_tfsf2d(self, start, phi: float, waveform: Waveform):
        _native.check_host(self._lib.xfh_sim_add_tfsf2d(
            self._h, _i3(start), phi, waveform.kind, waveform.tau, waveform.t0,
            waveform.freq), "addWaveformSource")
        self.waveform = waveform
        return self

    def add_nf2ff_fd(self, start, freqs: Sequence[float]):
        f = (ctypes.c_double * len(freqs))(*freqs)
        _native.check_host(self._lib.xfh_sim_add_nf2ff_fd(self._h, _i3(start), f, len(freqs)),
                           "addNF2FF")
        return self

    def add_nf2ff_td(self, start, theta: float, phi: float):
        """NF2FF time-domain box (NFFFTTimeDomain) at the inset `start` (cells,
        as add_nf2ff_fd), observed along (theta, phi)."""
        _native.check_host(self._lib.xfh_sim_add_nf2ff_td(self._h, _i3(start), theta, phi),
                           "addNF2FF(time domain)")
        return self

    def add_point_source(self, comp: int, index, waveform: Waveform, amplitude: float = 1.0):
        """Soft source on E_comp (0 x, 1 y, 2 z) at grid index `index` of that
        component's array (no counterpart in the reference's GPU path)."""
        _native.check_host(self._lib.xfh_sim_add_point_source(
            self._h, comp, _i3(index), waveform.kind, waveform.tau, waveform.t0, waveform.freq,
            amplitude), "addPointSource")
        return self

    def add_learned_boundary(self, weights, faces=FACES, thickness: int = 2):
        """The learned absorbing layer: `thickness` cells outside the domain on
        each face, outer node driven by the column MLP (XF_LB_NPARAM fp16
        weights as uint16 bits or a float array rounded to fp16)."""
        w = np.asarray(weights)
        if w.dtype != np.uint16:
            w = w.astype(np.float16).view(np.uint16)
        w = np.ascontiguousarray(w)
        self._lb_weights = w
        _native.check_host(self._lib.xfh_sim_add_learned_boundary(
            self._h, 1, _face_mask(faces), thickness,
            w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), w.size, 0, 0),
            "addLearnedBoundary")
        return self

    def add_boundary_recorder(self, inset: int, every: int, slots: int, faces=FACES):
        """Records learned-layer training samples on the plane `inset` nodes
        inside each face (no field is changed)."""
        _native.check_host(self._lib.xfh_sim_add_learned_boundary(
            self._h, 2, _face_mask(faces), inset, None, 0, every, slots), "addBoundaryRecorder")
        return self

    def lb_samples(self) -> np.ndarray:
        """Recorded samples, (recorded steps, columns, 13) float32 (a copy)."""
        p = ctypes.POINTER(ctypes.c_float)()
        cols = ctypes.c_int64()
        rec = ctypes.c_int64()
        _native.check_host(self._lib.xfh_sim_lb_samples(self._h, ctypes.byref(p),
                                                        ctypes.byref(cols), ctypes.byref(rec)),
                           "lb_samples")
        n = rec.value * cols.value * 13
        if n == 0:
            return np.zeros((0, cols.value, 13), np.float32)
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(rec.value, cols.value, 13).copy()

    def add_movie_monitor(self, field: str, lo, size, interval: int, output_dir: str = ""):
        """MovieMonitor(FieldMonitor(Cube(lo, size), field), interval): a frame
        of the component over the cube every `interval` steps (reference
        src/monitor/movie_monitor.cu); field is one of ex..hz."""
        f = ("ex", "ey", "ez", "hx", "hy", "hz").index(field.lower())
        _native.check_host(self._lib.xfh_sim_add_movie_monitor(
            self._h, f, _d3(lo), _d3(size), interval, output_dir.encode()), "addMonitor")
        return self

    def movie(self, index: int = 0) -> np.ndarray:
        """Frames of movie monitor `index`: (frames, box extents), a view."""
        return self.array(f"movie.{index}.data")

    def output_movies(self) -> int:
        """Writes every movie monitor's frames as %05d.npy; returns the count."""
        n = c_i64()
        _native.check_host(self._lib.xfh_sim_output_movies(self._h, ctypes.byref(n)), "output")
        return n.value

    def set_coefficient_layout(self, layout: str):
        _native.check_host(self._lib.xfh_sim_set_coefficient_layout(
            self._h, {"dense": 0, "compact": 1}[layout]), "setCoefficientLayout")
        return self

    def set_cell_sizes(self, axis: int, sizes):
        """A graded axis (0 x, 1 y, 2 z): the sizes of the domain's cells
        along it from its lower end; PML cells keep the based size."""
        v = np.ascontiguousarray(sizes, dtype=np.float64)
        _native.check_host(self._lib.xfh_sim_set_cell_sizes(
            self._h, axis, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), v.size), "setCellSizes")
        return self

    # ---- set-up and state
    def init(self, time_step: int):
        _native.check_host(self._lib.xfh_sim_init(self._h, time_step), "init")
        self.time_step = time_step
        return self

    def scene_ptr(self) -> int:
        p = c_vp()
        _native.check_host(self._lib.xfh_sim_scene(self._h, ctypes.byref(p)), "scene")
        return p.value

    def info(self):
        dims = (ctypes.c_int64 * 3)()
        dt = ctypes.c_double()
        dim = ctypes.c_int32()
        _native.check_host(self._lib.xfh_sim_info(self._h, dims, ctypes.byref(dt),
                                                  ctypes.byref(dim)), "info")
        return tuple(dims), dt.value, dim.value

    @property
    def shape(self):
        return self.info()[0]

    @property
    def dt(self):
        return self.info()[1]

    def array(self, name: str) -> np.ndarray:
        """A numpy view (no copy) of a named scene array (see xfh_scene_array)."""
        return scene_array(self.scene_ptr(), name)

    def nf2ff_current(self, box: int, freq: int, face: str, comp: str) -> np.ndarray:
        a = self.array(f"nf2ff.{box}.{freq}.{face}.{comp}")
        return a[..., 0] + 1j * a[..., 1]

    def waveform_samples(self):
        n = self.time_step
        t = np.zeros(n)
        v = np.zeros(n)
        _native.check_host(self._lib.xfh_sim_waveform(
            self._h, t.ctypes.data_as(c_dp), v.ctypes.data_as(c_dp), n), "waveform")
        return t, v

    def incident_spectrum(self, freq: float) -> complex:
        """DFT of the source waveform with the NF2FF E phasor convention."""
        t, v = self.waveform_samples()
        dt = self.dt
        return complex(np.sum(v * dt * np.exp(-2j * np.pi * freq * t)))

    def far_field(self, box: int, freq_index: int, theta: np.ndarray, phi: float,
                  incident: complex = 0j):
        """processFarField: E_theta, E_phi (normalised by `incident` if nonzero), RCS."""
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        n = theta.size
        et = np.zeros(2 * n)
        ep = np.zeros(2 * n)
        rcs = np.zeros(n)
        _native.check_host(self._lib.xfh_sim_far_field(
            self._h, box, freq_index, theta.ctypes.data_as(c_dp), n, phi, incident.real,
            incident.imag, et.ctypes.data_as(c_dp), ep.ctypes.data_as(c_dp),
            rcs.ctypes.data_as(c_dp)), "processFarField")
        return et[0::2] + 1j * et[1::2], ep[0::2] + 1j * ep[1::2], rcs


    def far_field_td(self, box: int = 0):
        """NFFFTTimeDomain::processFarField of time-domain box `box`: the time
        axis t_m = m dt and R E_theta(t_m), R E_phi(t_m)."""
        n = ctypes.c_int64()
        _native.check_host(self._lib.xfh_sim_far_field_td(self._h, box, ctypes.byref(n), None,
                                                           None, None), "processFarField(TD)")
        t, et, ep = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
        _native.check_host(self._lib.xfh_sim_far_field_td(
            self._h, box, ctypes.byref(n), t.ctypes.data_as(c_dp), et.ctypes.data_as(c_dp),
            ep.ctypes.data_as(c_dp)), "processFarField(TD)")
        return t, et, ep


    def output_far_field(self, box: int, theta, phi, sub_dir: str, output_dir: str | None = None):
        """NFFFTFrequencyDomain::processFarField(theta, phi, sub_dir) of the
        drivers: one of theta, phi is a scalar; writes theta.npy, phi.npy,
        freq.npy, e_theta.npy, e_phi.npy (complex, frequencies x directions)
        under <output_dir>/<sub_dir>/."""
        th = np.atleast_1d(np.ascontiguousarray(theta, dtype=np.float64))
        ph = np.atleast_1d(np.ascontiguousarray(phi, dtype=np.float64))
        _native.check_host(self._lib.xfh_sim_output_far_field(
            self._h, box, th.ctypes.data_as(c_dp), th.size, ph.ctypes.data_as(c_dp), ph.size,
            sub_dir.encode(), None if output_dir is None else str(output_dir).encode()),
            "processFarField")
        return self

    def output_far_field_td(self, box: int = 0, output_dir: str | None = None):
        """NFFFTTimeDomain::processFarField(): time.npy, e_theta.npy, e_phi.npy."""
        _native.check_host(self._lib.xfh_sim_output_far_field_td(
            self._h, box, None if output_dir is None else str(output_dir).encode()),
            "processFarField(TD)")
        return self

    def output_waveform(self, output_dir: str):
        """time.npy and incident_wave.npy of the TFSF waveform, as the drivers
        dump them."""
        _native.check_host(self._lib.xfh_sim_output_waveform(self._h, str(output_dir).encode()),
                           "output_waveform")
        return self


def scene_array(scene_ptr: int, name: str) -> np.ndarray:
    data = c_dp()
    shape = (ctypes.c_int64 * 5)()
    ndim = ctypes.c_int32()
    _native.check_host(_native.host().xfh_scene_array(
        ctypes.c_void_p(scene_ptr), name.encode(), ctypes.byref(data), shape, ctypes.byref(ndim)),
        f"array {name}")
    shp = tuple(int(shape[d]) for d in range(ndim.value))
    if not data:
        return np.zeros((0,))  # array absent from this scene (e.g. compact layout)
    return np.ctypeslib.as_array(data, shape=shp)


class SimulationHD:
    """xfdtd::cuda::SimulationHD(sim): setGridDim, setBlockDim, run(n).

    run() goes through the C++ drop-in class (include/xfdtd_cuda_amd/
    simulation_hd.h): init(n), copy to device, the DomainHD step loop on the
    GPU, copy back into the simulation's arrays.
    """

    def __init__(self, host: Simulation):
        self._sim = host
        self._grid = (1, 1, 1)
        self._block = (1, 1, 1)
        self.precision = "fp64"
        self.use_graph = False

    def host(self) -> Simulation:
        return self._sim

    def setGridDim(self, g):  # noqa: N802 (reference name)
        self._grid = tuple(int(x) for x in g)

    def setBlockDim(self, b):  # noqa: N802
        self._block = tuple(int(x) for x in b)

    def run(self, time_step: int):
        g = (c_i32 * 3)(*self._grid)
        b = (c_i32 * 3)(*self._block)
        _native.check_host(self._sim._lib.xfh_sim_run_hd(
            self._sim._h, time_step, g, b, PRECISIONS[self.precision], int(self.use_graph)),
            "SimulationHD::run")
        self._sim.time_step = time_step


class Engine:
    """Direct handle on the engine C ABI (xf_create / xf_upload / xf_run /
    xf_download), for callers that want the plan knobs and timings."""

    def __init__(self, sim: Simulation, precision: str = "fp64", use_graph: bool = False,
                 timing: bool = False, tile_y: int = 0, device: int = 0, coeff: str = "auto",
                 watchdog: int = 0):
        """coeff: "auto" (a table of distinct coefficient rows when there
        are at most 65536 per pass, else the dense per-cell arrays),
        "table" or "dense" (the scene's 18 coefficient arrays streamed)."""
        self._lib = _native.hip()
        self._sim = sim
        plan = XfPlan()
        plan.precision = PRECISIONS[precision]
        plan.use_graph = int(use_graph)
        plan.timing = int(timing)
        plan.tile_y = tile_y
        plan.device = device
        plan.coeff_mode = {"auto": _native.COEFF_AUTO, "table": _native.COEFF_TABLE,
                           "dense": _native.COEFF_DENSE}[coeff]
        # every `watchdog` steps the fields are checked for NaN/Inf on the
        # device; run() then raises XFError (status XF_ERR_NUMERIC)
        plan.watchdog = watchdog
        self._plan = plan
        h = c_vp()
        _native.check_hip(self._lib.xf_create(ctypes.c_void_p(sim.scene_ptr()),
                                              ctypes.byref(plan), ctypes.byref(h)), "xf_create")
        self._h = h

    def upload(self):
        _native.check_hip(self._lib.xf_upload(self._h), "xf_upload")
        return self

    def run(self, n: int):
        _native.check_hip(self._lib.xf_run(self._h, n), "xf_run")
        return self

    def download(self):
        _native.check_hip(self._lib.xf_download(self._h), "xf_download")
        return self

    def timing(self) -> dict:
        t = XfTiming()
        _native.check_hip(self._lib.xf_get_timing(self._h, ctypes.byref(t)), "xf_get_timing")
        return {k: getattr(t, k) for k, _ in XfTiming._fields_}

    def reset_timing(self):
        _native.check_hip(self._lib.xf_reset_timing(self._h), "xf_reset_timing")

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(512)
        _native.check_hip(self._lib.xf_describe(self._h, buf, 512), "xf_describe")
        return buf.value.decode()

    def set_tiling(self, tile_y: int = 0, chunks_x: int = 0, xcd_remap: bool | None = None,
                   nontemporal: bool | None = None, pml_split: bool | None = None,
                   k32: bool | None = None, pml_streams: bool | None = None,
                   pml_ztail: bool | None = None, tfsf_ztail: bool | None = None):
        """Tiling of the 3-D Yee kernels; None leaves a flag to the engine's
        default (no XCD remap, non-temporal streams on, z-face CPML split
        off).  pml_split moves the z-face CPML corrections into a launch of
        their own, pml_ztail into the end of each Yee tile (FP64/FP32; the
        results are the same either way; the tail takes precedence);
        tfsf_ztail, with pml_ztail, has the tiles apply the TF/SF
        corrections of the z = const faces too; k32 tiles k by 32 lanes with
        two j-rows per wave."""
        flags = 0
        if xcd_remap is not None or nontemporal is not None or pml_split is not None or \
                k32 is not None or pml_streams is not None or pml_ztail is not None or \
                tfsf_ztail is not None:
            flags = _native.TILE_SET
            flags |= _native.TILE_XCD_REMAP if xcd_remap else 0
            flags |= _native.TILE_PML_STREAMS if pml_streams else 0
            flags |= _native.TILE_NT if (nontemporal is None or nontemporal) else 0
            flags |= _native.TILE_PML_SPLIT if (pml_split is None or pml_split) else 0
            flags |= _native.TILE_PML_ZTAIL if pml_ztail else 0
            flags |= _native.TILE_TFSF_ZTAIL if tfsf_ztail else 0
            flags |= _native.TILE_K32 if k32 else 0
        _native.check_hip(self._lib.xf_set_tiling(self._h, tile_y, chunks_x, flags),
                          "xf_set_tiling")

    def set_timing(self, on: bool):
        _native.check_hip(self._lib.xf_set_timing(self._h, int(on)), "xf_set_timing")

    def autotune(self, trial_steps: int = 3) -> dict:
        """Times candidate tilings on this scene (state restored after each)
        and keeps the fastest; returns it."""
        t = XfTuning()
        _native.check_hip(self._lib.xf_autotune(self._h, trial_steps, ctypes.byref(t)),
                          "xf_autotune")
        return _tuning_dict(t)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.xf_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def _tuning_dict(t: XfTuning) -> dict:
    return {k: getattr(t, k) for k, _ in XfTuning._fields_}


class Batch:
    """Many independent 3-D scenes stepped together by the engine's batched
    kernels (xf_batch_* of include/xfdtd_hip.h): one persistent launch per
    Yee pass over the tiles of every grid, one launch for all TFSF faces and
    one for all NF2FF boxes per step, optionally replayed as a hipGraph."""

    def __init__(self, sims, precision: str = "fp64", use_graph: bool = False,
                 timing: bool = False, device: int = 0, watchdog: int = 0):
        self._lib = _native.hip()
        self._sims = list(sims)
        plan = XfPlan()
        plan.precision = PRECISIONS[precision]
        plan.use_graph = int(use_graph)
        plan.timing = int(timing)
        plan.device = device
        plan.watchdog = watchdog
        ptrs = (c_vp * len(self._sims))(*[s.scene_ptr() for s in self._sims])
        h = c_vp()
        self._h = None
        _native.check_hip(self._lib.xf_batch_create(ptrs, len(self._sims), ctypes.byref(plan),
                                                    ctypes.byref(h)), "xf_batch_create")
        self._h = h

    def upload(self):
        _native.check_hip(self._lib.xf_batch_upload(self._h), "xf_batch_upload")
        return self

    def run(self, n: int):
        _native.check_hip(self._lib.xf_batch_run(self._h, n), "xf_batch_run")
        return self

    def download(self):
        _native.check_hip(self._lib.xf_batch_download(self._h), "xf_batch_download")
        return self

    def set_tuning(self, tile_y: int = 1, planes: int = 4, blocks: int = 0,
                   nontemporal: bool = True, pml_split: bool = True, k32: bool = False,
                   xcd_remap: bool = False, pml_ztail: bool = False):
        """pml_split: z-face CPML in a launch of its own; pml_ztail (takes
        precedence): at the end of each Yee tile."""
        split = 2 if pml_ztail else int(bool(pml_split))
        t = XfTuning(tile_y, planes, blocks, int(nontemporal), int(xcd_remap), split, int(k32), 0.0)
        _native.check_hip(self._lib.xf_batch_set_tuning(self._h, ctypes.byref(t)),
                          "xf_batch_set_tuning")
        return self

    def tuning(self) -> dict:
        t = XfTuning()
        _native.check_hip(self._lib.xf_batch_get_tuning(self._h, ctypes.byref(t)),
                          "xf_batch_get_tuning")
        return _tuning_dict(t)

    def autotune(self, trial_steps: int = 3) -> dict:
        t = XfTuning()
        _native.check_hip(self._lib.xf_batch_autotune(self._h, trial_steps, ctypes.byref(t)),
                          "xf_batch_autotune")
        return _tuning_dict(t)

    def timing(self) -> dict:
        t = XfTiming()
        _native.check_hip(self._lib.xf_batch_get_timing(self._h, ctypes.byref(t)),
                          "xf_batch_get_timing")
        return {k: getattr(t, k) for k, _ in XfTiming._fields_}

    def reset_timing(self):
        _native.check_hip(self._lib.xf_batch_reset_timing(self._h), "xf_batch_reset_timing")

    def set_timing(self, on: bool):
        _native.check_hip(self._lib.xf_batch_set_timing(self._h, int(on)), "xf_batch_set_timing")

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(512)
        _native.check_hip(self._lib.xf_batch_describe(self._h, buf, 512), "xf_batch_describe")
        return buf.value.decode()

    def close(self):
        if getattr(self, "_h", None):
            self._lib.xf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
