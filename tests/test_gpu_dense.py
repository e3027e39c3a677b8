"""Dense per-cell Yee coefficients on the GPU (XF_COEFF_DENSE: the
reference's 18 per-field arrays streamed by the kernels,
fdtd_coefficient.cuh:149-167) and graded grids (GridSpace e/h node and size
arrays, grid_space.cuh:36-90), against the oracle, which reads the same
dense arrays (FP64: bit for bit)."""
import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes
from conftest import FIELDS

pytestmark = pytest.mark.gpu


def _state_names(sim):
    names = list(FIELDS)
    for p in range(6):
        try:
            sim.array(f"pml.{p}.ea_psi_hb")
        except xf.XFError:
            break
        names += [f"pml.{p}.{m}" for m in ("ea_psi_hb", "eb_psi_ha", "ha_psi_eb", "hb_psi_ea")]
    for face in xf.FACES:
        for c in ("ja", "jb", "ma", "mb"):
            try:
                names.append(f"nf2ff.0.0.{face}.{c}") if sim.array(f"nf2ff.0.0.{face}.{c}").size else None
            except xf.XFError:
                pass
    return names


def _gpu(sim, steps, coeff, precision="fp64"):
    e = xf.Engine(sim, precision=precision, coeff=coeff).upload()
    desc = e.describe()
    e.run(steps // 2).run(steps - steps // 2).download()
    e.close()
    return desc


def _assert_equal(g, c, names):
    for n in names:
        assert np.array_equal(g.array(n), c.array(n)), n


def _sizes(n, dl, seed, ratio=2.0):
    rng = np.random.default_rng(seed)
    t = np.linspace(0.0, 1.0, n)
    s = dl * (1.0 + (ratio - 1.0) * 0.5 * (1 - np.cos(2 * np.pi * t)))
    return s * (1.0 + 0.05 * rng.uniform(-1, 1, n))


def _graded_box(n, dl, steps, pml=4, disp=False, seed=0):
    """n^3 graded cells (every size distinct per axis), a CPML of the based
    size on every face, a dielectric (or Drude) sphere and three point
    sources."""
    sx, sy, sz = (_sizes(n, dl, seed + a, 1.8 + 0.3 * a) for a in range(3))
    s = xf.Simulation(dl, dl, dl, 0.9)
    ext = [float(np.sum(v)) for v in (sx, sy, sz)]
    s.add_cube("domain", (0.0, 0.0, 0.0), tuple(ext), xf.Material.air())
    tp = 2 * np.pi
    m = xf.Material.drude(1.5, [tp * 3e9], [tp * 0.3e9]) if disp else xf.Material.dielectric(3.0, 1.5, 0.01)
    s.add_sphere("sphere", tuple(0.5 * e for e in ext), 0.25 * min(ext), m)
    for a, v in enumerate((sx, sy, sz)):
        s.set_cell_sizes(a, v)
    s.add_pml_all(pml)
    rng = np.random.default_rng(seed)
    for q in range(3):
        f = 3e9 + 1e9 * q
        pos = [pml + n // 4 + int(rng.integers(0, n // 2)) for _ in range(3)]
        s.add_point_source(q % 3, pos, xf.Waveform.cosine_modulated_gaussian(1 / f, 3 / f, f), 1.0)
    s.add_movie_monitor("ez", (0.3 * ext[0], 0.5 * ext[1], 0.5 * ext[2]), (1e-9, 1e-9, 1e-9), 1)
    s.init(steps)
    return s


def test_dense_mode_sphere_fp64_bit_identical(gpu_available):
    """The TF/SF + CPML + NF2FF sphere with the coefficient arrays streamed
    instead of the table: fields, psi and NF2FF currents bit for bit."""
    steps = 80
    print("case: dense coefficients, 28^3 sphere FP64", flush=True)
    build = lambda: scenes.dielectric_sphere(n=20, pml=4, steps=steps, tfsf_inset=6, nf2ff_inset=5)
    g, c = build(), build()
    desc = _gpu(g, steps, "dense")
    assert "coeff=dense" in desc, desc
    oracle.run(c.scene_ptr(), steps)
    _assert_equal(g, c, _state_names(c))
    assert np.abs(c.array("ex")).max() > 0.1


def test_dense_mode_fp32_equals_table_fp32(gpu_available):
    """FP32: the dense arrays hold the table's float coefficients, so the two
    modes give the same bits."""
    steps = 80
    print("case: dense vs table FP32", flush=True)
    build = lambda: scenes.dielectric_sphere(n=20, pml=4, steps=steps, tfsf_inset=6, nf2ff_inset=5)
    a, b = build(), build()
    assert "coeff=dense" in _gpu(a, steps, "dense", "fp32")
    assert "coeff=table" in _gpu(b, steps, "table", "fp32")
    _assert_equal(a, b, _state_names(b))


def test_graded_grid_auto_dense_bit_identical(gpu_available):
    """A graded 44^3 grid has more than 65536 distinct coefficient rows per
    pass: the engine streams the dense arrays on its own, bit-identical to
    the oracle (fields, psi, the probe series)."""
    steps = 60
    print("case: graded 44^3 + CPML, auto coefficient mode", flush=True)
    g, c = _graded_box(44, 2e-3, steps), _graded_box(44, 2e-3, steps)
    desc = _gpu(g, steps, "auto")
    assert "coeff=dense" in desc, desc
    oracle.run(c.scene_ptr(), steps)
    _assert_equal(g, c, _state_names(c) + ["movie.0.data"])
    assert np.abs(c.array("movie.0.data")).max() > 0


def test_graded_grid_table_mode_bit_identical(gpu_available):
    """A small graded grid fits the table: the same bits through the ids."""
    steps = 60
    print("case: graded 16^3 + CPML, table", flush=True)
    g, c = _graded_box(16, 2e-3, steps, seed=3), _graded_box(16, 2e-3, steps, seed=3)
    assert "coeff=table" in _gpu(g, steps, "auto")
    oracle.run(c.scene_ptr(), steps)
    _assert_equal(g, c, _state_names(c) + ["movie.0.data"])


def test_graded_grid_dispersive_dense(gpu_available):
    """Dense coefficients feed the ADE kernel's curl terms too: a Drude
    sphere on a graded grid, J and the fields against the oracle."""
    steps = 60
    print("case: graded 20^3 Drude sphere, dense", flush=True)
    g, c = _graded_box(20, 2e-3, steps, disp=True, seed=5), _graded_box(20, 2e-3, steps, disp=True, seed=5)
    assert "coeff=dense" in _gpu(g, steps, "dense")
    oracle.run(c.scene_ptr(), steps)
    _assert_equal(g, c, _state_names(c))
    for n in ("ade.jx", "ade.jy", "ade.jz"):
        a, b = g.array(n), c.array(n)
        assert np.array_equal(a, b), n
    assert np.abs(c.array("ade.jx")).max() > 0


def test_graded_2d_dense(gpu_available):
    """2-D TMz on a graded grid with the dense arrays."""
    steps = 120
    print("case: graded 2-D 60^2, dense", flush=True)

    def build():
        dl = 2e-3
        sx, sy = _sizes(60, dl, 11), _sizes(60, dl, 12, 2.4)
        s = xf.Simulation(dl, dl, 1.0, 0.9)
        s.add_cube("domain", (0.0, 0.0, -np.inf), (float(np.sum(sx)), float(np.sum(sy)), np.inf),
                   xf.Material.air())
        s.add_cylinder("c", (0.06, 0.07, -np.inf), 0.02, np.inf, xf.Material.dielectric(4.0, 1.0, 0.0))
        s.set_cell_sizes(0, sx).set_cell_sizes(1, sy)
        s.add_pml_all(6, faces=("XN", "XP", "YN", "YP"))
        s.add_point_source(2, (30, 24, 0), xf.Waveform.cosine_modulated_gaussian(1 / 4e9, 3 / 4e9, 4e9), 1.0)
        s.init(steps)
        return s

    g, c = build(), build()
    assert "coeff=dense" in _gpu(g, steps, "dense")
    oracle.run(c.scene_ptr(), steps)
    _assert_equal(g, c, ["ez", "hx", "hy"] + [f"pml.{p}.{m}" for p in range(4)
                                              for m in ("ea_psi_hb", "eb_psi_ha", "ha_psi_eb", "hb_psi_ea")])
    assert np.abs(c.array("ez")).max() > 0
