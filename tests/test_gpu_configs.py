"""GPU <-> oracle parity at the stated sizes of two more workloads.

2-D TMz PEC cylinder, 1000^2 cells + 10-cell CPML (1020^2), dl = 1 mm,
  TF/SF Gaussian pulse, NF2FF at 1 GHz, three Ez probes (one-node movie
  monitors, every step), all 4000 steps (after test/cylinder_scatter_2d.cu:
  28-183): FP64 bit for bit on the fields, the CPML psi, the 2-D NF2FF
  currents and the probe series; FP32 within stated bars of the FP64 oracle.
Drude core + one-pole Lorentz shell, 256^3 + 10-cell CPML (276^3), three
  NF2FF frequencies (after test/dispersive_sphere_scatter/
  dispersive_sphere_scatter.cuh:31-194, m_lor_ade_updator.cu:27-88): FP64,
  a window of steps from a random state in which every field, psi and
  NF2FF accumulator and the J, J_prev and E_prev of every dispersive cell is
  live; bit for bit.
"""
import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes
from conftest import FIELDS, rel_l2

pytestmark = pytest.mark.gpu

PROBES = ((0.3, 0.0), (-0.25, 0.2), (0.0, -0.4))  # fractions of the domain edge


def _state_names(sim, n_freq=1):
    names = list(FIELDS)
    for p in range(6):
        try:
            sim.array(f"pml.{p}.ea_psi_hb")
        except xf.XFError:
            break
        names += [f"pml.{p}.{m}" for m in ("ea_psi_hb", "eb_psi_ha", "ha_psi_eb", "hb_psi_ea")]
    for f in range(n_freq):
        for face in xf.FACES:
            for c in ("ja", "jb", "ma", "mb"):
                try:
                    sim.array(f"nf2ff.0.{f}.{face}.{c}")
                except xf.XFError:
                    continue
                names.append(f"nf2ff.0.{f}.{face}.{c}")
    return names


def _cylinder(steps):
    s = scenes.cylinder_2d(n=1000, pml=10, steps=steps, init=False)
    half = 1000 * 1e-3 / 2
    for fx, fy in PROBES:
        s.add_movie_monitor("ez", (fx * 2 * half, fy * 2 * half, -np.inf), (1e-3, 1e-3, np.inf), 1)
    s.init(steps)
    return s


def _echo(sim):
    phi = np.linspace(-np.pi, np.pi, 360)
    inc = sim.incident_spectrum(1e9)
    return np.array([sim.far_field(0, 0, np.array([np.pi / 2]), p, inc)[2][0] for p in phi])


@pytest.fixture(scope="module")
def cylinder_oracle():
    steps = 4000
    print("case: 2-D cylinder 1020^2, oracle, 4000 steps", flush=True)
    c = _cylinder(steps)
    oracle.run(c.scene_ptr(), steps)
    return c


def test_cylinder_2d_full_size_fp64_bit_identical(gpu_available, cylinder_oracle):
    steps = 4000
    print("case: 2-D cylinder 1020^2 FP64, 4000 steps", flush=True)
    c = cylinder_oracle
    g = _cylinder(steps)
    e = xf.Engine(g).upload()
    e.run(1000).run(steps - 1000).download()
    e.close()
    names = _state_names(c)
    assert any(n.startswith("pml.") for n in names) and any(n.startswith("nf2ff.") for n in names)
    for n in names:
        a, b = g.array(n), c.array(n)
        assert np.array_equal(a, b), f"{n} differs (max |diff| {np.abs(a - b).max():.3e})"
    for q in range(len(PROBES)):
        pa, pb = g.movie(q), c.movie(q)
        assert pb.shape[0] == steps and np.abs(pb).max() > 0.1
        assert np.array_equal(pa, pb), f"probe {q} differs"
    assert np.abs(c.nf2ff_current(0, 0, "XP", "jb")).max() > 0
    for n in ("ex", "ey", "hz"):
        assert np.abs(g.array(n)).max() == 0


def test_cylinder_2d_full_size_fp32_against_fp64_oracle(gpu_available, cylinder_oracle):
    """FP32 fields and arithmetic (NF2FF sums in FP64) against the FP64
    oracle over the whole run.  Bars: probe series 1e-4 rel-L2 (measured
    5e-6), NF2FF currents 1e-4, echo width at 1 GHz 1e-5 (measured 5e-7),
    against the 1e-3 far-field bar the precision modes are held to."""
    steps = 4000
    print("case: 2-D cylinder 1020^2 FP32, 4000 steps", flush=True)
    c = cylinder_oracle
    g = _cylinder(steps)
    e = xf.Engine(g, precision="fp32").upload()
    e.run(steps).download()
    e.close()
    for q in range(len(PROBES)):
        err = rel_l2(g.movie(q), c.movie(q))
        assert err < 1e-4, (q, err)
    for n in _state_names(c):
        if n.startswith("nf2ff."):
            err = rel_l2(g.array(n), c.array(n))
            assert err < 1e-4, (n, err)
    err = rel_l2(_echo(g), _echo(c))
    assert err < 1e-5, err


def _dispersive_mask(sim):
    """(nx, ny, nz) cells whose ADE coefficients are not all zero."""
    nx, ny, nz = sim.shape
    m = sim.array("ade.coeff_e_j_sum") != 0
    for name in ("ade.coeff_e_e_p",):
        m |= sim.array(name) != 0
    for name in ("ade.coeff_j_j", "ade.coeff_j_j_p", "ade.coeff_j_e_n", "ade.coeff_j_e", "ade.coeff_j_e_p",
                 "ade.coeff_j_sum_j", "ade.coeff_j_sum_j_p"):
        m |= (sim.array(name) != 0).any(axis=-1)
    out = np.zeros_like(m)
    out[:nx, :ny, :nz] = m[:nx, :ny, :nz]
    return out


def test_dispersive_sphere_full_size_fp64_bit_identical(gpu_available):
    steps = 4
    print("case: 276^3 Drude + Lorentz sphere FP64, random state, 4 steps", flush=True)
    g = scenes.dispersive_sphere(n=256, pml=10, steps=steps + 1)
    c = scenes.dispersive_sphere(n=256, pml=10, steps=steps + 1)
    assert g.shape == (276, 276, 276)
    mask = _dispersive_mask(c)
    assert mask.sum() > 1_000_000
    rng = np.random.default_rng(31)
    for name in FIELDS:
        v = rng.random(c.array(name).shape, dtype=np.float32).astype(np.float64) - 0.5
        g.array(name)[...] = v
        c.array(name)[...] = v
    # J, J_prev (per pole) and E_prev live in the dispersive cells; the
    # engine keeps them for those cells only (the rest multiply zero
    # coefficients), so they start random there and zero elsewhere
    for name in ("ade.jx", "ade.jy", "ade.jz", "ade.jx_prev", "ade.jy_prev", "ade.jz_prev"):
        v = (rng.random(c.array(name).shape, dtype=np.float32).astype(np.float64) - 0.5) * 1e-3
        v[~mask] = 0.0
        g.array(name)[...] = v
        c.array(name)[...] = v
    for name in ("ade.ex_prev", "ade.ey_prev", "ade.ez_prev"):
        v = rng.random(c.array(name).shape, dtype=np.float32).astype(np.float64) - 0.5
        v[~mask] = 0.0
        g.array(name)[...] = v
        c.array(name)[...] = v
    e = xf.Engine(g).upload()
    desc = e.describe()
    e.run(1).run(steps - 1).download()
    e.close()
    oracle.run(c.scene_ptr(), steps)
    names = _state_names(c, n_freq=3)
    assert sum(n.startswith("nf2ff.") for n in names) == 3 * 6 * 4
    for n in names:
        a, b = g.array(n), c.array(n)
        assert np.array_equal(a, b), f"{n} differs (max |diff| {np.abs(a - b).max():.3e})"
    m4 = mask[..., None]
    for n in ("ade.jx", "ade.jy", "ade.jz", "ade.jx_prev", "ade.jy_prev", "ade.jz_prev"):
        a, b = g.array(n), c.array(n)
        assert np.array_equal(np.where(m4, a, 0), np.where(m4, b, 0)), n
        assert np.abs(b[mask]).max() > 0, n
    for n in ("ade.ex_prev", "ade.ey_prev", "ade.ez_prev"):
        a, b = g.array(n), c.array(n)
        assert np.array_equal(a[mask], b[mask]), n
    assert "ade_cells=" in desc and "ade_cells=0" not in desc
