"""Pins of the CPU oracle against mathematics that does not come from it.

The reference ships no golden vectors or known-answer tests (its drivers
only dump .npy files), so each part of the restated algorithm is pinned by
an independent closed form or invariant chosen so that a plausible mistake
(a dropped term, a wrong sign or index, a transposed operand) fails:

  Yee update + CPML + TFSF + NF2FF + far field   Mie series of the
      reference driver's sphere (tests/golden, scipy)
  Yee update                 PEC-cavity resonance = the Yee dispersion
                             relation; bounded energy; the flat and the
                             triple-loop forms of the reference agree
  TFSF corrector             empty-box leakage into the scattered region
  CPML corrector             reflection against a 3x larger domain
  ADE updators               the k = 0 (longitudinal) response of Drude,
                             Lorentz and Debye media, closed forms
  oracle                     bitwise determinism with OpenMP
"""
import os

import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes
from conftest import FIELDS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "mie_sphere_eps3_mu2_r0p1_1ghz.npz")


# ------------------------------------------------------------------ Mie
def test_mie_fixture_generator_rayleigh_limit():
    """The fixture's Mie code reproduces the Rayleigh small-sphere limit."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_mie
    th = np.linspace(0.1, np.pi - 0.1, 7)
    a, f = 1e-3, 1e9
    k = 2 * np.pi * f / make_mie.C0
    xz, yz = make_mie.rcs(a, 3.0, 1.0, f, th)
    r = 4 * np.pi * k**4 * a**6 * ((3 - 1) / (3 + 2)) ** 2
    np.testing.assert_allclose(yz, r, rtol=2e-3)
    np.testing.assert_allclose(xz, r * np.cos(th) ** 2, rtol=2e-3, atol=1e-6 * r)


def test_mie_far_field_of_the_driver_sphere():
    """test/dielectric_sphere_scatter.cu at its own dl = 5e-3: 70^3 vacuum +
    8-cell CPML, sphere r = 0.1 m (eps_r 3, mu_r 2), TFSF at 15, NF2FF at 11,
    1 GHz, the 360-point xz and yz cuts -- against the Mie series."""
    d = np.load(GOLDEN)
    s = scenes.dielectric_sphere(n=70, pml=8, steps=1500, dl=5e-3, radius=0.1)
    oracle.run(s.scene_ptr(), 1500)
    inc = s.incident_spectrum(1e9)
    _, _, rxz = s.far_field(0, 0, d["theta"], 0.0, inc)
    _, _, ryz = s.far_field(0, 0, d["theta"], np.pi / 2, inc)
    exz = np.linalg.norm(rxz - d["rcs_xz"]) / np.linalg.norm(d["rcs_xz"])
    eyz = np.linalg.norm(ryz - d["rcs_yz"]) / np.linalg.norm(d["rcs_yz"])
    assert exz < 0.05 and eyz < 0.05, (exz, eyz)


# --------------------------------------------------------------- cavity
def _cavity(nx=18, ny=6, nz=22, dl=1e-2):
    s = xf.Simulation(dl, dl, dl, 0.95)
    s.add_cube("domain", (0, 0, 0), (nx * dl, ny * dl, nz * dl), xf.Material.air())
    s.init(4000)
    return s


def test_pec_cavity_te101_resonance_matches_yee_dispersion():
    """A closed PEC box started in the TE101 pattern (Ey only) rings at the
    frequency the Yee dispersion relation gives for kx = pi/a, kz = pi/d:
    sin^2(w dt/2)/(c dt)^2 = sin^2(kx dx/2)/dx^2 + sin^2(kz dz/2)/dz^2."""
    nx, ny, nz, dl = 18, 6, 22, 1e-2
    s = _cavity(nx, ny, nz, dl)
    ey = s.array("ey")  # (nx+1, ny, nz+1), node i, j+1/2, k
    i = np.arange(nx + 1)[:, None, None]
    k = np.arange(nz + 1)[None, None, :]
    ey[...] = np.sin(np.pi * i / nx) * np.sin(np.pi * k / nz) * np.ones((1, ny, 1))
    ptr = s.scene_ptr()
    n = 3000
    probe = np.empty(n)
    for t in range(n):
        oracle.run(ptr, 1)
        probe[t] = ey[nx // 3, ny // 2, nz // 4]
    dt = s.dt
    c = xf.C_0
    rhs = (np.sin(np.pi / nx / 2) / dl) ** 2 + (np.sin(np.pi / nz / 2) / dl) ** 2
    w = 2 / dt * np.arcsin(c * dt * np.sqrt(rhs))
    # frequency from the phase of the analytic signal: fit cos(w t + p)
    tt = (np.arange(n) + 1) * dt
    best = None
    for ww in np.linspace(0.98 * w, 1.02 * w, 4001):
        a = np.vstack([np.cos(ww * tt), np.sin(ww * tt)]).T
        res = np.linalg.lstsq(a, probe, rcond=None)[1]
        if best is None or res[0] < best[0]:
            best = (res[0], ww)
    assert abs(best[1] - w) / w < 2e-5, (best[1], w)
    # energy stays bounded: no growth over thousands of steps
    assert np.abs(probe).max() < 1.01


def test_flat_and_triple_loop_yee_forms_agree():
    """update_scheme.cuh holds two forms of the Yee update (triple loop,
    :65-130, and the flat linear index loop the kernels use, :141-197)."""
    rng = np.random.default_rng(7)
    a = scenes.dielectric_sphere(n=12, pml=0, steps=4, tfsf_inset=4, nf2ff_inset=0)
    b = scenes.dielectric_sphere(n=12, pml=0, steps=4, tfsf_inset=4, nf2ff_inset=0)
    for name in FIELDS:
        v = rng.standard_normal(a.array(name).shape)
        a.array(name)[...] = v
        b.array(name)[...] = v
    for is_e in (False, True):
        for ax in range(3):
            oracle.yee_component(a.scene_ptr(), is_e, ax)
            oracle.yee_component_triple_loop(b.scene_ptr(), is_e, ax)
    for name in FIELDS:
        assert np.array_equal(a.array(name), b.array(name)), name


# ----------------------------------------------------------------- TFSF
@pytest.mark.parametrize("theta,phi,psi", [(0.0, 0.0, 0.0), (0.6, 0.9, 0.4), (2.2, 4.1, 1.1)])
def test_tfsf_empty_box_leakage(theta, phi, psi):
    """With no scatterer the total-field / scattered-field split is exact in
    the continuum: the scattered region only sees the leakage of the 1-D
    incident-field interpolation (< 1% of the incident peak)."""
    dl = 3e-3
    s = xf.Simulation(dl, dl, dl, 0.9)
    s.add_cube("domain", (-0.045,) * 3, (0.09,) * 3, xf.Material.air())
    s.add_pml_all(6)
    tau = 20 * dl / 6e8
    s.add_tfsf3d((10, 10, 10), theta, phi, psi, xf.Waveform.gaussian(tau, 4.5 * tau))
    s.init(260)
    ptr = s.scene_ptr()
    ex, ey, ez = s.array("ex"), s.array("ey"), s.array("ez")
    peak_tf = 0.0
    peak_sf = 0.0
    lo, hi = 10, 32  # 42^3 grid (30 + 2*6 CPML), TF box cells [10, 32)
    for t in range(26):
        oracle.run(ptr, 10)
        for f in (ex, ey, ez):
            inner = np.abs(f[lo + 2:hi - 2, lo + 2:hi - 2, lo + 2:hi - 2]).max()
            peak_tf = max(peak_tf, inner)
            sf = np.abs(f[7:9, 7:-7, 7:-7]).max()  # scattered-field slabs outside the box
            sf = max(sf, np.abs(f[-9:-7, 7:-7, 7:-7]).max())
            peak_sf = max(peak_sf, sf)
    assert peak_tf > 0.5
    assert peak_sf / peak_tf < 1e-2, peak_sf / peak_tf


# ----------------------------------------------------------------- CPML
def _probe_run(n, steps):
    """Sphere (r = 5 cells) in a 16-cell TF box centred in an n^3 domain +
    8-cell CPML; Ex recorded 12 cells below the centre, in the scattered-
    field region."""
    dl = 4e-3
    s = scenes.dielectric_sphere(n=n, pml=8, steps=steps, dl=dl, tfsf_inset=8 + (n - 16) // 2,
                                 nf2ff_inset=0, radius=0.02)
    ptr = s.scene_ptr()
    ex = s.array("ex")
    c = 8 + n // 2
    out = np.empty(steps)
    for t in range(steps):
        oracle.run(ptr, 1)
        out[t] = ex[c, c, c - 12]
    return out


def test_cpml_reflection_against_large_domain():
    """Scattered field at a probe 3 cells from the CPML of a 30^3 domain vs
    the same point of a 90^3 domain, whose boundary echo arrives after the
    200-step window: the CPML reflection stays below 1% (-40 dB) of the
    scattered peak."""
    steps = 200
    small = _probe_run(30, steps)
    large = _probe_run(90, steps)
    err = np.abs(small - large).max() / np.abs(large).max()
    assert np.abs(large).max() > 1e-3
    assert err < 1e-2, err


# ------------------------------------------------------------------ ADE
def _bulk(material, steps, j0=None):
    """Every cell dispersive, curl coefficients of E zeroed: each E is a
    0-D oscillator with the medium's k = 0 response, started by a step of
    the displacement D at t = 0 (E^-1 = 0, E^0 = 1, P^0 = 0).  j0 sets the
    polarisation current J^0 the history implies (a callable gets the
    scene)."""
    dl = 3e-3
    s = xf.Simulation(dl, dl, dl, 0.9)
    s.add_cube("domain", (0, 0, 0), (6 * dl,) * 3, xf.Material.air())
    s.add_cube("bulk", (-dl, -dl, -dl), (8 * dl,) * 3, material)
    s.init(steps)
    for n in (1, 2, 4, 5, 7, 8):
        s.array(f"coeff.{n}")[...] = 0.0
    s.array("ex")[...] = 1.0
    if j0 is not None:
        s.array("ade.jx")[...] = j0(s) if callable(j0) else j0
    ptr = s.scene_ptr()
    ex = s.array("ex")
    out = np.empty(steps)
    for t in range(steps):
        oracle.run(ptr, 1)
        out[t] = ex[3, 3, 3]
    return out, s.dt


def test_ade_drude_longitudinal_response():
    """eps_inf E'' + gamma eps_inf E' + wp^2 E = 0: E = exp(-g t/2)(cos Wt +
    g/(2W) sin Wt), W = sqrt(wp^2/eps_inf - g^2/4)."""
    wp, g, einf = 2 * np.pi * 2e9, 2 * np.pi * 0.2e9, 1.5
    e, dt = _bulk(xf.Material.drude(einf, [wp], [g]), 1500)
    t = (np.arange(1500) + 1) * dt
    W = np.sqrt(wp**2 / einf - g**2 / 4)
    ref = np.exp(-g * t / 2) * (np.cos(W * t) + g / (2 * W) * np.sin(W * t))
    assert np.abs(e - ref).max() < 5e-3, np.abs(e - ref).max()


def test_ade_lorentz_longitudinal_response():
    """E -> E_ss = eps_inf/(eps_inf + de) with a transient ringing at
    W = sqrt(w0^2 (eps_inf + de)/eps_inf - d^2), decaying as exp(-d t).
    The second-order recursion reads E^-1 through a central difference, so
    a step between n = -1 and 0 is only seen whole if J^0 carries the kick
    the recursion would have produced at n = -1 (cJEn * E^0).  The step's
    instant is then fixed to within dt/2, which bounds the waveform match
    (|dE/dt| dt/2 ~ 0.03); the steady state and the ringing frequency are
    checked tightly."""
    de, w0, dl_, einf = 2.0, 2 * np.pi * 1.5e9, 2 * np.pi * 0.1e9, 1.2
    n = 3000
    e, dt = _bulk(xf.Material.lorentz(einf, [de], [w0], [dl_]), n,
                  j0=lambda s: s.array("ade.coeff_j_e_n"))
    t = (np.arange(n) + 1) * dt
    W = np.sqrt(w0**2 * (einf + de) / einf - dl_**2)
    ess = einf / (einf + de)
    ref = ess + (1 - ess) * np.exp(-dl_ * t) * (np.cos(W * t) + dl_ / W * np.sin(W * t))
    assert np.abs(e - ref).max() < 0.04, np.abs(e - ref).max()
    assert abs(e[-1] - ess) < 2e-3, (e[-1], ess)
    # ringing frequency from the zero crossings of E - E_ss over the first ns
    x = e[:400] - ess
    cross = np.where(np.diff(np.sign(x)) != 0)[0]
    period = 2 * np.mean(np.diff(cross)) * dt
    assert abs(2 * np.pi / period - W) / W < 1e-2, (2 * np.pi / period, W)


def test_ade_debye_longitudinal_response():
    """P(0) = 0, J(0) = eps0 de / tau: E relaxes to eps_inf/(eps_inf + de)
    with time constant tau eps_inf/(eps_inf + de)."""
    de, tau, einf = 3.0, 1e-10, 2.0
    j0 = xf.EPSILON_0 * de / tau
    e, dt = _bulk(xf.Material.debye(einf, [de], [tau]), 800, j0=j0)
    t = (np.arange(800) + 1) * dt
    ess = einf / (einf + de)
    ref = ess + (1 - ess) * np.exp(-t * (einf + de) / (einf * tau))
    assert np.abs(e - ref).max() < 5e-3, np.abs(e - ref).max()


# ------------------------------------------------------------ determinism
def test_oracle_bitwise_deterministic():
    a = scenes.dielectric_sphere(n=20, pml=5, steps=60, tfsf_inset=8, nf2ff_inset=6)
    b = scenes.dielectric_sphere(n=20, pml=5, steps=60, tfsf_inset=8, nf2ff_inset=6)
    oracle.run(a.scene_ptr(), 60)
    oracle.run(b.scene_ptr(), 60)
    for n in FIELDS + ("pml.3.eb_psi_ha", "nf2ff.0.0.YP.mb"):
        assert np.array_equal(a.array(n), b.array(n)), n
