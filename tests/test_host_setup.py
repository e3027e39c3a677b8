"""Closed-form checks of the host set-up (the XFDTD_CORE subset restated in
xfdtd_cuda_amd/csrc/host): time step, Yee coefficients, CPML profiles,
TFSF incident table and projections, NF2FF phasors, ADE coefficients,
voxelisation and the reference drivers' grid sizes."""
import numpy as np
import pytest

import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes

EPS0, MU0, C0 = xf.EPSILON_0, xf.MU_0, xf.C_0


def test_driver_grid_sizes():
    """dielectric_sphere_scatter.cu at dl = 2.5e-3: 140 + 2*8 = 156 cells;
    benchmark.cu: 128 + 16; cylinder_scatter_2d: 2-D grid (nz = 1)."""
    s = scenes.dielectric_sphere(steps=10)
    assert s.shape == (156, 156, 156)
    b = scenes.vacuum_benchmark(kind=4, steps=10)
    assert b.shape == (144, 144, 144)
    c = scenes.cylinder_2d(n=100, pml=10, steps=10)
    (nx, ny, nz), _, dim = c.info()
    assert (nx, ny, nz, dim) == (120, 120, 1, 2)


def test_time_step_courant():
    s = scenes.dielectric_sphere(n=20, pml=4, steps=10, tfsf_inset=7, nf2ff_inset=5)
    dl = 2.5e-3
    assert s.dt == pytest.approx(0.9 / (C0 * np.sqrt(3 / dl**2)), rel=1e-14)
    c = scenes.cylinder_2d(n=40, pml=5, steps=10, tfsf_inset=8, nf2ff_inset=6)
    assert c.dt == pytest.approx(0.9 / (C0 * np.sqrt(2 / 1e-3**2)), rel=1e-14)


def test_yee_coefficients_vacuum_and_dielectric():
    """test/2d.cu:694-735 forms, signs of the cyclic (a, b, c) convention:
    E_c' = ce E_c + c_a (H_a(b) - H_a(b-1)) + c_b (H_b(a) - H_b(a-1)) with
    eps dE_x/dt = dHz/dy - dHy/dz etc."""
    s = scenes.dielectric_sphere(n=20, pml=0, steps=10, tfsf_inset=7, nf2ff_inset=0, eps_r=3,
                                 mu_r=2)
    dt, dl = s.dt, 2.5e-3
    c = [s.array(f"coeff.{n}") for n in range(18)]
    # a vacuum corner cell
    assert c[0][0, 0, 0] == 1.0
    assert c[1][0, 0, 0] == pytest.approx(-dt / (EPS0 * dl), rel=1e-14)  # cexhy
    assert c[2][0, 0, 0] == pytest.approx(dt / (EPS0 * dl), rel=1e-14)  # cexhz
    assert c[4][0, 0, 0] == pytest.approx(-dt / (EPS0 * dl), rel=1e-14)  # ceyhz
    assert c[5][0, 0, 0] == pytest.approx(dt / (EPS0 * dl), rel=1e-14)  # ceyhx
    assert c[10][0, 0, 0] == pytest.approx(dt / (MU0 * dl), rel=1e-14)  # chxey
    assert c[11][0, 0, 0] == pytest.approx(-dt / (MU0 * dl), rel=1e-14)  # chxez
    assert c[16][0, 0, 0] == pytest.approx(dt / (MU0 * dl), rel=1e-14)  # chzex
    # the sphere centre: eps_r 3, mu_r 2
    m = 10
    assert c[8][m, m, m] == pytest.approx(dt / (3 * EPS0 * dl), rel=1e-14)  # cezhy
    assert c[14][m, m, m] == pytest.approx(-dt / (2 * MU0 * dl), rel=1e-14)  # chyex


def test_lossy_coefficient_and_pec():
    s = xf.Simulation(1e-3, 1e-3, 1e-3, 0.9)
    s.add_cube("domain", (0, 0, 0), (0.01,) * 3, xf.Material.air())
    s.add_cube("lossy", (0, 0, 0), (0.005, 0.01, 0.01), xf.Material.dielectric(2.0, 1.0, 0.5))
    s.add_cube("pec", (0.005, 0, 0), (0.005, 0.01, 0.01), xf.Material.pec())
    s.init(5)
    dt = s.dt
    ce = s.array("coeff.0")
    eps = 2 * EPS0
    assert ce[2, 5, 5] == pytest.approx((2 * eps - 0.5 * dt) / (2 * eps + 0.5 * dt), rel=1e-14)
    # PEC as sigma = 1e10 (test/2d.cu:627-646): ce -> -1, curl terms -> 0
    assert ce[7, 5, 5] == pytest.approx(-1.0, abs=1e-6)
    vac = dt / (EPS0 * 1e-3)  # |curl coefficient| in vacuum
    # ratio 2 eps0 / (sigma dt) ~ 1e-9
    assert abs(s.array("coeff.1")[7, 5, 5]) < 1e-8 * vac


def test_cpml_profiles():
    """Roden & Gedney CFS-CPML: b = exp(-(sigma/kappa + alpha) dt/eps0),
    a = sigma/(sigma kappa + kappa^2 alpha) (b - 1); at the inner interface
    sigma = 0 so a = 0 and b = exp(-alpha_max dt / eps0); sigma grows as the
    cube of the depth (order 3)."""
    s = scenes.dielectric_sphere(n=20, pml=8, steps=10, tfsf_inset=11, nf2ff_inset=9)
    dt, dl = s.dt, 2.5e-3
    a_e = s.array("pml.0.coeff_a_e")  # XN face: E nodes 1..8, depth 7..0
    b_e = s.array("pml.0.coeff_b_e")
    assert a_e[-1] == 0.0
    assert b_e[-1] == pytest.approx(np.exp(-0.05 * dt / EPS0), rel=1e-14)
    sigma_max = 0.8 * 4 / (xf.ETA_0 * dl)
    rho = 7 / 8
    sig = sigma_max * rho**3
    alpha = 0.05 * (1 - rho)
    b = np.exp(-(sig + alpha) * dt / EPS0)
    assert b_e[0] == pytest.approx(b, rel=1e-13)
    assert a_e[0] == pytest.approx(sig / (sig + alpha) * (b - 1), rel=1e-13)
    # XP face mirrors XN (E depth 0..7 from the interface inward)
    a_p = s.array("pml.1.coeff_a_e")
    np.testing.assert_allclose(a_p, a_e[::-1], rtol=1e-13)
    # H nodes sit half a cell deeper/shallower
    assert s.array("pml.0.coeff_a_h")[0] != 0.0


def test_cpml_psi_shapes_and_cpsi():
    s = scenes.dielectric_sphere(n=20, pml=6, steps=10, tfsf_inset=9, nf2ff_inset=7)
    nx = 32
    assert s.array("pml.0.ea_psi_hb").shape == (6, nx, nx + 1)  # Ey lateral
    assert s.array("pml.0.eb_psi_ha").shape == (6, nx + 1, nx)  # Ez
    assert s.array("pml.2.ea_psi_hb").shape == (nx + 1, 6, nx)  # Y face, Ez
    assert s.array("pml.4.hb_psi_ea").shape == (nx, nx + 1, 6)  # Z face, Hy
    # c_psi of Ey on an X face is Ey's Hz curl coefficient (unscaled)
    np.testing.assert_array_equal(s.array("pml.0.c_ea_psi_hb")[0], s.array("coeff.4")[1])


def test_tfsf_incident_table_and_projections_normal_incidence():
    """theta = phi = psi = 0: Ex-polarised plane wave along +z; ratio 1; the
    incident point sits two aux cells before the first TF plane; the aux
    line is driven by the waveform sampled at t = n dt."""
    s = scenes.dielectric_sphere(n=20, pml=4, steps=50, tfsf_inset=7, nf2ff_inset=5)
    te = s.array("tfsf.e_inc")
    _, v = s.waveform_samples()
    np.testing.assert_array_equal(te[:, 0], v)
    np.testing.assert_allclose(s.array("tfsf.proj_z_int")[:5], [2, 3, 4, 5, 6], rtol=1e-15)
    np.testing.assert_allclose(s.array("tfsf.proj_z_half")[:3], [1.5, 2.5, 3.5], rtol=1e-15)
    assert np.all(s.array("tfsf.proj_x_int") == 0)
    tau = 20 * 2.5e-3 / 6e8
    t = np.arange(50) * s.dt
    np.testing.assert_allclose(v, np.exp(-((t - 4.5 * tau) / tau) ** 2), rtol=1e-14)


def test_tfsf_aux_line_propagates_at_c():
    """The 1-D aux grid carries the pulse at the (ratio-matched) speed of
    light: the peak moves c dt / (dl / ratio) aux cells per step."""
    s = scenes.dielectric_sphere(n=40, pml=4, steps=400, tfsf_inset=7, nf2ff_inset=5)
    e = s.array("tfsf.e_inc")
    # the peak enters at t0 = 4.5 tau (~87 steps) and crosses ~40 aux cells
    p1, p2 = np.argmax(e[100]), np.argmax(e[140])
    cells_per_step = C0 * s.dt / 2.5e-3
    assert p2 > p1
    assert abs((p2 - p1) / 40 - cells_per_step) < 0.03


def test_nf2ff_phasors():
    s = scenes.dielectric_sphere(n=20, pml=4, steps=30, tfsf_inset=7, nf2ff_inset=5, freqs=(1e9, 2e9))
    dt = s.dt
    for q, f in enumerate((1e9, 2e9)):
        te = s.array(f"nf2ff.0.{q}.transform_e")
        th = s.array(f"nf2ff.0.{q}.transform_h")
        n = np.arange(30)
        np.testing.assert_allclose(te[:, 0] + 1j * te[:, 1], dt * np.exp(-2j * np.pi * f * (n + 1) * dt), rtol=1e-12)
        np.testing.assert_allclose(th[:, 0] + 1j * th[:, 1], dt * np.exp(-2j * np.pi * f * (n + 0.5) * dt), rtol=1e-12)
    assert s.array("nf2ff.0.0.XN.ja").shape == (1, 18, 18, 2)  # 28 cells, inset 5


def test_material_permittivity_closed_forms():
    w = 2 * np.pi * 1e9
    d = xf.Material.drude(2.0, [3e9], [4e8])
    assert d.permittivity(w) == pytest.approx(2.0 - 9e18 / (w * w - 1j * w * 4e8), rel=1e-14)
    de = xf.Material.debye(2.0, [3.0], [1e-10])
    assert de.permittivity(w) == pytest.approx(2.0 + 3.0 / (1 + 1j * w * 1e-10), rel=1e-14)
    lo = xf.Material.lorentz(1.5, [2.0], [7e9], [3e8])
    assert lo.permittivity(w) == pytest.approx(1.5 + 2.0 * 49e18 / (49e18 + 2j * w * 3e8 - w * w),
                                               rel=1e-14)


def test_ade_storage_kinds():
    tp = 2 * np.pi
    s = scenes.dispersive_sphere(n=20, pml=4, steps=10)
    assert s.array("ade.jx").shape == (29, 29, 29, 1)
    # Drude + Lorentz together use the M-Lorentz storage (E history present)
    assert s.array("ade.ex_prev").shape == (29, 29, 29)
    # vacuum cells: all ADE coefficients zero
    assert s.array("ade.coeff_e_j_sum")[0, 0, 0] == 0.0
    assert s.array("ade.coeff_e_j_sum")[14, 14, 14] < 0.0
    del tp


def test_voxelisation_sphere_volume():
    """Cell-centre sampling of the 140^3 driver sphere: the voxel volume
    approaches 4/3 pi r^3."""
    s = scenes.dielectric_sphere(steps=5)
    ce = s.array("coeff.8")[:-1, :-1, :]  # cezhy: 1/eps inside the sphere
    inside = np.isclose(ce, ce.min())
    vol = inside.sum() * 2.5e-3**3
    assert vol == pytest.approx(4 / 3 * np.pi * 0.1**3, rel=0.01)


def test_compact_layout_has_no_dense_arrays():
    s = scenes.dielectric_sphere(n=16, pml=4, steps=5, tfsf_inset=6, nf2ff_inset=5,
                                 layout="compact")
    assert s.array("coeff.0").size == 0


def test_movie_monitor_box_and_frames(tmp_path):
    """The driver's movie (test/dielectric_sphere_scatter.cu:47-52): Ex over
    the xz plane y in [0, dl) of the 140^3 domain, a frame every 20 steps;
    the oracle fills frame t / 20 with Ex over the box at step t, and
    output() writes them as %05d.npy."""
    import oracle
    s = scenes.dielectric_sphere(n=20, pml=4, steps=45, dl=5e-3, tfsf_inset=7, nf2ff_inset=5,
                                 init=False)
    half = 20 * 5e-3 / 2
    s.add_movie_monitor("ex", (-half, 0.0, -half), (2 * half, 5e-3, 2 * half), 20,
                        str(tmp_path / "movie"))
    s.init(45)
    m = s.movie(0)
    assert m.shape == (3, 20, 1, 20)  # frames at t = 0, 20, 40; x cells 4..24; y node 14
    snaps = {}
    for t in range(45):
        if t % 20 == 0:
            snaps[t // 20] = None
        oracle.run(s.scene_ptr(), 1)
        if t % 20 == 0:
            snaps[t // 20] = s.array("ex")[4:24, 14:15, 4:24].copy()
    # frame t/20 is recorded in step t after the E pass, i.e. the field after the step
    for f in range(3):
        assert np.array_equal(m[f], snaps[f])
    assert np.abs(m[2]).max() > 0
    assert s.output_movies() == 3
    back = np.load(tmp_path / "movie" / "00002.npy")
    assert np.array_equal(back, m[2])
