"""Graded grids (Simulation::setCellSizes; the reference's GridSpace node and
size arrays, grid_space.cuh:36-90) pinned on the CPU by mathematics that
does not come from the coefficient builder.

Eigenmode pin.  In a lossless PEC box the Yee step maps E^n to E^{n+1} =
(2 - kappa) E^n - E^{n-1} on every discrete eigenvector v of the curl-curl
operator K (kappa = c^2 dt^2 lambda), starting from E^0 = v, H = 0 with E^1 =
(1 - kappa) v.  For a mode with E_z only, uniform along z, K separates into
one-dimensional finite-volume Laplacians along x and y:
  (K_x u)_i = -[(u_{i+1} - u_i)/e_i - (u_i - u_{i-1})/e_{i-1}] / h_i,
with e_i the cell edge between E nodes i, i+1 and h_i = (e_{i-1} + e_i)/2 the
dual edge around node i, both built here in numpy from the node coordinates.
The oracle run must keep the mode's shape to rounding and give the
recursion's amplitude; a curl coefficient with the wrong size (cell edge for
dual edge, an index off by one) turns the shape into a sum of modes within a
few hundred steps.  A second mode (E_x, uniform along x) covers z.
"""
import numpy as np
import pytest
import scipy.linalg

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes

C0 = xf.C_0


def _sizes(n, dl, seed, ratio=2.0):
    """n cell sizes of mean about dl, smoothly graded over a factor `ratio`
    with a seeded jitter (every size distinct)."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0.0, 1.0, n)
    s = dl * (1.0 + (ratio - 1.0) * 0.5 * (1 - np.cos(2 * np.pi * t)))
    return s * (1.0 + 0.05 * rng.uniform(-1, 1, n))


def _mode_1d(sizes):
    """Lowest eigenpair of the finite-volume Laplacian of E nodes 1..n-1 with
    zero ends (nodes from the cumulative sizes)."""
    e = np.asarray(sizes, float)
    n = e.size
    h = 0.5 * (e[:-1] + e[1:])  # dual edges around interior nodes 1..n-1
    T = np.zeros((n - 1, n - 1))
    for r in range(n - 1):
        T[r, r] = 1 / e[r] + 1 / e[r + 1]
        if r + 1 < n - 1:
            T[r, r + 1] = T[r + 1, r] = -1 / e[r + 1]
    lam, vec = scipy.linalg.eigh(T, np.diag(h))
    u = np.zeros(n + 1)
    u[1:n] = vec[:, 0]
    return lam[0], u / np.abs(u).max()


def _amplitude(kappa, steps):
    a_prev, a = 1.0, 1.0 - kappa
    for _ in range(steps - 1):
        a_prev, a = a, (2.0 - kappa) * a - a_prev
    return a if steps >= 1 else 1.0


def _cavity(sx, sy, sz, dl):
    s = xf.Simulation(dl, dl, dl, 0.9)
    s.add_cube("domain", (0.0, 0.0, 0.0), (float(np.sum(sx)), float(np.sum(sy)), float(np.sum(sz))),
               xf.Material.air())
    s.set_cell_sizes(0, sx).set_cell_sizes(1, sy).set_cell_sizes(2, sz)
    return s


@pytest.mark.parametrize("comp", ["ez", "ex"])
def test_graded_cavity_keeps_its_discrete_eigenmode(comp):
    dl = 2e-3
    sx, sy, sz = _sizes(24, dl, 1), _sizes(20, dl, 2, 2.5), _sizes(18, dl, 3, 1.7)
    steps = 400
    s = _cavity(sx, sy, sz, dl)
    s.init(steps)
    dt = s.dt
    assert dt == pytest.approx(0.9 / (C0 * np.sqrt(sum(1 / np.min(v) ** 2 for v in (sx, sy, sz)))), rel=1e-12)
    if comp == "ez":  # E_z(i, j) for all k: K = K_x + K_y
        (lx, ux), (ly, uy) = _mode_1d(sx), _mode_1d(sy)
        lam = lx + ly
        v = ux[:, None, None] * uy[None, :, None] * np.ones(len(sz))[None, None, :]
    else:  # E_x(j, k) for all i: K = K_y + K_z
        (ly, uy), (lz, uz) = _mode_1d(sy), _mode_1d(sz)
        lam = ly + lz
        v = np.ones(len(sx))[:, None, None] * uy[None, :, None] * uz[None, None, :]
    f = s.array(comp)
    assert f.shape == v.shape
    f[...] = v
    oracle.run(s.scene_ptr(), steps)
    a = _amplitude(C0 ** 2 * dt ** 2 * lam, steps)
    assert abs(a) > 0.05  # not at a node of the oscillation
    err = np.linalg.norm(f - a * v) / np.linalg.norm(a * v)
    assert err < 1e-9, err
    for other in ("ex", "ey", "ez"):
        if other != comp:
            assert np.abs(s.array(other)).max() < 1e-12


def test_wrong_size_breaks_the_eigenmode_pin():
    """The pin's power: the same check with the mode of a slightly different
    grading (the cell-edge Laplacian used for E, i.e. h_i = e_i) fails."""
    dl = 2e-3
    sx, sy, sz = _sizes(24, dl, 1), _sizes(20, dl, 2, 2.5), _sizes(18, dl, 3, 1.7)
    steps = 400
    s = _cavity(sx, sy, sz, dl)
    s.init(steps)

    def wrong_mode(sizes):
        e = np.asarray(sizes, float)
        n = e.size
        T = np.zeros((n - 1, n - 1))
        for r in range(n - 1):
            T[r, r] = 1 / e[r] + 1 / e[r + 1]
            if r + 1 < n - 1:
                T[r, r + 1] = T[r + 1, r] = -1 / e[r + 1]
        lam, vec = scipy.linalg.eigh(T, np.diag(e[1:]))
        u = np.zeros(n + 1)
        u[1:n] = vec[:, 0]
        return lam[0], u / np.abs(u).max()

    (lx, ux), (ly, uy) = wrong_mode(sx), wrong_mode(sy)
    v = ux[:, None, None] * uy[None, :, None] * np.ones(len(sz))[None, None, :]
    s.array("ez")[...] = v
    oracle.run(s.scene_ptr(), steps)
    a = _amplitude(C0 ** 2 * s.dt ** 2 * (lx + ly), steps)
    f = s.array("ez")
    assert np.linalg.norm(f - a * v) / np.linalg.norm(a * v) > 1e-3


def test_uniform_sizes_reproduce_the_uniform_grid():
    """setCellSizes with every size equal to the based size builds the same
    coefficients, dt and fields as the uniform grid, bit for bit."""
    steps = 60
    u = scenes.dielectric_sphere(n=20, pml=4, steps=steps, tfsf_inset=6, nf2ff_inset=5)
    g = scenes.dielectric_sphere(n=20, pml=4, steps=steps, tfsf_inset=6, nf2ff_inset=5, init=False)
    dl = 2.5e-3
    for ax in range(3):
        g.set_cell_sizes(ax, np.full(20, dl))
    g.init(steps)
    assert g.dt == u.dt
    for n in range(18):
        assert np.array_equal(g.array(f"coeff.{n}"), u.array(f"coeff.{n}")), n
    oracle.run(u.scene_ptr(), steps)
    oracle.run(g.scene_ptr(), steps)
    for n in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert np.array_equal(g.array(n), u.array(n)), n


def test_graded_grid_refusals():
    dl = 2e-3
    s = _cavity(_sizes(10, dl, 1), np.full(10, dl), np.full(10, dl), dl)
    s.set_coefficient_layout("compact")
    with pytest.raises(xf.XFError, match="dense coefficient layout"):
        s.init(10)
    # a TF/SF box over graded cells
    t = xf.Simulation(dl, dl, dl, 0.9)
    sx = _sizes(30, dl, 4)
    t.add_cube("domain", (0.0, 0.0, 0.0), (float(np.sum(sx)), 30 * dl, 30 * dl), xf.Material.air())
    t.set_cell_sizes(0, sx)
    tau = 20 * dl / 6e8
    t.add_tfsf3d((8, 8, 8), 0.0, 0.0, 0.0, xf.Waveform.gaussian(tau, 4.5 * tau))
    with pytest.raises(xf.XFError, match="TFSF box over graded cells"):
        t.init(10)
    # cells that do not add up to the domain
    b = xf.Simulation(dl, dl, dl, 0.9)
    b.add_cube("domain", (0.0, 0.0, 0.0), (20 * dl,) * 3, xf.Material.air())
    b.set_cell_sizes(1, np.full(19, dl))
    with pytest.raises(xf.XFError, match="add up"):
        b.init(10)
