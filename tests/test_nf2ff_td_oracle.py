"""Pins of the NF2FF time-domain restatement (oracle/xf_oracle.c
nf2ff_td_face, after src/nf2ff/time_domain/nf2ff_time_domain_data.cu:103-188)
and of the host far-field assembly, without a GPU.

* Moments: every contribution is a field increment dX split over the two
  bins around its fractional arrival s = D + off + t with weights (1 - f, f),
  f = s - floor(s).  So, exactly, sum_m p[m] = sum c ds dX,
  sum_m m p[m] = sum c ds dX s and sum_m m^2 p[m] = sum c ds dX (s^2 + f(1 - f)).
  The right-hand sides come from per-step field snapshots interpolated here
  from the Yee staggering (not from the oracle's interpolation), with D from
  the grid geometry.  A wrong sign, offset, position, area, interpolation
  stencil, prev update or floor split fails one of the three.
* Frequency domain: with z = exp(-j w dt), sum_m p[m] z^m equals
  2j sin(w dt / 2) / dt e^{-j w dmax / c} sum ds J(w) e^{j w r . rhat / c}
  plus the end-of-run term, J(w) being the frequency-domain box's phasor of
  the same current; what remains is the linear interpolation of the
  fractional delay, f(1 - f) (w dt)^2 / 2 at most per contribution, i.e.
  (w dt)^2 / 8.  A half-step slip would be w dt / 2 ~ 4e-2.
* Far field: the DFT of R E_theta(t), R E_phi(t) from processFarField
  equals sinc(w dt / 2) e^{-j w dmax / c} times the frequency-domain box's
  processFarField once the fields have left the box.
"""
import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes

C0 = 299792458.0
FACES = xf.FACES


def _face_tasks(n, inset):
    out = []
    for f in range(6):
        c = f // 2
        t = [(inset, n - inset)] * 3
        pos = n - inset if f % 2 else inset
        t[c] = (pos, pos + 1)
        out.append(t)
    return out


def _face_fields(sim, is_e, face, task):
    """(A, B): the tangential components at the face centres of `task`.  E_q
    sits at a cell centre along q and on nodes along the other axes, H_q on a
    node along q and at centres along the others; the face centre is a node
    along the normal and a centre along the two tangential axes.  Along each
    axis where a component is off that point it is the mean of the two
    neighbours on either side."""
    c = face // 2
    out = []
    for comp in ((c + 1) % 3, (c + 2) % 3):
        arr = sim.array(("e" if is_e else "h") + "xyz"[comp])
        shifts = []
        for ax in range(3):
            want_centre = ax != c
            at_centre = (ax == comp) if is_e else (ax != comp)
            if want_centre == at_centre:
                shifts.append((0,))
            elif want_centre:
                shifts.append((0, 1))
            else:
                shifts.append((-1, 0))
        acc, cnt = 0.0, 0
        for di in shifts[0]:
            for dj in shifts[1]:
                for dk in shifts[2]:
                    acc = acc + arr[task[0][0] + di:task[0][1] + di, task[1][0] + dj:task[1][1] + dj,
                                    task[2][0] + dk:task[2][1] + dk]
                    cnt += 1
        out.append(acc / cnt)
    return out


def _geometry(n_total, inset, origin, dl, u):
    """r . rhat at each surface cell (node along the normal, centres along
    the tangential axes) of each face."""
    rd = []
    for f, t in enumerate(_face_tasks(n_total, inset)):
        c = f // 2
        coords = []
        for ax in range(3):
            idx = np.arange(t[ax][0], t[ax][1], dtype=np.float64)
            coords.append(origin + dl * (idx if ax == c else idx + 0.5))
        X, Y, Z = np.meshgrid(*coords, indexing="ij")
        rd.append(X * u[0] + Y * u[1] + Z * u[2])
    return rd


def _unit(theta, phi):
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


N, PML, INSET, TFSF, DL = 16, 4, 5, 7, 2.5e-3
THETA, PHI = 0.7, 0.3


@pytest.fixture(scope="module")
def stepped():
    """A dielectric sphere with a TD box and a FD box on the same surface,
    stepped one step at a time with the face fields recorded."""
    steps = 200
    s = scenes.dielectric_sphere(n=N, pml=PML, steps=steps, tfsf_inset=TFSF, nf2ff_inset=INSET,
                                 freqs=(3e9,), td_dirs=[(THETA, PHI)])
    (n_total, _, _), dt, _ = s.info()
    tasks = _face_tasks(n_total, INSET)
    hist = {(f, e): [] for f in range(6) for e in (True, False)}
    ptr = s.scene_ptr()
    for _ in range(steps):
        oracle.run(ptr, 1)
        for f in range(6):
            for e in (True, False):
                a, b = _face_fields(s, e, f, tasks[f])
                hist[(f, e)].append((a.copy(), b.copy()))
    origin = -N * DL / 2 - PML * DL
    rd = _geometry(n_total, INSET, origin, DL, _unit(THETA, PHI))
    return s, hist, rd, dt, steps


def _coeffs(f, e):
    sgn = 1.0 if f % 2 else -1.0
    return (1.0 if e else -1.0) * sgn, (-1.0 if e else 1.0) * sgn


def test_moments_of_the_potentials(stepped):
    s, hist, rd, dt, steps = stepped
    dmax = max(r.max() for r in rd)
    assert s.array("nf2ff_td.0.XN.wa").shape == (steps + int(np.ceil(
        (dmax - min(r.min() for r in rd)) / (C0 * dt))) + 2,)
    checked = 0
    for f in range(6):
        D = (dmax - rd[f]) / (C0 * dt)
        for e in (True, False):
            off = 0.5 if e else 0.0
            ca, cb = _coeffs(f, e)
            A = np.array([h[0] for h in hist[(f, e)]])
            B = np.array([h[1] for h in hist[(f, e)]])
            dA = np.diff(A, axis=0, prepend=np.zeros_like(A[:1]))
            dB = np.diff(B, axis=0, prepend=np.zeros_like(B[:1]))
            sv = D[None] + off + np.arange(steps)[:, None, None, None]
            fr = sv - np.floor(sv)
            names = ("ua", "ub") if e else ("wa", "wb")
            for name, coef, d in ((names[0], ca, dB), (names[1], cb, dA)):
                p = s.array(f"nf2ff_td.0.{FACES[f]}.{name}")
                m = np.arange(p.size, dtype=np.float64)
                w = coef * DL * DL * d
                scale = np.abs(w).sum()
                assert scale > 0
                assert abs(p.sum() - w.sum()) <= 1e-13 * scale, (FACES[f], name, 0)
                assert abs((p * m).sum() - (w * sv).sum()) <= 1e-13 * scale * steps, (FACES[f], name, 1)
                assert abs((p * m * m).sum() - (w * (sv ** 2 + fr * (1 - fr))).sum()) <= \
                    1e-13 * scale * steps ** 2, (FACES[f], name, 2)
                checked += 1
            # the prev arrays hold the last step's face fields, indexed (a, b)
            c = f // 2
            for qn, X in zip(("ea", "eb") if e else ("ha", "hb"), (A[-1], B[-1])):
                ab = np.squeeze(X, axis=c)
                if (c + 1) % 3 > (c + 2) % 3:
                    ab = ab.T
                np.testing.assert_allclose(s.array(f"nf2ff_td.0.{FACES[f]}.prev.{qn}"), ab, rtol=0,
                                           atol=1e-15 * max(1.0, np.abs(X).max()))
    assert checked == 24


def test_potentials_match_frequency_domain_phasors(stepped):
    s, hist, rd, dt, steps = stepped
    dmax = max(r.max() for r in rd)
    w0 = 2 * np.pi * 3e9
    z = np.exp(-1j * w0 * dt)
    fac = 2j * np.sin(w0 * dt / 2) / dt * np.exp(-1j * w0 * dmax / C0)
    bound = (w0 * dt) ** 2 / 8
    worst = 0.0
    for f in range(6):
        ph = np.exp(1j * w0 * rd[f] / C0)
        for e in (True, False):
            off = 0.5 if e else 0.0
            ca, cb = _coeffs(f, e)
            a_last, b_last = hist[(f, e)][-1]
            pairs = ((("ua", "ma", ca, b_last), ("ub", "mb", cb, a_last)) if e else
                     (("wa", "ja", ca, b_last), ("wb", "jb", cb, a_last)))
            for pn, cn, coef, last in pairs:
                p = s.array(f"nf2ff_td.0.{FACES[f]}.{pn}")
                lhs = (p * z ** np.arange(p.size)).sum()
                cur = s.array(f"nf2ff.0.0.{FACES[f]}.{cn}")
                cur = cur[..., 0] + 1j * cur[..., 1]
                rhs = fac * (DL * DL * cur * ph).sum()
                rhs += np.exp(-1j * w0 * dmax / C0) * z ** (steps + off) * \
                    (DL * DL * coef * last * ph).sum()
                err = abs(lhs - rhs) / abs(rhs)
                worst = max(worst, err)
                assert err <= bound, (FACES[f], pn, err, bound)
    # the interpolation error averages f(1 - f) / 2 = 1/12 over uniform f
    assert worst > 0.3 * (w0 * dt) ** 2 / 12


def test_far_field_matches_frequency_domain_far_field():
    """A zero-mean incident pulse (cosine-modulated Gaussian at 3 GHz), so no
    static remainder keeps the truncated DFT of the frequency-domain box from
    converging; 800 steps let the scattered pulse leave the surface."""
    steps, f0, pml, inset = 800, 3e9, 6, INSET + 2
    dirs = [(0.7, 0.3), (np.pi / 2, 0.0), (2.5, 4.0)]
    edge = N * DL
    s = xf.Simulation(DL, DL, DL, 0.9)
    s.add_cube("domain", (-edge / 2,) * 3, (edge,) * 3, xf.Material.air())
    s.add_sphere("sphere", (0, 0, 0), 0.2857 * edge, xf.Material.dielectric(3.0, 2.0))
    s.add_tfsf3d((TFSF + 2,) * 3, 0.0, 0.0, 0.0,
                 xf.Waveform.cosine_modulated_gaussian(1 / f0, 3 / f0, f0))
    s.add_pml_all(pml)
    s.add_nf2ff_fd((inset,) * 3, [f0])
    for th, phi in dirs:
        s.add_nf2ff_td((inset,) * 3, th, phi)
    s.init(steps)
    oracle.run(s.scene_ptr(), steps)
    (n_total, _, _), dt, _ = s.info()
    w0 = 2 * np.pi * f0
    origin = -edge / 2 - pml * DL
    sinc = 2 * np.sin(w0 * dt / 2) / (w0 * dt)
    for q, (th, phi) in enumerate(dirs):
        t, et, ep = s.far_field_td(q)
        assert np.allclose(t, dt * np.arange(t.size))
        dmax = max(r.max() for r in _geometry(n_total, inset, origin, DL, _unit(th, phi)))
        ft, fp, _ = s.far_field(0, 0, np.array([th]), phi)
        ref_t = sinc * np.exp(-1j * w0 * dmax / C0) * ft[0]
        ref_p = sinc * np.exp(-1j * w0 * dmax / C0) * fp[0]
        z = np.exp(-1j * w0 * t)
        got_t, got_p = (et * z).sum() * dt, (ep * z).sum() * dt
        den = np.hypot(abs(ref_t), abs(ref_p))
        bound = (w0 * dt) ** 2 / 8
        assert abs(got_t - ref_t) <= bound * den, (q, abs(got_t - ref_t) / den)
        assert abs(got_p - ref_p) <= bound * den, (q, abs(got_p - ref_p) / den)


def test_two_d_grid_rejected():
    s = xf.Simulation(1e-3, 1e-3, 1.0, 0.9)
    s.add_cube("domain", (-0.02, -0.02, -np.inf), (0.04, 0.04, np.inf), xf.Material.air())
    s.add_nf2ff_td((5, 5, 0), 0.5, 0.0)
    with pytest.raises(xf.XFError, match="3-D"):
        s.init(10)


def test_engine_accepts_the_box_and_rejects_long_cells():
    """xf_create checks the TD box without a device: a scene built by the
    host passes; cells so long that a 16-cell tile spans more time bins than
    a tile's partial holds are refused with XF_ERR_UNSUPPORTED."""
    import ctypes
    from xfdtd_cuda_amd import _native

    def create(sim):
        p = _native.XfPlan()
        h = ctypes.c_void_p()
        st = _native.hip().xf_create(ctypes.c_void_p(sim.scene_ptr()), ctypes.byref(p),
                                     ctypes.byref(h))
        if st == 0:
            _native.hip().xf_destroy(h)
        return st, _native.hip().xf_last_error().decode()

    ok = scenes.dielectric_sphere(n=N, pml=PML, steps=30, tfsf_inset=TFSF, nf2ff_inset=INSET,
                                  td_dirs=[(0.3, 0.2)])
    assert create(ok)[0] == 0
    s = xf.Simulation(0.1, 1e-3, 1e-3, 0.9)
    s.add_cube("domain", (0, 0, 0), (2.0, 0.03, 0.03), xf.Material.air())
    s.add_nf2ff_td((2, 2, 2), np.pi / 2, 0.0)
    s.init(10)
    st, msg = create(s)
    assert st == 4 and "time bins" in msg, (st, msg)
