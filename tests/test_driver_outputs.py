"""The drivers' on-disk outputs (test/dielectric_sphere_scatter.cu:85-99,
test/benchmark.cu:181-203, dispersive_sphere_scatter.cuh:113-137) written by
the host set-up library after a run: the far-field cuts of
NFFFTFrequencyDomain::processFarField(theta, phi, sub_dir), the time-domain
box's processFarField(), and time.npy / incident_wave.npy of the TFSF
waveform.  The scene is advanced by the oracle, so no GPU is needed; the
files must hold exactly what the in-memory calls return, in a layout numpy
reads back, and the RCS the drivers' post-processing forms from them must
equal the library's own normalised RCS."""
import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes


@pytest.fixture(scope="module")
def run():
    s = scenes.dielectric_sphere(n=16, pml=4, steps=120, tfsf_inset=7, nf2ff_inset=5,
                                 freqs=(1e9, 2.5e9), td_dirs=[(np.pi, 0.0)])
    oracle.run(s.scene_ptr(), 120)
    return s


def test_far_field_cuts_on_disk(run, tmp_path):
    s = run
    theta = np.linspace(-np.pi, np.pi, 37)
    phi = np.linspace(-np.pi, np.pi, 19)
    s.output_far_field(0, theta, 0.0, "xz", tmp_path)
    s.output_far_field(0, np.pi / 2, phi, "xy", tmp_path)
    for sub, th, ph in (("xz", theta, np.zeros_like(theta)), ("xy", np.full_like(phi, np.pi / 2), phi)):
        d = tmp_path / sub
        assert np.array_equal(np.load(d / "theta.npy"), th)
        assert np.array_equal(np.load(d / "phi.npy"), ph)
        assert np.array_equal(np.load(d / "freq.npy"), [1e9, 2.5e9])
        et, ep = np.load(d / "e_theta.npy"), np.load(d / "e_phi.npy")
        assert et.dtype == np.complex128 and et.shape == (2, th.size) and ep.shape == et.shape
        for f in range(2):
            for n in range(th.size):
                a_t, a_p, _ = s.far_field(0, f, np.array([th[n]]), ph[n])
                assert et[f, n] == a_t[0] and ep[f, n] == a_p[0], (sub, f, n)
    # the drivers' post-processing: RCS from the files and the incident wave
    s.output_waveform(tmp_path)
    t, v = np.load(tmp_path / "time.npy"), np.load(tmp_path / "incident_wave.npy")
    inc = np.sum(v * s.dt * np.exp(-2j * np.pi * 1e9 * t))
    et = np.load(tmp_path / "xz" / "e_theta.npy")[0]
    ep = np.load(tmp_path / "xz" / "e_phi.npy")[0]
    rcs_files = 4 * np.pi * (np.abs(et / inc) ** 2 + np.abs(ep / inc) ** 2)
    _, _, rcs = s.far_field(0, 0, theta, 0.0, s.incident_spectrum(1e9))
    np.testing.assert_allclose(rcs_files, rcs, rtol=1e-12)


def test_waveform_files(run, tmp_path):
    s = run
    s.output_waveform(tmp_path / "w")
    t, v = s.waveform_samples()
    assert np.array_equal(np.load(tmp_path / "w" / "time.npy"), t)
    assert np.array_equal(np.load(tmp_path / "w" / "incident_wave.npy"), v)
    assert np.allclose(t, s.dt * np.arange(120))
    # every .npy header is padded so the data starts on a 64-byte boundary
    raw = (tmp_path / "w" / "time.npy").read_bytes()
    hlen = raw[8] | (raw[9] << 8)
    assert raw[:6] == b"\x93NUMPY" and (10 + hlen) % 64 == 0


def test_time_domain_far_field_on_disk(run, tmp_path):
    s = run
    s.output_far_field_td(0, tmp_path / "td")
    t, et, ep = s.far_field_td(0)
    assert np.array_equal(np.load(tmp_path / "td" / "time.npy"), t)
    assert np.array_equal(np.load(tmp_path / "td" / "e_theta.npy"), et)
    assert np.array_equal(np.load(tmp_path / "td" / "e_phi.npy"), ep)
    assert np.abs(et).max() > 0


def test_output_argument_errors(run, tmp_path):
    s = run
    with pytest.raises(xf.XFError, match="cut"):
        s.output_far_field(0, [0.1, 0.2], [0.0, 1.0], "bad", tmp_path)
    with pytest.raises(xf.XFError, match="index"):
        s.output_far_field(3, [0.1], 0.0, "bad", tmp_path)
    with pytest.raises(xf.XFError, match="index"):
        s.output_far_field_td(2, tmp_path)
    v = xf.Simulation(5e-3, 5e-3, 5e-3, 0.9)
    v.add_cube("domain", (0, 0, 0), (0.05,) * 3, xf.Material.air())
    v.init(5)
    with pytest.raises(xf.XFError, match="TFSF"):
        v.output_waveform(tmp_path)
