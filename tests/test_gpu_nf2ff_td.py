"""GPU <-> oracle parity of the NF2FF time-domain accumulation (k_nftd_tile +
k_nftd_reduce against oracle/xf_oracle.c nf2ff_td_face, after
src/nf2ff/time_domain/nf2ff_time_domain_data.cu:103-223).

The engine sums each time bin per 16x16 surface tile and then over the
tiles in a fixed order, where the reference (and the oracle) add every
contribution straight into the bin in (i, j, k) order; the sums therefore
agree to rounding, not bit for bit.  FP64 bar: rel-L2 <= 1e-12 per
potential array.  The prev arrays are plain copies of the face fields and
match bit for bit, as do the fields.  Runs are reproducible: two runs, and a
graph replay, give identical bits.
"""
import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes
from conftest import FIELDS, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12
POTS = ("wa", "wb", "ua", "ub")
PREVS = ("ea", "eb", "ha", "hb")


def _td_names(sim, boxes):
    names = []
    for q in range(boxes):
        for face in xf.FACES:
            try:
                sim.array(f"nf2ff_td.{q}.{face}.wa")
            except xf.XFError:
                continue
            names += [f"nf2ff_td.{q}.{face}.{p}" for p in POTS]
            names += [f"nf2ff_td.{q}.{face}.prev.{p}" for p in PREVS]
    return names


def _check(g, c, boxes, label, tol=TOL):
    worst = 0.0
    for n in _td_names(c, boxes):
        a, b = g.array(n), c.array(n)
        if ".prev." in n:
            assert np.array_equal(a, b), f"{label} {n} differs"
            continue
        err = rel_l2(a, b)
        worst = max(worst, err)
        assert err <= tol, f"{label} {n}: rel-L2 {err:.3e} (max|diff| {np.abs(a - b).max():.3e})"
    for n in FIELDS:
        assert np.array_equal(g.array(n), c.array(n)), f"{label} {n} not bit-identical"
    return worst


def _sphere(steps, dirs, n=24, pml=6, inset=7, tfsf=9):
    return scenes.dielectric_sphere(n=n, pml=pml, steps=steps, tfsf_inset=tfsf, nf2ff_inset=inset,
                                    freqs=(2e9,), td_dirs=dirs)


def test_td_matches_oracle(gpu_available):
    """Two observation directions on a 22x22-cell surface (2x2 tiles per face,
    ragged last tiles)."""
    dirs = [(0.7, 0.3), (2.2, 4.1)]
    g, c = _sphere(140, dirs), _sphere(140, dirs)
    e = xf.Engine(g).upload()
    e.run(140).download()
    e.close()
    oracle.run(c.scene_ptr(), 140)
    worst = _check(g, c, 2, "sphere")
    assert np.abs(c.array("nf2ff_td.0.ZP.ub")).max() > 0
    print(f"td sphere worst rel-L2 {worst:.2e}")


def test_td_uneven_grid_axis_directions_in_chunks(gpu_available):
    """Odd, non-cubic surface with observation along +z (a face whose cells
    all share one delay), along -x and oblique; three xf_run calls."""
    def build():
        s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.95)
        s.add_cube("domain", (-0.05, -0.04, -0.062), (0.1, 0.08, 0.124), xf.Material.air())
        s.add_sphere("s", (0.004, -0.003, 0.01), 0.021, xf.Material.dielectric(4.0, 1.5, 0.01))
        s.add_pml(5, "XN").add_pml(7, "XP").add_pml(4, "YN").add_pml(6, "YP")
        s.add_pml(5, "ZN").add_pml(5, "ZP")
        tau = 20 * 3e-3 / 6e8
        s.add_tfsf3d((9, 10, 11), 0.7, 0.4, 0.3, xf.Waveform.gaussian(tau, 4.5 * tau))
        s.add_nf2ff_td((7, 7, 8), 0.0, 0.0)
        s.add_nf2ff_td((7, 7, 8), np.pi / 2, np.pi)
        s.add_nf2ff_td((6, 8, 7), 1.1, 5.0)
        s.init(150)
        return s

    g, c = build(), build()
    e = xf.Engine(g).upload()
    e.run(40).run(60).run(50).download()
    e.close()
    oracle.run(c.scene_ptr(), 150)
    _check(g, c, 3, "uneven")


def test_td_long_face_many_tiles(gpu_available):
    """A 16 x 16 x 4116-cell slab: its x and y faces hold 258 tiles each
    (more than one 256-tile header chunk of the reduce) and span thousands of
    time bins along z (many reduce blocks per face)."""
    def build():
        dl = 1e-3
        s = xf.Simulation(dl, dl, dl, 0.9)
        s.add_cube("domain", (0, 0, 0), (20 * dl, 20 * dl, 4120 * dl), xf.Material.air())
        s.add_sphere("s", (10 * dl, 10 * dl, 2060 * dl), 4 * dl, xf.Material.dielectric(2.0, 1.0))
        tau = 20 * dl / 6e8
        s.add_tfsf3d((5, 5, 5), 0.3, 0.2, 0.0, xf.Waveform.gaussian(tau, 4.5 * tau))
        s.add_nf2ff_td((2, 2, 2), 0.05, 0.7)
        s.init(30)
        return s

    g, c = build(), build()
    e = xf.Engine(g).upload()
    e.run(30).download()
    e.close()
    oracle.run(c.scene_ptr(), 30)
    _check(g, c, 1, "slab")
    assert np.abs(c.array("nf2ff_td.0.XN.wb")).max() > 0


def test_td_reproducible_and_graph_replay(gpu_available):
    dirs = [(0.9, 1.3)]
    a, b, c = _sphere(90, dirs, n=20, pml=5, inset=6, tfsf=8), \
        _sphere(90, dirs, n=20, pml=5, inset=6, tfsf=8), _sphere(90, dirs, n=20, pml=5, inset=6, tfsf=8)
    for sim, graph in ((a, False), (b, False)):
        e = xf.Engine(sim, use_graph=graph).upload()
        e.run(90).download()
        e.close()
    e = xf.Engine(c, use_graph=True).upload()
    e.run(30).run(30).run(30).download()
    e.close()
    for n in _td_names(a, 1) + list(FIELDS):
        assert np.array_equal(a.array(n), b.array(n)), f"eager runs differ: {n}"
        assert np.array_equal(a.array(n), c.array(n)), f"graph replay differs: {n}"


def test_td_fp32_close_to_fp64(gpu_available):
    dirs = [(0.7, 0.3)]
    g, c = _sphere(140, dirs), _sphere(140, dirs)
    e = xf.Engine(g, precision="fp32").upload()
    e.run(140).download()
    e.close()
    oracle.run(c.scene_ptr(), 140)
    for face in xf.FACES:
        for p in POTS:
            n = f"nf2ff_td.0.{face}.{p}"
            assert rel_l2(g.array(n), c.array(n)) < 1e-4, n


def test_td_run_past_potentials_rejected(gpu_available):
    s = _sphere(20, [(0.7, 0.3)], n=16, pml=4, inset=5, tfsf=7)
    e = xf.Engine(s).upload()
    e.run(20)
    with pytest.raises(xf.XFError, match="past the end"):
        e.run(1)
    e.close()
