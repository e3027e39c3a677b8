"""Batched stepping of many independent grids (xf_batch_*) and the tile
autotuner, against the oracle run grid by grid.

The batched kernels run the same tile bodies as the single-grid kernels,
only scheduled from one persistent launch over the tiles of all grids, so
every grid of a FP64 batch must match its own oracle run bit for bit; FP32
batches must match single-grid FP32 engine runs bit for bit.
"""
import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes
from conftest import FIELDS, selection_weights

pytestmark = pytest.mark.gpu


def _uneven(steps, count=5, seed=3, layout="compact"):
    """Small grids of uneven sizes (odd extents, different CPML widths,
    one without TFSF, one without NF2FF)."""
    rng = np.random.default_rng(seed)
    sims = []
    for g in range(count):
        n = int(rng.integers(9, 24))
        pml = int(rng.integers(3, 6))
        t_in = 4 if n >= 12 else 3
        radius = float(rng.uniform(0.2, 0.35)) * n * 3e-3
        eps_r = float(rng.uniform(1.5, 8))
        if g != 1:
            s = scenes.dielectric_sphere(
                n=n, pml=pml, steps=steps, dl=3e-3, tfsf_inset=pml + t_in,
                nf2ff_inset=(pml + 1) if g != 2 else 0, radius=radius, eps_r=eps_r, mu_r=1.0,
                layout=layout, init=False)
        else:  # no TFSF: seed the field instead
            s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.9)
            h = n * 3e-3 / 2
            s.add_cube("domain", (-h, -h, -h), (n * 3e-3,) * 3, xf.Material.air())
            s.add_sphere("s", (0, 0, 0), 0.3 * n * 3e-3, xf.Material.dielectric(2.5, 1.0))
            s.add_pml_all(pml)
            s.add_nf2ff_fd((pml + 1,) * 3, [1e9])
            s.set_coefficient_layout(layout)
        s.init(steps)
        if g == 1:
            ez = s.array("ez")
            c = ez.shape[0] // 2
            ez[c - 1:c + 2, c - 1:c + 2, :] = 1.0
        sims.append(s)
    return sims


def _oracle_copies(build, steps):
    sims = build()
    for s in sims:
        oracle.run(s.scene_ptr(), steps)
    return sims


def _assert_identical(gpu_sims, ref_sims, label):
    for g, (a, b) in enumerate(zip(gpu_sims, ref_sims)):
        for n in FIELDS:
            assert np.array_equal(a.array(n), b.array(n)), f"{label}: grid {g} {n} differs"
        for p in range(6):
            try:
                pa = a.array(f"pml.{p}.ea_psi_hb")
            except xf.XFError:
                break
            assert np.array_equal(pa, b.array(f"pml.{p}.ea_psi_hb")), f"{label}: grid {g} psi {p}"
        try:
            ja = a.nf2ff_current(0, 0, "ZP", "ja")
        except xf.XFError:
            continue
        assert np.array_equal(ja, b.nf2ff_current(0, 0, "ZP", "ja")), f"{label}: grid {g} nf2ff"


def test_batch_matches_oracle_grid_by_grid(gpu_available):
    steps = 50
    build = lambda: _uneven(steps)
    ref = _oracle_copies(build, steps)
    sims = build()
    b = xf.Batch(sims).upload()
    b.run(20).run(steps - 20).download()
    desc = b.describe()
    b.close()
    assert "grids=5" in desc
    _assert_identical(sims, ref, "batch")
    assert np.abs(ref[0].array("ex")).max() > 1e-3


@pytest.mark.parametrize("tile_y,planes,blocks,nt,split,k32,xcd", [
    (1, 1, 0, True, False, False, False), (2, 3, 7, False, True, False, False),
    (4, 100, 1, True, False, False, False), (3, 2, 0, False, True, False, False),
    (1, 4, 13, True, True, False, False), (1, 4, 0, False, True, True, False),
    (4, 2, 5, True, False, True, False),
    # XCD-contiguous tile walk: resident grid, block counts not a multiple of
    # 8, fewer than 8 blocks (plain walk), and with 32-lane tiles
    (1, 4, 0, True, True, False, True), (2, 2, 13, True, True, False, True),
    (4, 8, 5, False, False, False, True), (1, 3, 0, True, True, True, True),
    # z-face CPML at the end of each tile (split = 2)
    (1, 4, 0, True, 2, False, False), (2, 3, 13, False, 2, True, True), (4, 16, 0, True, 2, True, False)])
def test_batch_tunings_bit_identical(gpu_available, tile_y, planes, blocks, nt, split, k32, xcd):
    steps = 30
    build = lambda: _uneven(steps, count=4, seed=11)
    ref = _oracle_copies(build, steps)
    sims = build()
    b = xf.Batch(sims).upload()
    b.set_tuning(tile_y=tile_y, planes=planes, blocks=blocks, nontemporal=nt, pml_split=bool(split), k32=k32,
                 xcd_remap=xcd, pml_ztail=split == 2)
    assert ("pml_split=0" not in b.describe()) == (split is True)
    assert ("pml_ztail=1" in b.describe()) == (split == 2)
    assert ("k32=1" in b.describe()) == k32
    assert ("xcd=1" in b.describe()) == xcd
    b.run(steps).download()
    b.close()
    _assert_identical(sims, ref, f"tuning ty={tile_y} planes={planes} blocks={blocks} split={split}")


@pytest.mark.parametrize("split", [1, 2])
def test_batch_cpsi_rows_mixed(gpu_available, split):
    """Grids whose CPML c_psi values all come from the coefficient table and
    grids with kappa-stretched faces that keep the c_psi arrays, in one
    batch, with the z faces split off or at the tile end: each grid
    bit-identical to the oracle."""
    steps = 40

    def build():
        sims = []
        for g in range(3):
            s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.95)
            s.set_coefficient_layout("dense")
            e = 0.03 + 0.006 * g
            s.add_cube("domain", (-e, -e, -e - 0.003 * g), (2 * e, 2 * e, 2 * e + 0.006 * g),
                       xf.Material.air())
            s.add_sphere("s", (0.0, 0.001, 0.0), 0.6 * e, xf.Material.dielectric(2.0 + g, 1.0, 0.01))
            for f in xf.FACES:
                s.add_pml(4, f, kappa_max=2.5 if (g == 1 and f in ("YP", "ZN")) else 1.0)
            tau = 20 * 3e-3 / 6e8
            s.add_tfsf3d((6, 6, 6), 0.4, 0.7, 0.1, xf.Waveform.gaussian(tau, 4.5 * tau))
            s.add_nf2ff_fd((5, 5, 5), [1.5e9])
            s.init(steps)
            sims.append(s)
        return sims

    ref = _oracle_copies(build, steps)
    sims = build()
    b = xf.Batch(sims).upload()
    b.set_tuning(tile_y=1, planes=3, pml_split=True, k32=True, pml_ztail=split == 2)
    desc = b.describe()
    assert "cpsi_tab=2" in desc, desc  # the kappa grid keeps its arrays
    b.run(15).run(steps - 15).download()
    b.close()
    _assert_identical(sims, ref, f"c_psi rows split={split}")


def test_batch_graph_replay_matches_eager(gpu_available):
    steps = 40
    build = lambda: _uneven(steps, count=4, seed=5)
    ref = _oracle_copies(build, steps)
    sims = build()
    b = xf.Batch(sims, use_graph=True).upload()
    # several graph runs with host work and synchronisation between them
    b.run(10)
    b.download()
    b.run(15).run(15).download()
    b.close()
    _assert_identical(sims, ref, "graph")


def test_batch_autotune_leaves_state_unchanged(gpu_available):
    steps = 40
    build = lambda: _uneven(steps + 20, count=3, seed=9)
    ref = build()
    for s in ref:
        oracle.run(s.scene_ptr(), steps)
    sims = build()
    b = xf.Batch(sims).upload()
    b.run(5)
    best = b.autotune(trial_steps=2)
    assert 1 <= best["tile_y"] <= 4 and best["planes"] >= 1 and best["step_ms"] > 0
    assert b.tuning()["tile_y"] == best["tile_y"]
    b.run(steps - 5).download()
    b.close()
    _assert_identical(sims, ref, "autotune")


def test_engine_autotune_leaves_state_unchanged(gpu_available):
    steps = 40
    build = lambda: scenes.dielectric_sphere(n=20, pml=5, steps=steps + 30, tfsf_inset=9,
                                             nf2ff_inset=7)
    c = build()
    oracle.run(c.scene_ptr(), steps)
    g = build()
    eng = xf.Engine(g).upload()
    eng.run(7)
    best = eng.autotune(trial_steps=2)
    assert best["step_ms"] > 0
    eng.run(steps - 7).download()
    eng.close()
    for n in FIELDS:
        assert np.array_equal(g.array(n), c.array(n)), n


@pytest.mark.parametrize("prec", ["fp32", "bf16c", "fp16"])
def test_batch_reduced_precision_matches_single_grid(gpu_available, prec):
    steps = 40
    build = lambda: _uneven(steps, count=4, seed=21)
    singles = build()
    for s in singles:
        e = xf.Engine(s, precision=prec).upload()
        e.run(steps).download()
        e.close()
    sims = build()
    b = xf.Batch(sims, precision=prec).upload()
    b.run(steps).download()
    b.close()
    _assert_identical(sims, singles, prec)


def test_batch_mixed_id_widths(gpu_available):
    """One grid needs 16-bit coefficient ids (> 256 distinct rows): the
    whole batch switches to 16-bit ids and stays exact."""
    steps = 25

    def build():
        sims = _uneven(steps, count=3, seed=2, layout="dense")
        s = sims[0]
        ce = s.array("coeff.0")  # cexe
        rng = np.random.default_rng(0)
        ce[...] = ce * (1.0 - 1e-4 * rng.integers(0, 300, ce.shape))
        return sims

    ref = _oracle_copies(build, steps)
    sims = build()
    b = xf.Batch(sims).upload()
    assert "id_bytes=2" in b.describe()
    b.run(steps).download()
    b.close()
    _assert_identical(sims, ref, "mixed ids")


def test_batch_with_learned_layers_sources_movies_and_time_domain(gpu_available):
    """Grids with a learned absorbing layer and point sources, one with a
    movie monitor, one with an NF2FF time-domain box and one recording
    training samples, stepped as one batch (point sources and learned layers
    of all grids in one launch per phase, the learned layers' predictor on
    the matrix cores, movies and time-domain boxes per grid) and under a
    hipGraph: every grid bit-identical to the oracle, the
    time-domain potentials within rel-L2 1e-12 (their tile sums are ordered
    differently from the oracle's serial adds)."""
    steps = 40
    w = selection_weights(5)  # exact matrix products: the batch stays bit-identical

    def build():
        sims = [scenes.point_source_box(n=14, boundary="learned", size=2, steps=steps, seed=1,
                                        weights=w),
                scenes.point_source_box(n=17, boundary="mur", size=2, steps=steps, seed=2)]
        s = scenes.point_source_box(n=12, boundary="cpml", size=4, steps=steps, seed=3, init=False)
        s.add_movie_monitor("ez", (0.0, 0.0, 0.03), (np.inf, np.inf, 5e-3), 3)
        s.init(steps)
        sims.append(s)
        sims.append(scenes.dielectric_sphere(n=16, pml=4, steps=steps, tfsf_inset=7, nf2ff_inset=5,
                                             freqs=(2e9,), td_dirs=[(0.6, 0.4)]))
        sims.append(scenes.point_source_box(n=15, boundary="cpml", size=6, steps=steps, seed=4,
                                            record=(3, 3, 20)))
        return sims

    ref = _oracle_copies(build, steps)
    for graph in (False, True):
        sims = build()
        b = xf.Batch(sims, use_graph=graph).upload()
        b.run(15).run(steps - 15).download()
        b.close()
        _assert_identical(sims, ref, f"additions graph={graph}")
        assert np.array_equal(sims[2].movie(0), ref[2].movie(0))
        assert np.abs(ref[2].movie(0)).max() > 0
        assert np.array_equal(sims[0].array("lb.history"), ref[0].array("lb.history"))
        assert np.array_equal(sims[4].lb_samples(), ref[4].lb_samples())
        assert np.abs(ref[4].lb_samples()).max() > 0
        for face in xf.FACES:
            for q in ("wa", "ub"):
                n = f"nf2ff_td.0.{face}.{q}"
                a, c = sims[3].array(n), ref[3].array(n)
                assert np.linalg.norm(a - c) <= 1e-12 * np.linalg.norm(c), n
