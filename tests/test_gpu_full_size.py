"""Parity at the benchmarked sizes: the 536^3 FP64 sphere (512^3 + 12-cell
CPML, TFSF, NF2FF) and the 64-grid batch (edges U[48, 160] + 8-cell CPML), a
few steps each from a random state, against the CPU oracle, bit for bit.

From a random state every path is live at once: every Yee cell, every CPML
psi (zero-initialised psi picks up the random curls in the first step), the
TFSF faces and the NF2FF accumulators.  The smaller parity tests cover long
runs; these cover the full-size layout (strides, tile counts, 32-bit psi
indices, the persistent batch over 64 grids).  About 40 s each, most of it
building the scenes and stepping the oracle.
"""
import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes
from conftest import FIELDS

pytestmark = pytest.mark.gpu


def _randomise(sims, seed):
    """The same random fields in every scene of each pair (pairs share the
    seed), magnitudes of order one."""
    rng = np.random.default_rng(seed)
    for name in FIELDS:
        shape = sims[0].array(name).shape
        v = rng.random(shape, dtype=np.float32).astype(np.float64) - 0.5
        for s in sims:
            s.array(name)[...] = v


def _state(sim):
    """Fields, every psi array and the first NF2FF box's currents."""
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
                sim.array(f"nf2ff.0.0.{face}.{c}")
            except xf.XFError:
                continue
            names.append(f"nf2ff.0.0.{face}.{c}")
    return names


def _assert_same(g, c, label):
    names = _state(c)
    for name in names:
        a, b = g.array(name), c.array(name)
        assert np.array_equal(a, b), f"{label}: {name} differs (max |diff| {np.abs(a - b).max():.3e})"
    return len(names)


@pytest.mark.parametrize("tiling", ["default", "split"])
def test_full_size_sphere_fp64_bit_identical(gpu_available, tiling):
    """The bench's 536^3 scene from a random state: the default FP64 tiling
    (z-face CPML at the end of each 3-plane, 32-lane tile) and the split
    z-face launch."""
    steps = 3
    print(f"case: 536^3 FP64 sphere, random state, 3 steps, {tiling} tiling", flush=True)
    g = scenes.large_sphere(n=512, pml=12, steps=steps + 1, layout="compact")
    c = scenes.large_sphere(n=512, pml=12, steps=steps + 1, layout="compact")
    _randomise([g, c], 21)
    e = xf.Engine(g).upload()
    if tiling == "split":
        e.set_tiling(tile_y=1, pml_split=True)
    desc = e.describe()
    e.run(1).run(steps - 1).download()
    e.close()
    oracle.run(c.scene_ptr(), steps)
    n = _assert_same(g, c, "536^3")
    if tiling == "split":
        assert "pml_split=1" in desc and "pml_ztail=0" in desc, desc
    else:
        assert "pml_ztail=1" in desc and "klanes=32" in desc and "chunk=3" in desc, desc
    assert n > 6 + 24
    assert np.abs(c.array("pml.5.ea_psi_hb")).max() > 0


def test_full_size_batch_fp64_bit_identical(gpu_available):
    steps = 2
    print("case: 64-grid batch FP64, random state, 2 steps", flush=True)
    gs = scenes.batch_spheres(count=64, steps=steps + 1)
    cs = scenes.batch_spheres(count=64, steps=steps + 1)
    for q, (g, c) in enumerate(zip(gs, cs)):
        _randomise([g, c], 100 + q)
    b = xf.Batch(gs, precision="fp64").upload()
    b.run(steps).download()
    b.close()
    for c in cs:
        oracle.run(c.scene_ptr(), steps)
    for q, (g, c) in enumerate(zip(gs, cs)):
        _assert_same(g, c, f"batch grid {q}")
