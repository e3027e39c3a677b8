"""GPU <-> CPU-oracle parity of the HIP time-stepping path (FP64).

Each case builds the same scene twice with the host set-up, advances one
copy with the HIP engine through the C ABI (or the SimulationHD drop-in) and
the other with the oracle (oracle/xf_oracle.c, the reference kernels
restated), then compares every state array the reference writes back:
fields, CPML psi, NF2FF currents, ADE J / E_prev.  The FP64 bar is
rel-L2 <= 1e-12 (the engine and oracle share evaluation order and neither
contracts to FMA, so in practice the arrays are bit-identical; the test also
reports max |diff|).
"""
import re

import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes
from conftest import FIELDS, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _state_names(sim):
    names = list(FIELDS)
    for p in range(6):
        try:
            sim.array(f"pml.{p}.ea_psi_hb")
        except xf.XFError:
            break
        names += [f"pml.{p}.{m}" for m in ("ea_psi_hb", "eb_psi_ha", "ha_psi_eb", "hb_psi_ea")]
    for q in range(4):
        for f in range(4):
            try:
                sim.array(f"nf2ff.{q}.{f}.transform_e")
            except xf.XFError:
                break
            for face in xf.FACES:
                for c in ("ja", "jb", "ma", "mb"):
                    try:
                        sim.array(f"nf2ff.{q}.{f}.{face}.{c}")
                    except xf.XFError:
                        continue
                    names.append(f"nf2ff.{q}.{f}.{face}.{c}")
    return names


def _compare(gpu_sim, cpu_sim, names=None, tol=TOL, label=""):
    names = names or _state_names(cpu_sim)
    worst = 0.0
    for n in names:
        a, b = gpu_sim.array(n), cpu_sim.array(n)
        if a.size == 0:
            continue
        err = rel_l2(a, b)
        worst = max(worst, err)
        assert err <= tol, f"{label} {n}: rel-L2 {err:.3e} > {tol:.1e} (max|diff| {np.abs(a - b).max():.3e})"
    return worst


def _run_engine(sim, steps, **kw):
    eng = xf.Engine(sim, **kw).upload()
    eng.run(steps).download()
    eng.close()


def _pair(build, steps, run_gpu=_run_engine, **kw):
    g, c = build(), build()
    run_gpu(g, steps, **kw)
    oracle.run(c.scene_ptr(), steps)
    return g, c


def test_sphere_pml_tfsf_nf2ff(gpu_available):
    build = lambda: scenes.dielectric_sphere(n=24, pml=6, steps=140, tfsf_inset=9, nf2ff_inset=7)
    g, c = _pair(build, 140)
    _compare(g, c, label="sphere")
    # the run must have produced non-trivial state everywhere we compare
    assert np.abs(c.array("ex")).max() > 0.1
    assert np.abs(c.array("pml.0.ea_psi_hb")).max() > 0
    assert np.abs(c.nf2ff_current(0, 0, "ZP", "ja")).max() > 0
    for n in FIELDS:
        assert np.array_equal(g.array(n), c.array(n)), f"{n} not bit-identical"


def test_sphere_uneven_sizes_and_steps_in_chunks(gpu_available):
    """Non-cubic grid with odd extents, run as 3 separate xf_run calls."""
    def build():
        s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.95)
        s.add_cube("domain", (-0.05, -0.04, -0.062), (0.1, 0.08, 0.124), xf.Material.air())
        s.add_sphere("s", (0.004, -0.003, 0.01), 0.021, xf.Material.dielectric(4.0, 1.5, 0.01))
        s.add_pml(5, "XN").add_pml(7, "XP").add_pml(4, "YN").add_pml(6, "YP")
        s.add_pml(5, "ZN").add_pml(5, "ZP")
        tau = 20 * 3e-3 / 6e8
        s.add_tfsf3d((9, 10, 11), 0.7, 0.4, 0.3, xf.Waveform.gaussian(tau, 4.5 * tau))
        s.add_nf2ff_fd((7, 7, 8), [1e9, 2.5e9])
        s.init(150)
        return s

    def run3(sim, steps):
        eng = xf.Engine(sim).upload()
        eng.run(40).run(60).run(steps - 100).download()
        eng.close()

    g, c = _pair(build, 150, run_gpu=run3)
    _compare(g, c, label="uneven")


@pytest.mark.parametrize("tile_y,chunks,xcd,nt,split,k32", [
    (1, 0, False, True, False, False), (2, 0, False, False, False, False),
    (3, 7, True, True, False, False), (4, 1, False, True, False, False),
    (1, 1000, True, False, False, False), (2, 3, False, True, False, False),
    (1, 0, False, True, True, False), (4, 1, False, False, True, False),
    (2, 3, True, True, True, False), (1, 0, False, True, True, True),
    (4, 5, False, False, True, True), (3, 7, True, True, False, True)])
def test_tiling_variants_bit_identical(gpu_available, tile_y, chunks, xcd, nt, split, k32):
    """Every tile shape (64 or 32 k-lanes) / march length / block order /
    streaming flag, and the z-face CPML corrections fused or in a launch of
    their own, gives the same bits: the tiling only reorders independent
    cell updates, and a node's z-face corrections come last either way."""
    build = lambda: scenes.dielectric_sphere(n=21, pml=5, steps=60, tfsf_inset=8, nf2ff_inset=6)
    g, c = build(), build()
    eng = xf.Engine(g).upload()
    eng.set_tiling(tile_y=tile_y, chunks_x=chunks, xcd_remap=xcd, nontemporal=nt, pml_split=split,
                   k32=k32)
    assert ("pml_split=1" in eng.describe()) == split
    assert ("klanes=32" in eng.describe()) == k32
    eng.run(60).download()
    eng.close()
    oracle.run(c.scene_ptr(), 60)
    for n in FIELDS:
        assert np.array_equal(g.array(n), c.array(n)), f"{n} differs (ty={tile_y} ch={chunks})"
    _compare(g, c, label=f"tiling ty={tile_y} ch={chunks}")


@pytest.mark.parametrize("tile_y,chunks,split,k32,graph", [
    (1, 0, True, False, False), (2, 3, False, False, False), (1, 1000, True, True, False),
    (4, 5, True, False, True)])
def test_pml_streams_bit_identical(gpu_available, tile_y, chunks, split, k32, graph):
    """CPML-free tiles and CPML tiles of each pass in two launches on two
    streams: the same bits as the oracle, eagerly and as a replayed graph."""
    build = lambda: scenes.dielectric_sphere(n=21, pml=5, steps=60, tfsf_inset=8, nf2ff_inset=6)
    g, c = build(), build()
    eng = xf.Engine(g, use_graph=graph).upload()
    eng.set_tiling(tile_y=tile_y, chunks_x=chunks, nontemporal=True, pml_split=split, k32=k32,
                   pml_streams=True)
    assert "pml_streams=1" in eng.describe()
    eng.run(20).run(40).download()
    eng.close()
    oracle.run(c.scene_ptr(), 60)
    for n in _state_names(c):
        assert np.array_equal(g.array(n), c.array(n)), f"{n} differs"


@pytest.mark.parametrize("kappa", [False, True])
def test_cpsi_from_table(gpu_available, kappa):
    """The CPML correction's c_psi read from the node's coefficient row
    where the engine found every value equal (the usual air layer: all 24
    psi arrays of six faces), and from the per-node array on faces whose
    kappa stretching makes the Yee coefficient differ: bit-identical to the
    oracle either way, in the default tiling and with the z faces split."""
    def build():
        s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.95)
        s.set_coefficient_layout("dense")
        s.add_cube("domain", (-0.04, -0.04, -0.05), (0.08, 0.08, 0.1), xf.Material.air())
        s.add_sphere("s", (0.002, 0.0, 0.004), 0.015, xf.Material.dielectric(3.0, 1.0, 0.02))
        for f in xf.FACES:
            s.add_pml(5, f, kappa_max=3.0 if kappa and f in ("XN", "ZP") else 1.0)
        tau = 20 * 3e-3 / 6e8
        s.add_tfsf3d((8, 8, 9), 0.5, 0.9, 0.2, xf.Waveform.gaussian(tau, 4.5 * tau))
        s.add_nf2ff_fd((6, 6, 7), [1.5e9])
        s.init(80)
        return s

    for split in (None, True):
        print(f"case: cpsi kappa={kappa} split={split}", flush=True)
        g, c = build(), build()
        e = xf.Engine(g, coeff="table").upload()
        if split:
            e.set_tiling(tile_y=1, pml_split=True)
        desc = e.describe()
        n = int(re.search(r"cpsi_tab=(\d+)", desc).group(1))
        assert (n < 24 and n > 0) if kappa else n == 24, desc
        e.run(80).download()
        e.close()
        oracle.run(c.scene_ptr(), 80)
        for name in _state_names(c):
            assert np.array_equal(g.array(name), c.array(name)), f"{name} differs (kappa={kappa})"


def test_pml_streams_dispersive(gpu_available):
    """The two-stream E pass with ADE cells (their Yee update skipped, the
    ADE kernel after the join)."""
    build = _dispersive("mix")
    g, c = build(), build()
    eng = xf.Engine(g).upload()
    eng.set_tiling(tile_y=1, nontemporal=True, pml_split=True, pml_streams=True)
    eng.run(160).download()
    eng.close()
    oracle.run(c.scene_ptr(), 160)
    _compare(g, c, label="pml streams ade")


def test_second_upload_resets_device_state(gpu_available):
    """xf_upload twice on one handle (what a batch does when it widens the
    coefficient ids): the second upload frees and rebuilds every device
    buffer and the run stays exact."""
    build = lambda: scenes.dielectric_sphere(n=20, pml=5, steps=50, tfsf_inset=8, nf2ff_inset=6)
    g, c = build(), build()
    eng = xf.Engine(g).upload()
    eng.upload()
    eng.run(50).download()
    eng.close()
    oracle.run(c.scene_ptr(), 50)
    for n in FIELDS:
        assert np.array_equal(g.array(n), c.array(n)), n


def test_oblique_incidence_all_quadrants(gpu_available):
    """TFSF with incident directions in every octant sign pattern."""
    for theta, phi in ((2.3, 4.0), (0.4, 2.2), (1.9, 5.5)):
        def build(theta=theta, phi=phi):
            s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.9)
            s.add_cube("domain", (-0.045,) * 3, (0.09,) * 3, xf.Material.air())
            s.add_pml_all(5)
            tau = 20 * 3e-3 / 6e8
            s.add_tfsf3d((8, 8, 8), theta, phi, 0.5, xf.Waveform.gaussian(tau, 4.5 * tau))
            s.init(100)
            return s
        g, c = _pair(build, 100)
        _compare(g, c, label=f"oblique {theta},{phi}")


def test_compact_layout_matches_dense(gpu_available):
    dense = scenes.dielectric_sphere(n=20, pml=5, steps=100, tfsf_inset=8, nf2ff_inset=6)
    comp = scenes.dielectric_sphere(n=20, pml=5, steps=100, tfsf_inset=8, nf2ff_inset=6,
                                    layout="compact")
    cpu = scenes.dielectric_sphere(n=20, pml=5, steps=100, tfsf_inset=8, nf2ff_inset=6,
                                   layout="compact")
    _run_engine(dense, 100)
    _run_engine(comp, 100)
    oracle.run(cpu.scene_ptr(), 100)
    for n in FIELDS:
        assert np.array_equal(dense.array(n), comp.array(n)), n
    _compare(comp, cpu, label="compact")


def test_graph_replay_matches_eager(gpu_available):
    build = lambda: scenes.dielectric_sphere(n=20, pml=5, steps=90, tfsf_inset=8, nf2ff_inset=6)
    a, b = build(), build()
    ea = xf.Engine(a, use_graph=True).upload()
    # several graph replays with eager work and syncs between them
    ea.run(30).run(30).run(30).download()
    ea.close()
    _run_engine(b, 90)
    for n in _state_names(b):
        assert np.array_equal(a.array(n), b.array(n)), n


def test_simulation_hd_dropin(gpu_available):
    """The reference's calling sequence: SimulationHD(&s); setGridDim;
    setBlockDim; run(n) -- through the C++ class."""
    build = lambda: scenes.dielectric_sphere(n=22, pml=6, steps=110, tfsf_inset=9, nf2ff_inset=7,
                                             init=False)
    g, c = build(), build()
    hd = xf.SimulationHD(g)
    hd.setGridDim((4, 4, 4))
    hd.setBlockDim((8, 8, 8))
    hd.run(110)
    c.init(110)
    oracle.run(c.scene_ptr(), 110)
    _compare(g, c, label="SimulationHD")


def test_cylinder_2d(gpu_available):
    build = lambda: scenes.cylinder_2d(n=80, pml=8, steps=200, tfsf_inset=14, nf2ff_inset=11)
    g, c = _pair(build, 200)
    _compare(g, c, label="2d")
    assert np.abs(c.array("ez")).max() > 0.1
    # Ex, Ey, Hz stay zero in the TE updator
    for n in ("ex", "ey", "hz"):
        assert np.abs(g.array(n)).max() == 0


def _dispersive(kind):
    def build():
        dl = 3e-3
        s = xf.Simulation(dl, dl, dl, 0.9)
        s.add_cube("domain", (-0.042,) * 3, (0.084,) * 3, xf.Material.air())
        tp = 2 * np.pi
        if kind == "drude":
            m = xf.Material.drude(1.5, [tp * 2e9], [tp * 0.2e9])
        elif kind == "debye":
            m = xf.Material.debye(2.0, [3.0, 1.0], [1e-10, 3e-11])
        else:
            m = xf.Material.lorentz(1.2, [2.0], [tp * 1.5e9], [tp * 0.1e9])
        s.add_sphere("shell", (0, 0, 0), 0.02, m)
        if kind == "mix":
            s.add_sphere("core", (0, 0, 0), 0.01, xf.Material.drude(1.0, [tp * 2e9], [tp * 0.2e9]))
        s.add_pml_all(5)
        tau = 20 * dl / 6e8
        s.add_tfsf3d((8, 8, 8), 0.0, 0.0, 0.0, xf.Waveform.gaussian(tau, 4.5 * tau))
        s.add_nf2ff_fd((6, 6, 6), [1e9])
        s.init(160)
        return s
    return build


@pytest.mark.parametrize("kind", ["drude", "debye", "lorentz", "mix"])
def test_ade(gpu_available, kind):
    build = _dispersive(kind)
    g, c = _pair(build, 160)
    _compare(g, c, label=kind)
    names = ["ade.jx", "ade.jy", "ade.jz"]
    if kind in ("lorentz", "mix"):
        names += ["ade.jx_prev", "ade.jz_prev"]
    for n in names:
        # E_prev and J of non-dispersive cells are not tracked on the device
        # (multiplied by zero coefficients); J stays exactly zero there.
        assert rel_l2(g.array(n), c.array(n)) <= TOL, n
    assert np.abs(c.array("ade.jx")).max() > 0


def test_ade_coefficient_table(gpu_available):
    """The ADE kernel reads each dispersive cell's coefficients from a table
    of the distinct rows (a row id per cell); the dense coefficient mode
    keeps the per-cell arrays.  Both equal the oracle."""
    build = _dispersive("mix")
    steps = 120
    c = build()
    oracle.run(c.scene_ptr(), steps)
    for mode in ("auto", "dense"):
        print(f"case: ADE coefficients, coeff={mode}", flush=True)
        g = build()
        e = xf.Engine(g, coeff=mode).upload()
        desc = e.describe()
        e.run(steps).download()
        e.close()
        rows = int(re.search(r"ade_rows=(-?\d+)", desc).group(1))
        assert (rows > 0) if mode == "auto" else (rows == -1), desc
        _compare(g, c, label=f"ade {mode}")
        for n in ("ade.jx", "ade.jy", "ade.jz", "ade.jx_prev"):
            assert rel_l2(g.array(n), c.array(n)) <= TOL, (mode, n)


def test_pml_split_uneven_faces_fp32_and_bf16(gpu_available):
    """The z-face split on a grid with a different CPML thickness on every
    face and three chunked runs: FP64 bit-identical to the oracle, FP32
    bit-identical to the fused FP32 run; BF16 storage keeps the corrections
    fused (the split would round the intermediate field)."""
    def build():
        s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.95)
        s.add_cube("domain", (-0.05, -0.04, -0.062), (0.1, 0.08, 0.124), xf.Material.air())
        s.add_sphere("s", (0.004, -0.003, 0.01), 0.021, xf.Material.dielectric(4.0, 1.5, 0.01))
        s.add_pml(5, "XN").add_pml(7, "XP").add_pml(4, "YN").add_pml(6, "YP")
        s.add_pml(3, "ZN").add_pml(9, "ZP")
        tau = 20 * 3e-3 / 6e8
        s.add_tfsf3d((11, 9, 12), 0.7, 0.4, 0.3, xf.Waveform.gaussian(tau, 4.5 * tau))
        s.add_nf2ff_fd((9, 8, 10), [1.5e9])
        s.init(120)
        return s

    def run(sim, precision, split):
        e = xf.Engine(sim, precision=precision).upload()
        e.set_tiling(tile_y=2, pml_split=split)
        desc = e.describe()
        e.run(30).run(50).run(40).download()
        e.close()
        return desc

    g, c = build(), build()
    assert "pml_split=1" in run(g, "fp64", True)
    oracle.run(c.scene_ptr(), 120)
    _compare(g, c, label="split fp64")
    a, b = build(), build()
    run(a, "fp32", True)
    run(b, "fp32", False)
    for n in FIELDS:
        assert np.array_equal(a.array(n), b.array(n)), n
    assert "pml_split=0" in run(build(), "bf16", True)


def test_pml_split_single_z_face(gpu_available):
    """The z-face split with CPML on XN and ZP only (one z face in the
    launch; PEC elsewhere): bit-identical to the oracle, fields and psi."""
    def build():
        s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.95)
        s.add_cube("domain", (-0.04, -0.04, -0.05), (0.08, 0.08, 0.1), xf.Material.air())
        s.add_sphere("s", (0.002, 0.0, 0.004), 0.015, xf.Material.dielectric(3.0, 1.0, 0.02))
        s.add_pml(5, "XN").add_pml(6, "ZP")
        tau = 20 * 3e-3 / 6e8
        s.add_tfsf3d((8, 8, 9), 0.5, 0.9, 0.2, xf.Waveform.gaussian(tau, 4.5 * tau))
        s.init(100)
        return s

    g, c = build(), build()
    e = xf.Engine(g).upload()
    e.set_tiling(tile_y=1, pml_split=True)
    assert "pml_split=1" in e.describe()
    e.run(100).download()
    e.close()
    oracle.run(c.scene_ptr(), 100)
    _compare(g, c, label="split one z face")
    assert np.abs(c.array("pml.1.ea_psi_hb")).max() > 0


def _uneven_faces(steps):
    s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.95)
    s.add_cube("domain", (-0.05, -0.04, -0.062), (0.1, 0.08, 0.124), xf.Material.air())
    s.add_sphere("s", (0.004, -0.003, 0.01), 0.021, xf.Material.dielectric(4.0, 1.5, 0.01))
    s.add_pml(5, "XN").add_pml(7, "XP").add_pml(4, "YN").add_pml(6, "YP")
    s.add_pml(3, "ZN").add_pml(9, "ZP")
    tau = 20 * 3e-3 / 6e8
    s.add_tfsf3d((11, 9, 12), 0.7, 0.4, 0.3, xf.Waveform.gaussian(tau, 4.5 * tau))
    s.add_nf2ff_fd((9, 8, 10), [1.5e9])
    s.init(steps)
    return s


@pytest.mark.parametrize("scene,tile_y,chunks,k32,streams", [
    ("sphere", 1, 0, False, False), ("sphere", 4, 1, False, False), ("sphere", 2, 7, True, False),
    ("sphere", 3, 1000, False, True), ("uneven", 2, 5, False, False), ("uneven", 1, 0, True, True),
    ("one_z", 1, 3, False, False)])
def test_pml_ztail_bit_identical(gpu_available, scene, tile_y, chunks, k32, streams):
    """The z-face CPML corrections applied at the end of each Yee tile by
    the slab columns' own lanes (XF_TILE_PML_ZTAIL): bit-identical to the
    oracle, fields and every psi array, for tile shapes and march lengths
    that put the z slabs at tile starts, inside tiles and across several
    k-tiles, with one or two z faces of different thicknesses."""
    steps = 90
    if scene == "sphere":
        build = lambda: scenes.dielectric_sphere(n=21, pml=5, steps=steps, tfsf_inset=8, nf2ff_inset=6)
    elif scene == "uneven":
        build = lambda: _uneven_faces(steps)
    else:
        def build():
            s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.95)
            s.add_cube("domain", (-0.04, -0.04, -0.05), (0.08, 0.08, 0.1), xf.Material.air())
            s.add_sphere("s", (0.002, 0.0, 0.004), 0.015, xf.Material.dielectric(3.0, 1.0, 0.02))
            s.add_pml(5, "XN").add_pml(6, "ZP")
            tau = 20 * 3e-3 / 6e8
            s.add_tfsf3d((8, 8, 9), 0.5, 0.9, 0.2, xf.Waveform.gaussian(tau, 4.5 * tau))
            s.init(steps)
            return s
    print(f"case: ztail {scene} ty={tile_y} ch={chunks} k32={k32} streams={streams}", flush=True)
    g, c = build(), build()
    e = xf.Engine(g).upload()
    e.set_tiling(tile_y=tile_y, chunks_x=chunks, pml_ztail=True, k32=k32, pml_streams=streams)
    desc = e.describe()
    assert "pml_ztail=1" in desc and "pml_split=0" in desc, desc
    e.run(35).run(steps - 35).download()
    e.close()
    oracle.run(c.scene_ptr(), steps)
    for n in _state_names(c):
        assert np.array_equal(g.array(n), c.array(n)), f"{n} differs ({scene} ty={tile_y} ch={chunks})"


def test_pml_ztail_fp32_and_bf16(gpu_available):
    """FP32 with the tail equals FP32 with the z faces in the march, bit for
    bit; BF16 storage keeps the z faces in the march."""
    def run(sim, precision, tail):
        e = xf.Engine(sim, precision=precision).upload()
        e.set_tiling(tile_y=2, pml_split=False, pml_ztail=tail)
        desc = e.describe()
        e.run(40).run(40).download()
        e.close()
        return desc

    a, b = _uneven_faces(80), _uneven_faces(80)
    assert "pml_ztail=1" in run(a, "fp32", True)
    assert "pml_ztail=0" in run(b, "fp32", False)
    for n in FIELDS:
        assert np.array_equal(a.array(n), b.array(n)), n
    assert "pml_ztail=0" in run(_uneven_faces(80), "bf16", True)


@pytest.mark.parametrize("scene,tile_y,chunks,k32,streams", [
    ("sphere", 1, 0, True, False), ("sphere", 4, 1, False, False), ("sphere", 2, 7, True, False),
    ("sphere", 3, 1000, False, False), ("uneven", 2, 5, False, False), ("uneven", 1, 0, True, False),
    ("one_z", 1, 3, True, False), ("uneven", 1, 0, True, True)])
def test_tfsf_ztail_bit_identical(gpu_available, scene, tile_y, chunks, k32, streams):
    """The TF/SF corrections of the z = const faces applied by the Yee tiles
    after their march (XF_TILE_TFSF_ZTAIL), the face edges that other faces
    also correct left to k_tfsf: bit-identical to the oracle, fields, psi
    and NF2FF currents, for tile shapes and march lengths that put the TF/SF
    z faces inside and at the ends of k-tiles and marches.  With the two-
    stream CPML split the tail stays off."""
    steps = 90
    if scene == "sphere":
        build = lambda: scenes.dielectric_sphere(n=21, pml=5, steps=steps, tfsf_inset=8, nf2ff_inset=6)
    elif scene == "uneven":
        build = lambda: _uneven_faces(steps)
    else:
        def build():
            s = xf.Simulation(3e-3, 3e-3, 3e-3, 0.95)
            s.add_cube("domain", (-0.04, -0.04, -0.05), (0.08, 0.08, 0.1), xf.Material.air())
            s.add_sphere("s", (0.002, 0.0, 0.004), 0.015, xf.Material.dielectric(3.0, 1.0, 0.02))
            s.add_pml(5, "XN").add_pml(6, "ZP")
            tau = 20 * 3e-3 / 6e8
            s.add_tfsf3d((8, 8, 9), 0.5, 0.9, 0.2, xf.Waveform.gaussian(tau, 4.5 * tau))
            s.init(steps)
            return s
    print(f"case: tfsf ztail {scene} ty={tile_y} ch={chunks} k32={k32} streams={streams}", flush=True)
    g, c = build(), build()
    e = xf.Engine(g).upload()
    e.set_tiling(tile_y=tile_y, chunks_x=chunks, pml_ztail=True, tfsf_ztail=True, k32=k32,
                 pml_streams=streams)
    desc = e.describe()
    assert f"tfsf_ztail={0 if streams else 1}" in desc, desc
    e.run(35).run(steps - 35).download()
    e.close()
    oracle.run(c.scene_ptr(), steps)
    for n in _state_names(c):
        assert np.array_equal(g.array(n), c.array(n)), f"{n} differs ({scene} ty={tile_y} ch={chunks})"


def test_tfsf_ztail_fp32_graph_and_ade(gpu_available):
    """FP32 with the TF/SF z-face tail equals FP32 without it, bit for bit,
    stepped eagerly and as hipGraph replays (the tail reads the step from the
    device counter there); a scene with ADE media keeps the tail off."""
    def run(sim, tail, graph=False):
        e = xf.Engine(sim, precision="fp32", use_graph=graph).upload()
        e.set_tiling(tile_y=1, k32=True, pml_ztail=True, tfsf_ztail=tail)
        desc = e.describe()
        e.run(40).run(40).download()
        e.close()
        return desc

    a, b, c = _uneven_faces(80), _uneven_faces(80), _uneven_faces(80)
    assert "tfsf_ztail=1" in run(a, True)
    assert "tfsf_ztail=0" in run(b, False)
    assert "tfsf_ztail=1" in run(c, True, graph=True)
    for n in FIELDS:
        assert np.array_equal(a.array(n), b.array(n)), n
        assert np.array_equal(c.array(n), b.array(n)), n
    d = _dispersive("mix")()
    e = xf.Engine(d).upload()
    e.set_tiling(tile_y=1, k32=True, pml_ztail=True, tfsf_ztail=True)
    assert "tfsf_ztail=0" in e.describe()
    e.close()


def test_fp32_close_to_fp64(gpu_available):
    build = lambda: scenes.dielectric_sphere(n=24, pml=6, steps=140, tfsf_inset=9, nf2ff_inset=7)
    g, c = _pair(build, 140, precision="fp32")
    for n in ("ex", "ey", "ez", "hy"):
        assert rel_l2(g.array(n), c.array(n)) < 1e-4, n
    assert rel_l2(g.array("nf2ff.0.0.XP.jb"), c.array("nf2ff.0.0.XP.jb")) < 1e-4


def test_ade_fp32_close_to_fp64(gpu_available):
    """FP32 fields with FP64 pole recursions (the dispersive sphere's mixed
    precision): Drude core + Lorentz shell against the FP64 oracle."""
    build = _dispersive("mix")
    g, c = _pair(build, 160, precision="fp32")
    for n in ("ex", "ey", "ez", "hx"):
        assert rel_l2(g.array(n), c.array(n)) < 1e-4, n
    for n in ("ade.jx", "ade.jz"):
        assert rel_l2(g.array(n), c.array(n)) < 1e-4, n
    assert rel_l2(g.array("nf2ff.0.0.ZP.ma"), c.array("nf2ff.0.0.ZP.ma")) < 1e-4


def test_cylinder_2d_fp32(gpu_available):
    build = lambda: scenes.cylinder_2d(n=80, pml=8, steps=200, tfsf_inset=14, nf2ff_inset=11)
    g, c = _pair(build, 200, precision="fp32")
    for n in ("ez", "hx", "hy"):
        assert rel_l2(g.array(n), c.array(n)) < 1e-4, n
    assert rel_l2(g.array("nf2ff.0.0.XP.ma"), c.array("nf2ff.0.0.XP.ma")) < 1e-4


def test_compensated_bf16_storage_close_to_fp64(gpu_available):
    """bf16 + 8-bit residual storage (XF_BF16_COMP): about 16 significant
    bits per stored value, so it lands between bf16 storage and FP32."""
    build = lambda: scenes.dielectric_sphere(n=24, pml=6, steps=140, tfsf_inset=9, nf2ff_inset=7)
    errs = {}
    for prec in ("bf16c", "bf16", "fp32"):
        g, c = _pair(build, 140, precision=prec)
        errs[prec] = max(rel_l2(g.array(n), c.array(n)) for n in ("ex", "ey", "ez", "hy"))
    assert errs["bf16c"] < 2e-3, errs
    assert errs["bf16c"] < errs["bf16"] / 30, errs
    assert errs["fp32"] < errs["bf16c"], errs


def test_fp16_storage_between_bf16_and_compensated(gpu_available):
    """IEEE fp16 storage (XF_FP16_STORE, H scaled by the free-space
    impedance): 11 significant bits against bf16's 8, so well inside bf16's
    error and outside the compensated format's 16 bits."""
    build = lambda: scenes.dielectric_sphere(n=24, pml=6, steps=140, tfsf_inset=9, nf2ff_inset=7)
    errs = {}
    for prec in ("fp16", "bf16", "bf16c"):
        g, c = _pair(build, 140, precision=prec)
        errs[prec] = max(rel_l2(g.array(n), c.array(n)) for n in ("ex", "ey", "ez", "hy"))
    print(errs)
    assert np.isfinite(errs["fp16"])
    assert errs["fp16"] < errs["bf16"] / 4, errs
    assert errs["bf16c"] < errs["fp16"], errs


def test_bf16_storage_runs(gpu_available):
    build = lambda: scenes.dielectric_sphere(n=24, pml=6, steps=140, tfsf_inset=9, nf2ff_inset=7)
    g, c = _pair(build, 140, precision="bf16")
    err = rel_l2(g.array("ex"), c.array("ex"))
    assert np.isfinite(err) and err < 0.2


def test_config0_geometry(gpu_available):
    """The 156^3 dielectric-sphere driver scene (140^3 + 8-cell CPML), first
    150 of its 2400 steps (the oracle finishes these in seconds)."""
    build = lambda: scenes.dielectric_sphere(steps=2400)
    g, c = build(), build()
    eng = xf.Engine(g).upload()
    eng.run(150).download()
    eng.close()
    oracle.run(c.scene_ptr(), 150)
    _compare(g, c, names=list(FIELDS) + ["pml.0.ea_psi_hb", "pml.5.hb_psi_ea",
                                         "nf2ff.0.0.ZN.ja", "nf2ff.0.0.XP.mb"], label="cfg0")


def test_cpp_dropin_driver(gpu_available, tmp_path):
    """tests/cpp/simulation_hd_driver.cpp: the reference driver's C++
    calling sequence against xfdtd::cuda::SimulationHD; its Ex and far field
    match the oracle run of the same scene."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "simulation_hd_driver")
    out = tmp_path / "drv.bin"
    r = subprocess.run([exe, "20", "90", str(out), str(tmp_path / "data")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.int64, count=4)
    nx, ny, nz, steps = (int(v) for v in raw)
    data = np.fromfile(out, dtype=np.float64)[4:]
    ex = data[: nx * (ny + 1) * (nz + 1)].reshape(nx, ny + 1, nz + 1)
    rcs = data[nx * (ny + 1) * (nz + 1):]
    edge = 20 * 5e-3
    c = scenes.dielectric_sphere(n=20, pml=6, steps=90, dl=5e-3, tfsf_inset=9, nf2ff_inset=7,
                                 radius=edge * 0.1 / 0.35)
    oracle.run(c.scene_ptr(), 90)
    assert steps == 90
    assert rel_l2(ex, c.array("ex")) <= TOL
    theta = -np.pi + 2 * np.pi * np.arange(360) / 359.0
    _, _, rcs_c = c.far_field(0, 0, theta, 0.0, c.incident_spectrum(1e9))
    assert rel_l2(rcs, rcs_c) <= 1e-10
    # the driver's files: far-field cuts, time axis and incident wave
    d = tmp_path / "data"
    t, v = c.waveform_samples()
    assert np.array_equal(np.load(d / "time.npy"), t)
    assert np.array_equal(np.load(d / "incident_wave.npy"), v)
    for sub, phi in (("xz", 0.0), ("yz", np.pi / 2)):
        et_c, ep_c, _ = c.far_field(0, 0, theta, phi)
        assert np.array_equal(np.load(d / "fd" / sub / "theta.npy"), theta)
        for name, ref in (("e_theta", et_c), ("e_phi", ep_c)):
            got = np.load(d / "fd" / sub / f"{name}.npy")
            assert got.dtype == np.complex128 and got.shape == (1, 360)
            assert np.linalg.norm(got[0] - ref) <= 1e-10 * np.linalg.norm(ref), (sub, name)


def test_movie_monitors_match_oracle(gpu_available):
    """Movie frames (reference MovieMonitor, movie_monitor.cu:17-45) of an E
    and an H component in 3-D and of Ez in 2-D, run in two xf_run calls."""
    def build3():
        s = scenes.dielectric_sphere(n=20, pml=4, steps=70, dl=5e-3, tfsf_inset=7, nf2ff_inset=5,
                                     init=False)
        h = 20 * 5e-3 / 2
        s.add_movie_monitor("ex", (-h, 0.0, -h), (2 * h, 5e-3, 2 * h), 10)
        s.add_movie_monitor("hy", (-h, -h, 0.0), (2 * h, 2 * h, 5e-3), 7)
        s.init(70)
        return s

    def build2():
        s = scenes.cylinder_2d(n=60, pml=6, steps=60, dl=2e-3, tfsf_inset=10, nf2ff_inset=8,
                               init=False)
        s.add_movie_monitor("ez", (-0.06, -0.06, -np.inf), (0.12, 0.12, np.inf), 15)
        s.init(60)
        return s

    for build, steps in ((build3, 70), (build2, 60)):
        g, c = build(), build()
        e = xf.Engine(g).upload()
        e.run(steps // 3).run(steps - steps // 3).download()
        e.close()
        oracle.run(c.scene_ptr(), steps)
        q = 0
        while True:
            try:
                a = g.movie(q)
            except xf.XFError:
                break
            assert np.array_equal(a, c.movie(q)), f"movie {q}"
            assert np.abs(a).max() > 0
            q += 1
        assert q >= 1
