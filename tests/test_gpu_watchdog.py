"""The field watchdog (xf_plan.watchdog): a device check of every stored field
value every N steps; xf_run / xf_batch_run return XF_ERR_NUMERIC when one is
a NaN or an infinity.  The reference only checks for launch errors at the
end of a run (simulation_hd.cu:251-256).  The failures here are forced from
the host: a point source far beyond the 16-bit storage range (65504), so
the stored field overflows to infinity -- arithmetic on infinities, no fault.
"""
import numpy as np
import pytest

import oracle
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes
from conftest import FIELDS

pytestmark = pytest.mark.gpu


def _box(amp, steps=60, n=16):
    s = xf.Simulation(5e-3, 5e-3, 5e-3, 0.9)
    s.add_cube("domain", (0.0, 0.0, 0.0), (n * 5e-3,) * 3, xf.Material.air())
    s.add_pml_all(4)
    f = 4e9
    s.add_point_source(2, (12, 12, 12), xf.Waveform.cosine_modulated_gaussian(1 / f, 2 / f, f), amp)
    s.init(steps)
    return s


@pytest.mark.parametrize("graph", [False, True])
def test_fp16_overflow_is_reported(gpu_available, graph):
    print(f"case: watchdog, fp16 overflow, graph={graph}", flush=True)
    s = _box(1e7)
    e = xf.Engine(s, precision="fp16", watchdog=5, use_graph=graph).upload()
    with pytest.raises(xf.XFError, match=r"status 6\).*non-finite field value found after step \d+"):
        e.run(40)
    e.close()
    # the device is fine afterwards: a normal run on a new handle
    ok = _box(1.0)
    e2 = xf.Engine(ok, precision="fp16", watchdog=5).upload()
    e2.run(40).download()
    e2.close()
    assert np.isfinite(ok.array("ez")).all() and np.abs(ok.array("ez")).max() > 0


def test_first_bad_step_and_period(gpu_available):
    """The step named is the first checked step with a bad value: checks
    run every N steps and after the last step of a run."""
    s = _box(1e7)
    e = xf.Engine(s, precision="fp16", watchdog=1000).upload()
    with pytest.raises(xf.XFError, match=r"after step 37 "):
        e.run(37)  # only the end-of-run check (step 37) runs
    e.close()


def test_watchdog_leaves_results_unchanged(gpu_available):
    """FP64 with the watchdog on every 7 steps: still bit-identical to the
    oracle (the check only reads)."""
    steps = 60
    build = lambda: scenes.dielectric_sphere(n=20, pml=4, steps=steps, tfsf_inset=6, nf2ff_inset=5)
    g, c = build(), build()
    e = xf.Engine(g, watchdog=7).upload()
    e.run(30).run(30).download()
    e.close()
    oracle.run(c.scene_ptr(), steps)
    for n in FIELDS:
        assert np.array_equal(g.array(n), c.array(n)), n


def test_batch_names_the_grid(gpu_available):
    print("case: watchdog, batch of 3 with grid 1 overflowing", flush=True)
    sims = [_box(1.0), _box(1e7), _box(1.0)]
    b = xf.Batch(sims, precision="fp16", watchdog=10).upload()
    with pytest.raises(xf.XFError, match=r"status 6\).*grid 1 found after step \d+"):
        b.run(40)
    b.close()
