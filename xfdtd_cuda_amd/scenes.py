"""The reference drivers' scenes as builders of Simulation objects.

dielectric_sphere   test/dielectric_sphere_scatter.cu:20-107 at dl = 2.5e-3:
                    a 0.35 m vacuum cube (140^3 cells) + 8-cell CPML on six
                    faces, a r = 0.1 m sphere (eps_r 3, mu_r 2), a Gaussian
                    TFSF plane wave (tau = 20 dl / 6e8, t0 = 4.5 tau) at a
                    15-cell inset, a 1 GHz NF2FF box at an 11-cell inset.
vacuum_benchmark    test/benchmark.cu:105-163 (type 0..4).
cylinder_2d         test/cylinder_scatter_2d.cu:28-183 in the 1000 x 1000,
                    dl = 1e-3 form: PEC cylinder r = 0.1 m, 10-cell CPML.
large_sphere        the dielectric sphere scaled to n^3 cells (+ pml), sphere
                    radius 0.35 * edge, eps_r ~ U[2, 6] from a seed.
batch_spheres       64 sphere scatterers of uneven sizes for a batched run.
point_source_box    vacuum cube with seeded point sources and a chosen boundary
                    (CPML, learned layer, Mur, PEC): the learned layer's data
                    and its judge.
dispersive_sphere   Drude core + Lorentz shell, n^3 cells + 10-cell CPML.

Insets the reference does not fix (2-D NF2FF/TFSF, scaled boxes) are chosen
to keep the CPML < NF2FF < TFSF order of the 3-D driver.
"""
from __future__ import annotations

import numpy as np

from .simulation import FACES, Material, Simulation, Waveform


def dielectric_sphere(n: int = 140, pml: int = 8, steps: int = 2400, dl: float = 2.5e-3,
                      tfsf_inset: int = 15, nf2ff_inset: int = 11, freqs=(1e9,),
                      radius: float | None = None, eps_r: float = 3.0, mu_r: float = 2.0,
                      layout: str = "dense", init: bool = True, td_dirs=(),
                      lb_weights=None) -> Simulation:
    """n interior cells per edge (140 reproduces the driver's 0.35 m cube);
    td_dirs: (theta, phi) observation directions of NF2FF time-domain boxes on
    the frequency-domain box's surface; lb_weights: a `pml`-cell learned
    absorbing layer with these fp16 weights instead of the CPML."""
    edge = n * dl
    half = edge / 2
    radius = (0.1 / 0.35) * edge if radius is None else radius
    s = Simulation(dl, dl, dl, 0.9)
    s.add_cube("domain", (-half, -half, -half), (edge, edge, edge), Material.air())
    s.add_sphere("dielectric_sphere", (0, 0, 0), radius, Material.dielectric(eps_r, mu_r))
    tau = (dl * 20) / 6e8
    s.add_tfsf3d((tfsf_inset,) * 3, 0.0, 0.0, 0.0, Waveform.gaussian(tau, 4.5 * tau))
    if pml > 0 and lb_weights is not None:
        s.add_learned_boundary(lb_weights, thickness=pml)
    elif pml > 0:
        s.add_pml_all(pml)
    if nf2ff_inset > 0 and freqs:
        s.add_nf2ff_fd((nf2ff_inset,) * 3, list(freqs))
    for theta, phi in td_dirs:
        s.add_nf2ff_td((nf2ff_inset,) * 3, theta, phi)
    s.set_coefficient_layout(layout)
    if init:
        s.init(steps)
    return s


def vacuum_benchmark(kind: int = 4, n: int = 128, dl: float = 2.5e-3, steps: int = 2400,
                     layout: str = "dense", init: bool = True) -> Simulation:
    """benchmark --type kind: 0 updator only, 1 +TFSF, 2 +PML, 3 +TFSF+PML,
    4 +TFSF+PML+NF2FF, on an n^3 vacuum cube (+8-cell PML when used)."""
    half = n * dl / 2
    s = Simulation(dl, dl, dl, 0.9)
    s.add_cube("domain", (-half, -half, -half), (n * dl,) * 3, Material.air())
    if kind >= 4:
        s.add_nf2ff_fd((11, 11, 11), [1e9])
    if kind in (2, 3, 4):
        s.add_pml_all(8)
    if kind in (1, 3, 4):
        tau = dl * 20 / 6e8
        s.add_tfsf3d((15, 15, 15), 0.0, 0.0, 0.0, Waveform.gaussian(tau, 4.5 * tau))
    s.set_coefficient_layout(layout)
    if init:
        s.init(steps)
    return s


def cylinder_2d(n: int = 1000, pml: int = 10, steps: int = 4000, dl: float = 1e-3,
                radius: float | None = None, tfsf_inset: int = 15, nf2ff_inset: int = 12,
                freqs=(1e9,), phi: float = 0.0, init: bool = True) -> Simulation:
    """2-D TMz PEC cylinder (n x n cells + pml on four faces)."""
    edge = n * dl
    half = edge / 2
    radius = 0.1 * edge if radius is None else radius
    s = Simulation(dl, dl, 1.0, 0.9)
    s.add_cube("domain", (-half, -half, -np.inf), (edge, edge, np.inf), Material.air())
    s.add_cylinder("cylinder", (0.0, 0.0, -np.inf), radius, np.inf, Material.pec())
    tau = 20 * dl / 6e8
    s.add_tfsf2d((tfsf_inset, tfsf_inset), phi, Waveform.gaussian(tau, 4.5 * tau))
    if pml > 0:
        s.add_pml_all(pml, faces=("XN", "XP", "YN", "YP"))
    if nf2ff_inset > 0 and freqs:
        s.add_nf2ff_fd((nf2ff_inset, nf2ff_inset, 0), list(freqs))
    if init:
        s.init(steps)
    return s


def large_sphere(n: int = 512, pml: int = 12, steps: int = 4000, seed: int = 2024,
                 dl: float = 2.5e-3, layout: str = "compact", freqs=(1e9,),
                 init: bool = True) -> Simulation:
    """The dielectric-sphere scene scaled to n^3 (+2*pml) cells; sphere radius
    0.35 * domain edge, eps_r drawn from U[2, 6] with numpy's default_rng(seed).
    TF/SF and NF2FF insets keep the 8 < 11 < 15 gaps of the 140^3 driver."""
    eps_r = float(np.random.default_rng(seed).uniform(2.0, 6.0))
    edge = n * dl
    return dielectric_sphere(n=n, pml=pml, steps=steps, dl=dl, tfsf_inset=pml + 7,
                             nf2ff_inset=pml + 3, freqs=freqs, radius=0.35 * edge,
                             eps_r=eps_r, mu_r=1.0, layout=layout, init=init)


def batch_spheres(count: int = 64, seed: int = 7, edge_range=(48, 160), pml: int = 8,
                  steps: int = 2000, dl: float = 2.5e-3, freqs=(1e9,), layout: str = "compact",
                  init: bool = True, lb_weights=None) -> list[Simulation]:
    """A batch of independent sphere scatterers of uneven sizes: cube edge
    (interior cells) ~ U{edge_range}, sphere radius ~ U[0.2, 0.4] x edge,
    eps_r ~ U[1.5, 8], all from numpy's default_rng(seed), each with a
    pml-cell CPML, a Gaussian TFSF plane wave and an NF2FF box.  The TF box
    inset is min(7, floor(edge/2 - r) - 1) cells inside the CPML so the box
    always encloses the sphere (a 0.4 x edge sphere in a 48-cell cube leaves
    4 cells); the NF2FF surface sits max(1, inset - 4) cells inside the CPML.
    lb_weights: a `pml`-cell learned absorbing layer with these weights on
    every grid instead of the CPML (the same spheres, sources and boxes).
    """
    rng = np.random.default_rng(seed)
    sims = []
    for _ in range(count):
        n = int(rng.integers(edge_range[0], edge_range[1] + 1))
        frac = float(rng.uniform(0.2, 0.4))
        eps_r = float(rng.uniform(1.5, 8.0))
        r_cells = frac * n
        t_in = max(2, min(7, int(np.floor(n / 2 - r_cells)) - 1))
        n_in = max(1, t_in - 4)
        sims.append(dielectric_sphere(n=n, pml=pml, steps=steps, dl=dl, tfsf_inset=pml + t_in,
                                      nf2ff_inset=pml + n_in, freqs=freqs, radius=r_cells * dl,
                                      eps_r=eps_r, mu_r=1.0, layout=layout, init=init,
                                      lb_weights=lb_weights))
    return sims


LB_NPARAM = 12 * 32 + 32 + 32 * 32 + 32 + 32 + 1  # XF_LB_NPARAM


def point_source_box(n: int = 40, boundary: str = "cpml", size: int = 16, steps: int = 400,
                     seed: int = 0, n_sources: int = 3, dl: float = 5e-3, weights=None,
                     record=None, init: bool = True) -> Simulation:
    """Vacuum cube of n^3 cells with seeded point sources, bounded by
    boundary = "cpml" (size-cell CPML), "learned" (size-cell learned layer
    with `weights`), "mur" (the learned layer with a zero network: first-order
    Mur) or "pec" (nothing).  Sources (seeded by `seed`, the same physical
    positions whatever the boundary): random E component, position in the
    central half of the cube, cosine-modulated Gaussian pulses with carrier
    f ~ U[2, 5] GHz (at least 12 cells per wavelength at dl = 5 mm), envelope
    tau = 1 / f (zero mean to 5e-5), delay 3 tau + U[0, 0.3] ns, amplitude
    +-U[0.5, 1].  record = (inset, every, slots)
    adds a training-sample recorder on the plane `inset` nodes inside each
    face."""
    rng = np.random.default_rng(seed)
    s = Simulation(dl, dl, dl, 0.9)
    s.add_cube("domain", (0.0, 0.0, 0.0), (n * dl,) * 3, Material.air())
    margin = 0
    if boundary == "cpml":
        s.add_pml_all(size)
        margin = size
    elif boundary in ("learned", "mur"):
        w = np.zeros(LB_NPARAM, np.float16) if boundary == "mur" else weights
        s.add_learned_boundary(w, thickness=size)
        margin = size
    elif boundary != "pec":
        raise ValueError(f"unknown boundary {boundary!r}")
    for _ in range(n_sources):
        comp = int(rng.integers(0, 3))
        pos = [margin + n // 4 + int(rng.integers(0, n // 2)) for _ in range(3)]
        f = float(rng.uniform(2e9, 5e9))
        tau = 1.0 / f
        t0 = 3 * tau + float(rng.uniform(0, 0.3e-9))
        amp = float(rng.choice([-1.0, 1.0]) * rng.uniform(0.5, 1.0))
        s.add_point_source(comp, pos, Waveform.cosine_modulated_gaussian(tau, t0, f), amp)
    if record is not None:
        inset, every, slots = record
        s.add_boundary_recorder(inset, every, slots)
    if init:
        s.init(steps)
    return s


def dispersive_sphere(n: int = 256, pml: int = 10, steps: int = 6000, dl: float = 2.5e-3,
                      freqs=(0.5e9, 1e9, 2e9), init: bool = True) -> Simulation:
    """Drude core (omega_p = 2 pi 2 GHz, gamma = 2 pi 0.2 GHz) inside a
    one-pole Lorentz shell (delta_eps = 2, omega_0 = 2 pi 1.5 GHz,
    delta = 2 pi 0.1 GHz).  Radii: shell 0.3 * edge, core 0.2 * edge."""
    edge = n * dl
    half = edge / 2
    s = Simulation(dl, dl, dl, 0.9)
    s.add_cube("domain", (-half, -half, -half), (edge,) * 3, Material.air())
    tp = 2 * np.pi
    s.add_sphere("shell", (0, 0, 0), 0.3 * edge,
                 Material.lorentz(1.0, [2.0], [tp * 1.5e9], [tp * 0.1e9]))
    s.add_sphere("core", (0, 0, 0), 0.2 * edge, Material.drude(1.0, [tp * 2e9], [tp * 0.2e9]))
    tau = 20 * dl / 6e8
    s.add_tfsf3d((pml + 7,) * 3, 0.0, 0.0, 0.0, Waveform.gaussian(tau, 4.5 * tau))
    s.add_pml_all(pml)
    s.add_nf2ff_fd((pml + 3,) * 3, list(freqs))
    if init:
        s.init(steps)
    return s


__all__ = ["FACES", "LB_NPARAM", "batch_spheres", "cylinder_2d", "dielectric_sphere",
           "dispersive_sphere", "large_sphere", "point_source_box", "vacuum_benchmark"]
