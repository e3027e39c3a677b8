"""The learned absorbing layer: training data from thick-CPML runs of the
engine itself, training on the GPU, and the reflection judge.

Pipeline (no reference counterpart):
  1. record_samples: vacuum boxes with seeded point sources bounded by a
     16-cell CPML run on the GPU with a recorder on the plane one node inward
     of the CPML's inner edge; each sample is the 12 column inputs and the
     true next value of the plane node (xf_learned_boundary, XF_LB_RECORD).
  2. Trainer: the HIP trainer (xf_lb_trainer_*) fits the network's residual
     over the first-order Mur update.
  3. reflection_errors: the same point-source scene (a seed not used for
     training) bounded by the learned 2-cell layer, by Mur, by thin CPMLs,
     compared over the interior with the 16-cell CPML run, the judge.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native, scenes
from ._native import XfTiming  # noqa: F401  (re-exported for callers)
from .simulation import C_0, Engine

N_PARAM = scenes.LB_NPARAM


def mur_k(dt: float, dl: float) -> float:
    """First-order Mur coefficient (c dt - dl) / (c dt + dl)."""
    return (C_0 * dt - dl) / (C_0 * dt + dl)


def init_weights(seed: int = 0, out_scale: float = 0.0) -> np.ndarray:
    """He-initialised hidden layers; output layer scaled by out_scale
    (0: the network starts as the plain Mur update)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / 12), (32, 12))
    W2 = rng.normal(0, np.sqrt(2 / 32), (32, 32))
    w3 = rng.normal(0, np.sqrt(1 / 32), 32) * out_scale
    return np.concatenate([W1.ravel(), np.zeros(32), W2.ravel(), np.zeros(32), w3,
                           [0.0]]).astype(np.float32)


def _lib():
    lib = _native.hip()
    if not getattr(lib, "_xf_lb_declared", False):
        pf = ctypes.POINTER(ctypes.c_float)
        sigs = {
            "xf_lb_trainer_create": [pf, ctypes.c_int64, ctypes.c_double, ctypes.c_int32,
                                     ctypes.POINTER(ctypes.c_void_p)],
            "xf_lb_trainer_set_weights": [ctypes.c_void_p, pf],
            "xf_lb_trainer_get_weights": [ctypes.c_void_p, pf, ctypes.POINTER(ctypes.c_uint16)],
            "xf_lb_trainer_grad": [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
                                   ctypes.POINTER(ctypes.c_double), pf],
            "xf_lb_trainer_train": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_float, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_double)],
            "xf_lb_trainer_samples": [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)],
            "xf_lb_trainer_set_weighting": [ctypes.c_void_p, ctypes.c_int32],
            "xf_lb_trainer_destroy": [ctypes.c_void_p],
        }
        for name, args in sigs.items():
            getattr(lib, name).argtypes = args
            getattr(lib, name).restype = ctypes.c_int
        lib._xf_lb_declared = True
    return lib


class Trainer:
    """GPU trainer of the column network (xf_lb_trainer_*)."""

    def __init__(self, samples: np.ndarray, mur_k: float, device: int = 0):
        self._lib = _lib()
        s = np.ascontiguousarray(samples, dtype=np.float32).reshape(-1, 13)
        h = ctypes.c_void_p()
        self._h = None
        _native.check_hip(self._lib.xf_lb_trainer_create(
            s.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), s.shape[0], mur_k, device,
            ctypes.byref(h)), "xf_lb_trainer_create")
        self._h = h
        self.n = s.shape[0]

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == N_PARAM
        _native.check_hip(self._lib.xf_lb_trainer_set_weights(
            self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "set_weights")
        return self

    def weights(self):
        """(fp32 master weights, fp16 bits as uint16)."""
        w = np.empty(N_PARAM, np.float32)
        b = np.empty(N_PARAM, np.uint16)
        _native.check_hip(self._lib.xf_lb_trainer_get_weights(
            self._h, w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            b.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))), "get_weights")
        return w, b

    def set_weighting(self, power: int):
        """Loss weight (s / s_max)^power per sample (0: uniform)."""
        _native.check_hip(self._lib.xf_lb_trainer_set_weighting(self._h, power), "set_weighting")
        return self

    def grad(self, first: int = 0, count: int | None = None):
        count = self.n - first if count is None else count
        loss = ctypes.c_double()
        g = np.empty(N_PARAM, np.float32)
        _native.check_hip(self._lib.xf_lb_trainer_grad(
            self._h, first, count, ctypes.byref(loss),
            g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "grad")
        return loss.value, g

    def loss(self) -> float:
        return self.grad()[0]

    def train(self, steps: int, batch: int = 4096, lr: float = 1e-3, seed: int = 0) -> float:
        last = ctypes.c_double()
        _native.check_hip(self._lib.xf_lb_trainer_train(self._h, steps, batch, lr, seed,
                                                         ctypes.byref(last)), "train")
        return last.value

    def close(self):
        if getattr(self, "_h", None):
            self._lib.xf_lb_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def record_samples(n: int = 40, seeds=(0, 1, 2), steps: int = 400, pml: int = 16,
                   depth: int = 1, every: int = 2, precision: str = "fp32",
                   n_sources: int = 3):
    """Training samples from point-source boxes bounded by a `pml`-cell CPML,
    recorded every `every` steps on the plane `depth` nodes inward of the
    CPML's inner edge (depth 1 trains a layer that reflects less in the
    closed loop than depth 2 or 3: measured over three network inits and
    three unseen seeds, 1.5e-2 against 2.4e-2).  Returns (samples (N, 13)
    float32, mur_k)."""
    out = []
    k = None
    for seed in seeds:
        slots = steps // every + 1
        s = scenes.point_source_box(n=n, boundary="cpml", size=pml, steps=steps, seed=seed,
                                    n_sources=n_sources, record=(pml + depth, every, slots))
        e = Engine(s, precision=precision).upload()
        e.run(steps).download()
        e.close()
        smp = s.lb_samples()
        out.append(smp.reshape(-1, 13))
        k = mur_k(s.dt, 5e-3)
    return np.concatenate(out), k


def augment(samples: np.ndarray, noise: float, seed: int = 0) -> np.ndarray:
    """The samples plus a copy whose 12 inputs carry Gaussian noise of
    `noise` times the row's largest input magnitude (targets unchanged):
    the closed loop feeds the layer its own slightly-off values, which the
    one-step samples never show it."""
    rng = np.random.default_rng(seed)
    x = np.array(samples, np.float32, copy=True)
    s = np.max(np.abs(x[:, :12]), axis=1, keepdims=True)
    x[:, :12] = (x[:, :12] + noise * s * rng.standard_normal((x.shape[0], 12))).astype(np.float32)
    return np.concatenate([np.asarray(samples, np.float32), x])


def train_layer(samples: np.ndarray, k: float, seed: int = 0, device: int = 0,
                noise: float = 0.01):
    """The training recipe: the samples plus a copy with 1 % input noise
    (augment), Adam from the plain Mur start, 1500 steps at lr 2e-3 on
    minibatches of 8192, 750 at 5e-4, 750 at 1.25e-4 on 16384 (longer runs
    fit the one-step target better but reflect more in the closed loop).
    With depth-1 samples the noise copy lowers the reflection further and
    steadies it across network inits (mean 1.44e-2, worst 1.59e-2 over three
    inits, against 1.54e-2 and 1.66e-2 without it).  Returns (fp16 weight
    bits, loss before, loss after); the losses are on the training set."""
    if noise > 0:
        samples = augment(samples, noise, seed + 100)
    t = Trainer(samples, k, device).set_weights(init_weights(seed))
    l0 = t.loss()
    t.train(1500, batch=8192, lr=2e-3, seed=seed + 1)
    t.train(750, batch=8192, lr=5e-4, seed=seed + 2)
    t.train(750, batch=16384, lr=1.25e-4, seed=seed + 3)
    l1 = t.loss()
    _, w16 = t.weights()
    t.close()
    return w16, l0, l1


def _interior(sim, margin: int, n: int):
    return [sim.array(c)[margin:margin + n, margin:margin + n, margin:margin + n].copy()
            for c in ("ex", "ey", "ez")]


def run_snapshots(boundary: str, n: int, seed: int, steps: int, every: int, size: int,
                  weights=None, precision: str = "fp64", n_sources: int = 3):
    """Interior E snapshots (every `every` steps) of the point-source box."""
    s = scenes.point_source_box(n=n, boundary=boundary, size=size, steps=steps, seed=seed,
                                n_sources=n_sources, weights=weights)
    margin = size if boundary != "pec" else 0
    e = Engine(s, precision=precision).upload()
    snaps = []
    for _ in range(steps // every):
        e.run(every).download()
        snaps.append(_interior(s, margin, n))
    t = e.timing()
    e.close()
    cells = int(np.prod(s.info()[0]))
    return snaps, t["run_ms"], cells


def reflection_errors(weights, n: int = 40, seed: int = 7, steps: int = 400, every: int = 20,
                      precision: str = "fp64", variants=None, ref=None):
    """Relative L2 error of the interior E field over time of each boundary
    against the 16-cell CPML run (the judge).  `ref` may pass that run's
    snapshots (run_snapshots("cpml", ..., 16)) when several boundaries are
    judged on one seed."""
    if ref is None:
        ref, _, _ = run_snapshots("cpml", n, seed, steps, every, 16, precision=precision)
    den = np.sqrt(sum(float((c ** 2).sum()) for snap in ref for c in snap))
    variants = variants or [("learned", 2), ("mur", 2), ("cpml", 2), ("cpml", 4), ("cpml", 8)]
    out = {}
    for kind, size in variants:
        snaps, _, _ = run_snapshots(kind, n, seed, steps, every, size,
                                    weights=weights if kind == "learned" else None,
                                    precision=precision)
        num = np.sqrt(sum(float(((a - b) ** 2).sum()) for sa, sb in zip(snaps, ref)
                          for a, b in zip(sa, sb)))
        out[f"{kind}{size}"] = num / den
    return out


def train_selected(samples: np.ndarray, k: float, inits=(0, 1, 2), noises=(0.01, 0.02),
                   val_seeds=(11, 12, 13), n: int = 40, steps: int = 400, every: int = 20,
                   device: int = 0):
    """One layer per (input-noise level, network init) by train_layer, and
    the one whose closed-loop reflection against the 16-cell CPML, averaged
    over the validation seeds, is lowest.  The validation seeds are scenes
    neither the training data (seeds 0-5) nor the judge (seeds 7-10) use:
    one-step training loss does not rank layers by their closed-loop
    reflection (over six inits and three noise levels the validation means
    ranged from 1.3e-2 to 2.5e-2 at nearly equal loss), so the choice is
    made on held-out closed-loop runs.  Returns (fp16 weight bits, table of
    the candidates: noise, init, loss, validation errors and mean)."""
    refs = {sd: run_snapshots("cpml", n, sd, steps, every, 16)[0] for sd in val_seeds}
    table = []
    best = None
    for noise in noises:
        for init in inits:
            w16, _, l1 = train_layer(samples, k, seed=init, device=device, noise=noise)
            val = [reflection_errors(w16, n=n, seed=sd, steps=steps, every=every,
                                     variants=[("learned", 2)], ref=refs[sd])["learned2"]
                   for sd in val_seeds]
            row = {"noise": noise, "init": init, "loss": float(l1),
                   "val": [float(v) for v in val], "val_mean": float(np.mean(val))}
            table.append(row)
            if best is None or row["val_mean"] < best[1]["val_mean"]:
                best = (w16, row)
    return best[0], table
