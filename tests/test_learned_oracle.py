"""Pins of the learned absorbing layer's CPU restatements (oracle/xf_oracle.c
lb_step / lb_predict and oracle/lb_mlp.py), without a GPU.

* the C predictor and the numpy predictor agree bit for bit;
* the predictor's matrix-product model against outputs of the matrix cores
  measured on an MI355X;
* a hand-built network with a known output (r = x^_4) gives
  mur + s * fp16(x4 / s) exactly;
* the numpy gradient matches central finite differences;
* the recorder's columns are the field values a direct gather from per-step
  snapshots gives (inputs, Laplacian, target, column order);
* with a zero network (first-order Mur) the layer drains a point-source box
  like a 16-cell CPML, where a PEC box keeps its energy;
* a point source adds its waveform to its node.
"""
import numpy as np
import pytest

import oracle
from oracle import lb_mlp
import xfdtd_cuda_amd as xf
from xfdtd_cuda_amd import scenes


def _rand_x(n, seed=0):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((n, 12)) * 10.0 ** rng.uniform(-6, 2, (n, 1))).astype(np.float32)
    x[0] = 0.0          # an all-zero column: pure Mur (= 0)
    x[1, :] = 0.75      # max |x| in [0.5, 1): exponent 0
    return x


def test_c_and_numpy_predictors_agree_bitwise():
    w = lb_mlp.init_weights(3, out_scale=0.7)
    w[1505 - 1] = 0.01
    w16 = w.astype(np.float16).view(np.uint16)
    x = _rand_x(500)
    k = -0.2763
    a = oracle.lb_predict(w16, k, x)
    b = lb_mlp.predict(w16, x, k)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert a[0] == 0.0


def test_product_model_against_measured_matrix_cores():
    """The predictor's product model (two exact 8-term groups, one fp32
    rounding each) against v_mfma_f32_32x32x16_f16 outputs measured on an
    MI355X (tests/golden/make_mfma_fixture.py; one case per operand regime).
    The model is exact for most outputs and always within 8 units of 2^-23
    of the largest term; a single exact sum of all 16 terms or plain
    sequential fp32 adds fit the hardware less well, which is what the
    model is chosen by."""
    import os
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "mfma_f16_32x32x16_gfx950.npz"))
    a, b, c, d = f["a"].astype(np.float32), f["b"].astype(np.float32), f["c"], f["d"]
    rows_a = np.repeat(a, 32, axis=1).reshape(-1, 16)                   # (case, m, n) -> a[m]
    rows_b = np.tile(np.swapaxes(b, 1, 2), (1, 32, 1)).reshape(-1, 16)  # -> b[:, n]
    got = oracle.lb_mfma16(c.reshape(-1), rows_a, rows_b)
    want = d.reshape(-1)
    terms = np.abs(rows_a.astype(np.float64) * rows_b)
    scale = np.maximum(terms.max(axis=1), np.abs(c.reshape(-1)))
    err = np.abs(got.astype(np.float64) - want) / (scale * 2.0 ** -23)
    exact = np.mean(got == want)
    assert exact > 0.75, exact
    assert err.max() <= 8.0, err.max()
    # the alternatives the model was chosen over fit worse
    seq = c.reshape(-1).copy()
    for k in range(16):
        seq = (seq + (rows_a[:, k] * rows_b[:, k]).astype(np.float32)).astype(np.float32)
    assert np.mean(seq == want) < exact


def test_identity_network_closed_form():
    """W1 rows 0/1 = +-e4, W2 = I on 0/1, w3 = (1, -1): MLP(x^) = x^_4, the
    scaled input after its fp16 rounding, so the prediction is
    mur + s * fp16(x4 / s) exactly (every sum has one non-zero term)."""
    W1 = np.zeros((32, 12)); W1[0, 4] = 1; W1[1, 4] = -1
    W2 = np.zeros((32, 32)); W2[0, 0] = 1; W2[1, 1] = 1
    w3 = np.zeros(32); w3[0] = 1; w3[1] = -1
    w = np.concatenate([W1.ravel(), np.zeros(32), W2.ravel(), np.zeros(32), w3, [0.0]])
    w16 = w.astype(np.float16).view(np.uint16)
    x = _rand_x(300, 5)
    k = 0.31
    pred = oracle.lb_predict(w16, k, x)
    mur = (x[:, 4] + np.float32(k) * (x[:, 0] - x[:, 3])).astype(np.float32)
    e = lb_mlp.scale_exponent(x)
    x4h = np.ldexp(np.ldexp(x[:, 4], -e).astype(np.float16).astype(np.float32), e)
    assert np.array_equal(pred, (mur + x4h).astype(np.float32))
    assert np.array_equal(lb_mlp.predict(w16, x, k), pred)
    assert np.count_nonzero(x4h != x[:, 4]) > 100  # the fp16 rounding is in effect


def test_numpy_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = lb_mlp.init_weights(2, out_scale=1.0)
    w[:] = w.astype(np.float16).astype(np.float32)  # exactly representable
    xs = rng.uniform(-1, 1, (64, 12))
    r = rng.standard_normal(64) * 0.1
    _, g = lb_mlp.loss_and_grad(w, xs, r)
    w64 = w.astype(np.float64)
    for p in list(rng.choice(1505, 40, replace=False)) + [1504, 0, 384 + 5]:
        h = 1e-6
        wp, wm = w64.copy(), w64.copy()
        wp[p] += h
        wm[p] -= h
        fd = (lb_mlp.loss_only(wp, xs, r) - lb_mlp.loss_only(wm, xs, r)) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 + 1e-5 * abs(g[p]), (p, fd, g[p])


def _gather_columns(snaps_e, t, n, inset):
    """Columns of step t from snapshots E^t (snaps_e[t] = (ex, ey, ez) after t
    steps): inputs per xf_learned_boundary and target E^{t+1} at layer 0."""
    rows = []
    nn = [n, n, n]
    for f in range(6):
        c, pos = f // 2, f % 2
        a, b = (c + 1) % 3, (c + 2) % 3
        for cc, comp in ((0, a), (1, b)):
            p1 = nn[a] - inset - (1 if cc == 0 else 0)
            q1 = nn[b] - inset - (1 if cc == 1 else 0)

            def val(arr, layer, p, q):
                idx = [0, 0, 0]
                idx[c] = nn[c] - inset - layer if pos else inset + layer
                idx[a], idx[b] = p, q
                return np.float32(arr[comp][tuple(idx)])

            zero = (np.zeros((n + 1,) * 3),) * 3
            cur = snaps_e[t]
            prev = snaps_e[t - 1] if t >= 1 else zero
            nxt = snaps_e[t + 1]
            for p in range(inset, p1 + 1):
                for q in range(inset, q1 + 1):
                    x = [val(nxt, l, p, q) for l in (1, 2, 3)]
                    x += [val(cur, l, p, q) for l in range(4)]
                    x += [val(prev, l, p, q) for l in range(3)]
                    for layer in (1, 0):
                        u = val(cur, layer, p, q)
                        up = val(cur, layer, p + 1, q) if p < p1 else u
                        um = val(cur, layer, p - 1, q) if p > inset else u
                        vp = val(cur, layer, p, q + 1) if q < q1 else u
                        vm = val(cur, layer, p, q - 1) if q > inset else u
                        x.append(np.float32(np.float32(np.float32(up + um) + np.float32(vp + vm))
                                            - np.float32(4.0) * u))
                    rows.append(x + [val(nxt, 0, p, q)])
    return np.array(rows, np.float32)


def test_recorder_columns_match_snapshot_gather():
    steps, inset, every = 12, 7, 3
    s = scenes.point_source_box(n=10, boundary="cpml", size=5, steps=steps, seed=4,
                                record=(inset, every, 10))
    (n, _, _), _, _ = s.info()
    ptr = s.scene_ptr()

    def snap():
        # pad every component to (n+1)^3 so one index form serves all
        out = []
        for name in ("ex", "ey", "ez"):
            a = s.array(name)
            b = np.zeros((n + 1,) * 3)
            b[:a.shape[0], :a.shape[1], :a.shape[2]] = a
            out.append(b)
        return tuple(out)

    snaps = [snap()]
    for _ in range(steps):
        oracle.run(ptr, 1)
        snaps.append(snap())
    rec = s.lb_samples()
    assert rec.shape[0] == 4  # t = 0, 3, 6, 9
    for slot in range(4):
        t = slot * every
        want = _gather_columns(snaps, t, n, inset)
        assert want.shape == rec[slot].shape
        assert np.array_equal(rec[slot], want), f"slot {slot}"
    assert np.abs(rec[3]).max() > 0


def _energy(s, m):
    e = 0.0
    for name in ("ex", "ey", "ez"):
        a = s.array(name)
        e += float((a[m:a.shape[0] - m, m:a.shape[1] - m, m:a.shape[2] - m] ** 2).sum())
    return e


def test_zero_network_layer_absorbs_like_cpml():
    out = {}
    for b, size in (("cpml", 16), ("mur", 2), ("pec", 0)):
        s = scenes.point_source_box(n=24, boundary=b, size=size, steps=400, seed=1)
        oracle.run(s.scene_ptr(), 400)
        out[b] = _energy(s, size)
    # all pulses have left a 24-cell box with an absorbing boundary; only the
    # sources' static remainder (identical in every variant) stays
    assert out["pec"] > 1e3 * out["cpml"]
    assert out["mur"] < 1.05 * out["cpml"]


def test_point_source_adds_its_waveform():
    s = scenes.point_source_box(n=12, boundary="pec", steps=5, seed=2, n_sources=1)
    v = s.array("point.0.value")
    before = {c: s.array(c).copy() for c in ("ex", "ey", "ez")}
    oracle.run(s.scene_ptr(), 1)
    changed = [(c, np.argwhere(s.array(c) != before[c])) for c in before]
    changed = [(c, w) for c, w in changed if len(w)]
    if v[0] == 0.0:
        assert not changed
    else:
        assert len(changed) == 1 and len(changed[0][1]) == 1
        c, w = changed[0]
        assert s.array(c)[tuple(w[0])] == v[0]


def test_training_augmentation_keeps_targets_and_scales_noise_per_row():
    """learned.augment: the original samples, then a copy whose inputs carry
    noise proportional to each row's largest input (rows of very different
    scale get proportionate noise), targets untouched, reproducible."""
    from xfdtd_cuda_amd import learned
    rng = np.random.default_rng(3)
    x = (rng.standard_normal((4000, 13)) * 10.0 ** rng.uniform(-6, 3, (4000, 1))).astype(np.float32)
    a = learned.augment(x, 0.01, seed=5)
    assert a.shape == (8000, 13) and a.dtype == np.float32
    assert np.array_equal(a[:4000], x)
    assert np.array_equal(a[4000:, 12], x[:, 12])
    rel = (a[4000:, :12] - x[:, :12]) / np.max(np.abs(x[:, :12]), axis=1, keepdims=True)
    assert 0.009 < rel.std() < 0.011
    assert abs(rel.mean()) < 1e-3
    assert np.array_equal(learned.augment(x, 0.01, seed=5), a)


def test_learned_layer_validation_errors():
    s = xf.Simulation(5e-3, 5e-3, 5e-3, 0.9)
    s.add_cube("domain", (0, 0, 0), (0.05,) * 3, xf.Material.air())
    with pytest.raises(xf.XFError, match="weights"):
        s.add_learned_boundary(np.zeros(10, np.float16))
    s = xf.Simulation(5e-3, 5e-3, 5e-3, 0.9)
    s.add_cube("domain", (0, 0, 0), (0.05,) * 3, xf.Material.air())
    s.add_pml_all(4)
    s.add_learned_boundary(np.zeros(scenes.LB_NPARAM, np.float16), faces=("XN",))
    with pytest.raises(xf.XFError, match="both a PML and a learned layer"):
        s.init(5)
