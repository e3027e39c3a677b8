"""xf_lb_grad_noalloc, the gradient entry point that allocates nothing, against
the trainer's gradient (xf_lb_trainer_grad) on the same samples and weights:
bit for bit, on two calls in a row (the second with fewer samples, a smaller
range and another weighting), through the C ABI.  The caller's buffers are
allocated here with the HIP runtime the engine itself links."""
import ctypes

import numpy as np
import pytest

from oracle import lb_mlp
import xfdtd_cuda_amd as xf  # noqa: F401  (loads the engine and its HIP runtime)
from xfdtd_cuda_amd import _native, learned

pytestmark = pytest.mark.gpu

N_PARAM = learned.N_PARAM
XF_ERR_INVALID, XF_ERR_UNSUPPORTED = 1, 4  # include/xfdtd_hip.h


def _hip():
    rt = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    vp = ctypes.c_void_p
    rt.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
    rt.hipFree.argtypes = [vp]
    rt.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
    rt.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
    rt.hipStreamSynchronize.argtypes = [vp]
    rt.hipStreamDestroy.argtypes = [vp]
    rt.hipSetDevice.argtypes = [ctypes.c_int]
    for f in (rt.hipMalloc, rt.hipFree, rt.hipMemcpy, rt.hipStreamCreate, rt.hipStreamSynchronize,
              rt.hipStreamDestroy, rt.hipSetDevice):
        f.restype = ctypes.c_int
    return rt


H2D, D2H = 1, 2


class _Dev:
    """A device buffer holding a copy of a host array."""

    def __init__(self, rt, nbytes):
        self.rt, self.p = rt, ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.p), max(nbytes, 16)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.rt.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D) == 0

    def get(self, a):
        assert self.rt.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.rt.hipFree(self.p)
            self.p = ctypes.c_void_p()


def _lib():
    lib = _native.hip()
    f = lib.xf_lb_grad_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int32, ctypes.c_void_p,
                  ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    return lib


def _samples(n, seed):
    """Raw rows of 12 inputs + target over five decades of scale, with two
    all-zero rows (the predictor's unscaled case)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 13)).astype(np.float32)
    x *= (10.0 ** rng.uniform(-5, 0, (n, 1))).astype(np.float32)
    x[7] = 0.0
    x[n // 2] = 0.0
    return x


def _trainer_grad(smp, k, power, w, first, count):
    t = learned.Trainer(smp, k).set_weights(w).set_weighting(power)
    loss, g = t.grad(first, count)
    t.close()
    return loss, g


def test_noalloc_grad_matches_trainer_twice(gpu_available):
    lib, rt = _lib(), _hip()
    assert rt.hipSetDevice(0) == 0
    stream = ctypes.c_void_p()
    assert rt.hipStreamCreate(ctypes.byref(stream)) == 0
    k = 0.31
    cases = [  # (n, seed, power, first, count): the second call is smaller
        (50_000, 1, 2, 1_000, 45_000),
        (3_000, 2, 0, 17, 2_500),
    ]
    bufs = []
    try:
        for n, seed, power, first, count in cases:
            print(f"case: noalloc grad n={n} power={power} range=[{first}, {first + count})", flush=True)
            smp = _samples(n, seed)
            w = lb_mlp.init_weights(seed + 4, out_scale=0.5).astype(np.float32)
            raw, wd, ld, gd = (_Dev(rt, smp.nbytes), _Dev(rt, w.nbytes), _Dev(rt, 8), _Dev(rt, 4 * N_PARAM))
            bufs += [raw, wd, ld, gd]
            raw.put(smp)
            wd.put(w)
            st = lib.xf_lb_grad_noalloc(raw.p, n, k, power, wd.p, first, count, ld.p, gd.p, 0, stream)
            assert st == 0, _native.hip().xf_last_error()
            assert rt.hipStreamSynchronize(stream) == 0
            loss = ld.get(np.zeros(1, np.float64))[0]
            grad = gd.get(np.zeros(N_PARAM, np.float32))
            loss_t, grad_t = _trainer_grad(smp, k, power, w, first, count)
            assert np.isfinite(loss) and loss > 0
            assert loss == loss_t, (loss, loss_t)
            assert np.array_equal(grad, grad_t)
            assert np.abs(grad).max() > 0
    finally:
        for b in bufs:
            b.free()
        rt.hipStreamDestroy(stream)


def test_noalloc_grad_rejects_bad_arguments(gpu_available):
    lib = _lib()
    one = ctypes.c_void_p(16)  # never dereferenced: every call below is refused first
    assert lib.xf_lb_grad_noalloc(None, 10, 0.3, 0, one, 0, 10, one, one, 0, None) == XF_ERR_INVALID
    assert lib.xf_lb_grad_noalloc(one, 10, 0.3, 0, one, 5, 10, one, one, 0, None) == XF_ERR_INVALID
    assert lib.xf_lb_grad_noalloc(one, 10, 0.3, 7, one, 0, 10, one, one, 0, None) == XF_ERR_INVALID
    assert lib.xf_lb_grad_noalloc(one, 1 << 33, 0.3, 0, one, 0, 10, one, one, 0,
                                  None) == XF_ERR_UNSUPPORTED
    assert lib.xf_lb_grad_noalloc(one, 10, 0.3, 0, one, 0, 10, one, one, 999, None) == XF_ERR_INVALID
