"""The FFT pair entry point that allocates nothing
(psx_fft_scaled_pair_noalloc): its scratch is a static device array per GPU.
Two calls in a row, the second on a smaller grid, each agree with the plan's
psx_fft_scaled_pair on the same inputs and with the oracle; grids beyond the
scratch's limits are refused before any launch."""
import ctypes

import numpy as np
import pytest

import oracle as O
from polyswift_plusplus_amd import PsxError, _abi
from polyswift_plusplus_amd._abi import PSX_FP32, PSX_FP64, check

pytestmark = pytest.mark.gpu


class Dev:
    """A device buffer from psx_buffer_create."""

    def __init__(self, L, host):
        self.L, self.nbytes = L, host.nbytes
        self.p = ctypes.c_void_p()
        check(L.psx_buffer_create(self.nbytes, ctypes.byref(self.p)))
        check(L.psx_buffer_upload(self.p, host.ctypes.data, self.nbytes))

    def get(self, dtype, shape):
        out = np.empty(shape, dtype)
        check(self.L.psx_buffer_download(out.ctypes.data, self.p, self.nbytes))
        return out

    def free(self):
        check(self.L.psx_buffer_destroy(self.p))


def run_pair(L, shape, prec, data, kdata):
    """(no-alloc result, plan result) for the same device inputs."""
    dt = np.float64 if prec == PSX_FP64 else np.float32
    d, k = Dev(L, np.ascontiguousarray(data, dt)), Dev(L, np.ascontiguousarray(kdata, dt))
    r1, r2 = Dev(L, np.zeros(shape, dt)), Dev(L, np.zeros(shape, dt))
    plan = ctypes.c_void_p()
    try:
        check(L.psx_fft_scaled_pair_noalloc(*shape, prec, None, d.p, k.p, r1.p))
        check(L.psx_fft_plan_create(*shape, prec, None, ctypes.byref(plan)))
        check(L.psx_fft_scaled_pair(plan, d.p, k.p, r2.p))
        check(L.psx_fft_synchronize(plan))
        return r1.get(dt, shape), r2.get(dt, shape)
    finally:
        if plan:
            check(L.psx_fft_plan_destroy(plan))
        for b in (d, k, r1, r2):
            b.free()


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("prec", [PSX_FP64, PSX_FP32], ids=["fp64", "fp32"])
def test_noalloc_pair_two_calls(prec):
    L = _abi.load()
    rng = np.random.default_rng(71)
    for shape in ((48, 40, 32), (8, 12, 10)):      # the second call is smaller
        print(f"no-alloc pair {shape} prec {prec}", flush=True)
        data, kdata = rng.standard_normal(shape), rng.random(shape)
        got, plan = run_pair(L, shape, prec, data, kdata)
        want = O.scaled_fft_pair(data, kdata)
        if prec == PSX_FP64:
            assert rel(got, plan) < 1e-13 and rel(got, want) < 1e-12
        else:
            assert rel(got, plan) < 2e-6 and rel(got, want) < 2e-5


def test_noalloc_pair_largest_grid():
    L = _abi.load()
    shape = (128, 128, 128)
    print(f"no-alloc pair {shape} fp64", flush=True)
    rng = np.random.default_rng(72)
    data, kdata = rng.standard_normal(shape), rng.random(shape)
    got, plan = run_pair(L, shape, PSX_FP64, data, kdata)
    assert rel(got, plan) < 1e-13


def test_noalloc_pair_refuses_oversized_grids():
    L = _abi.load()
    z = np.zeros(8)
    bufs = [Dev(L, z) for _ in range(3)]
    try:
        for shape, prec in (((256, 128, 128), PSX_FP64), ((2048, 2, 2), PSX_FP32),
                            ((512, 128, 128), PSX_FP32)):
            st = L.psx_fft_scaled_pair_noalloc(*shape, prec, None, *[b.p for b in bufs])
            assert st == 2, shape
            assert "no-alloc FFT pair" in _abi.last_error()
        assert L.psx_fft_scaled_pair_noalloc(4, 4, 4, 7, None, *[b.p for b in bufs]) == 1
        assert L.psx_fft_scaled_pair_noalloc(4, 4, 4, PSX_FP64, None, None, bufs[1].p, bufs[2].p) == 1
    finally:
        for b in bufs:
            b.free()
