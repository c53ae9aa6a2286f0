"""The allocation-free entry point mcml_run_batch_sums_noalloc (GPU): per-(set, SDS) detected
weight and counts written to caller-owned device arrays, computed in the library's fixed
per-GPU scratch array, against mcml_run_batch's reflectance and detected arrays for the
same sets and seed.  Two calls in a row, the second with a smaller input, so the scratch's
leftovers from the first call must not reach the second.

Bar: counts equal; sums equal bit for bit where every weight is an integer number of
W-units and the FP64 sums are exact (fibre detection), and within 1e-12 relative where the
ring arc fraction makes weights fractional (the two sums round differently)."""
import ctypes as C

import numpy as np
import pytest

import fixtures as F
from mcml_gpu_amd import abi, make_opts

SEED = 97


class DeviceArray:
    """Caller-owned device memory from the HIP runtime libmcml.so is linked against (a second
    runtime in the process, such as the one bundled with PyTorch, cannot share the GPU)."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), self.host.nbytes) == 0
        assert self.hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # host to device

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def __del__(self):
        if self.ptr:
            self.hip().hipFree(self.ptr)


def noalloc(ctx, sims, opts):
    g = len(sims) * sims[0].num_sds
    refl = DeviceArray(np.full(g, -1.0))          # sentinels: every element must be written
    det = DeviceArray(np.full(g, 2 ** 63, np.uint64))
    arr = (abi.Sim * len(sims))(*sims)
    rc = ctx.lib.mcml_run_batch_sums_noalloc(ctx.h, arr, len(sims), C.byref(opts), SEED, refl.ptr, det.ptr)
    assert rc == 0, ctx.lib.mcml_last_error(ctx.h)
    return refl.get().reshape(len(sims), -1), det.get().reshape(len(sims), -1)


@pytest.mark.gpu
def test_noalloc_sums_equal_run_batch_on_two_calls(ctx):
    big = [F.fiber_busy(300000), F.fiber_busy(90000), F.fiber_busy(1), F.fiber_busy(150000), F.fiber_busy(0)]
    big[1].layers[0].mua = 0.9
    big[3].layers[1].mus = 60.0
    small = [F.ring_busy(40000), F.ring_busy(7000)]
    small[1].layers[0].mua = 0.8
    for sims, exact in ((big, True), (small, False)):
        print("allocation-free call: %d sets, %d photons" % (len(sims), sum(s.number_photons for s in sims)),
              flush=True)
        ref = ctx.run_batch(sims, make_opts(skip_records=True), seed=SEED)
        for opts in (make_opts(), make_opts(block_size=128, chunk=64, refill_min=1, boundary_min=1)):
            refl, det = noalloc(ctx, sims, opts)
            assert np.array_equal(det, ref.detected)
            assert det.sum() > 100
            if exact:
                assert np.array_equal(refl, ref.reflectance)
            else:
                assert np.allclose(refl, ref.reflectance, rtol=1e-12, atol=0)
                assert not np.array_equal(refl, np.zeros_like(refl))


@pytest.mark.gpu
def test_noalloc_rejects_what_it_cannot_hold(ctx):
    sims = [F.ring_busy(10)] * 1025
    buf = DeviceArray(np.zeros(4))
    arr = (abi.Sim * len(sims))(*sims)
    rc = ctx.lib.mcml_run_batch_sums_noalloc(ctx.h, arr, len(sims), C.byref(make_opts()), SEED,
                                             buf.ptr, buf.ptr)
    assert rc == abi.ERR_LIMIT and b"at most 1024" in ctx.lib.mcml_last_error(ctx.h)
    rc = ctx.lib.mcml_run_batch_sums_noalloc(ctx.h, arr, 2, C.byref(make_opts(replay=True)), SEED,
                                             buf.ptr, buf.ptr)
    assert rc == abi.ERR_ARG
    wide = [F.max_stack(10)] * 137  # 30 detectors: 4110 (set, SDS) groups
    arr = (abi.Sim * len(wide))(*wide)
    rc = ctx.lib.mcml_run_batch_sums_noalloc(ctx.h, arr, len(wide), C.byref(make_opts()), SEED,
                                             buf.ptr, buf.ptr)
    assert rc == abi.ERR_LIMIT and b"groups" in ctx.lib.mcml_last_error(ctx.h)
