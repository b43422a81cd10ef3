"""The allocation-free entry point mgo_ensemble_percentiles_noalloc against mgo_ensemble_percentiles_values (the
allocating entry point it is a variant of): bitwise equal on two calls in a row, the second with fewer runs and
positions (the scratch keeps the first call's contents), on the context stream and on a caller's stream; limits
of the scratch are refused before any launch. The test owns the device memory of the inputs and outputs (HIP
runtime calls through ctypes)."""
import ctypes as C

import numpy as np
import pytest

from memilio_amd import capi
from memilio_amd.engine import Context

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree / streams of the HIP runtime the library already loaded."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipFree", [C.c_void_p]), ("hipStreamCreate", [C.POINTER(C.c_void_p)]),
                           ("hipStreamSynchronize", [C.c_void_p]), ("hipStreamDestroy", [C.c_void_p])):
            getattr(self.lib, name).argtypes = args
            getattr(self.lib, name).restype = C.c_int

    def ok(self, e):
        assert e == 0, f"HIP error {e}"

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        self.ok(self.lib.hipMalloc(C.byref(p), a.nbytes))
        self.ok(self.lib.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D))
        return p

    def download(self, p, shape):
        a = np.empty(shape)
        self.ok(self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, a.nbytes, D2H))
        return a


def _noalloc(ctx, dv, runs, positions, ps, dout, stream=None):
    p = np.ascontiguousarray(ps, dtype=np.float64)
    return ctx.lib.mgo_ensemble_percentiles_noalloc(ctx.ptr, dv, runs, positions, capi.dptr(p), len(p), dout, stream)


def test_percentiles_noalloc_equals_allocating_entry():
    rng = np.random.default_rng(11)
    ps = [0.05, 0.25, 0.5, 0.75, 0.95]
    with Context() as ctx:
        hip = Hip()
        for runs, positions in ((1000, 3000), (37, 1234)):
            vals = rng.lognormal(0.0, 2.0, (runs, positions))
            vals[:, ::7] = 0.0   # ties
            ref = ctx.ensemble_percentiles_values(vals, ps).reshape(len(ps), positions)
            dv = hip.upload(vals)
            dout = hip.upload(np.full((len(ps), positions), np.nan))
            assert _noalloc(ctx, dv, runs, positions, ps, dout) == capi.OK
            assert np.array_equal(hip.download(dout, (len(ps), positions)), ref), (runs, positions)
            # enqueued on a caller's stream, completed by the caller
            s = C.c_void_p()
            hip.ok(hip.lib.hipStreamCreate(C.byref(s)))
            dout2 = hip.upload(np.zeros((len(ps), positions)))
            assert _noalloc(ctx, dv, runs, positions, ps, dout2, stream=s) == capi.OK
            hip.ok(hip.lib.hipStreamSynchronize(s))
            assert np.array_equal(hip.download(dout2, (len(ps), positions)), ref), (runs, positions)
            hip.ok(hip.lib.hipStreamDestroy(s))
            for p in (dv, dout, dout2):
                hip.ok(hip.lib.hipFree(p))
        dv = hip.upload(np.zeros((8193, 2)))
        dout = hip.upload(np.zeros((1, 2)))
        assert _noalloc(ctx, dv, 8193, 2, [0.5], dout) == capi.ERR_OUT_OF_RANGE
        assert "scratch" in ctx.lib.mgo_last_error(ctx.ptr).decode()
        for p in (dv, dout):
            hip.ok(hip.lib.hipFree(p))
