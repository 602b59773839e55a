"""The allocation-free crater product (btpm_crater_contract_noalloc, device
memory in and out, its packed view-factor tiles in a fixed device array)
against the allocating entry point (btpm_crater_contract) on the same
inputs: bit for bit, on two calls in a row with the second smaller and with
another view-factor matrix (so the tiles are packed again), once on the
library's stream (blocking) and once enqueued on a caller's stream; sizes
outside its limit are argument errors."""
import ctypes as C

import numpy as np
import pytest

from binarytpm_amd import _lib, inputs
from binarytpm_amd import thermal as th

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers and a stream from the HIP runtime libbtpm itself uses
    (already loaded in this process), through ctypes."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipStreamCreate",
                     "hipStreamSynchronize", "hipStreamDestroy", "hipDeviceSynchronize"):
            getattr(self.L, name).restype = C.c_int

    def ok(self, rc):
        assert rc == 0, f"HIP error {rc}"

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), C.c_size_t(a.nbytes)))
        self.ok(self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1))
        return p.value

    def download(self, p, shape):
        out = np.empty(shape, np.float32)
        self.ok(self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p),
                                 C.c_size_t(out.nbytes), 2))
        return out

    def free(self, p):
        self.ok(self.L.hipFree(C.c_void_p(p)))


def _noalloc(ctx, vf_d, e_d, n, q_d, stream=None):
    L = _lib.lib()
    return L.btpm_crater_contract_noalloc(ctx.handle, C.c_void_p(vf_d), C.c_void_p(e_d), int(n),
                                          C.c_void_p(q_d), C.c_void_p(stream) if stream else None)


def test_noalloc_contract_matches_allocating(ctx):
    hip = _Hip()
    rng = np.random.default_rng(5)
    cg = inputs.crater_cap(45.0)
    cases = [(np.ascontiguousarray(cg.vf, np.float32), 777),
             (rng.uniform(0.0, 0.01, 10000).astype(np.float32), 333)]
    stream = C.c_void_p()
    hip.ok(hip.L.hipStreamCreate(C.byref(stream)))
    try:
        for use_stream in (False, True):
            for vf, n in cases:
                e = rng.uniform(0.3, 1500.0, (n, 100)).astype(np.float32)
                ref = th.crater_contract(ctx, vf, e)
                dv, de = hip.upload(vf), hip.upload(e)
                dq = hip.upload(np.full((n, 100), np.nan, np.float32))
                if use_stream:
                    _lib.check(_noalloc(ctx, dv, de, n, dq, stream.value))
                    hip.ok(hip.L.hipStreamSynchronize(stream))
                else:
                    _lib.check(_noalloc(ctx, dv, de, n, dq))
                got = hip.download(dq, (n, 100))
                for p in (dv, de, dq):
                    hip.free(p)
                print(f"[noalloc] n={n} caller stream={use_stream}: identical "
                      f"{np.array_equal(got, ref)}", flush=True)
                np.testing.assert_array_equal(got, ref)
        dv = hip.upload(np.zeros(10000, np.float32))
        for bad in (0, -1, (1 << 24) + 1):
            assert _noalloc(ctx, dv, dv, bad, dv) == _lib.BTPM_ERR_ARG
        hip.free(dv)
    finally:
        hip.ok(hip.L.hipStreamDestroy(stream))
