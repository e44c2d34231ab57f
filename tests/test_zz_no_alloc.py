"""arbo_sur_forward_noalloc (the surrogate's forward pass on device buffers
with its activations in the library's fixed scratch array) against
arbo_sur_forward: bit-identical outputs on two calls in a row, the second
smaller and on a stream of the caller's, and the size limit enforced.

Device buffers come from the HIP runtime the engine library itself links
(libamdhip64), called through ctypes."""
import ctypes as C

import numpy as np
import pytest

from arbo_gpu_fit_amd import abi
from arbo_gpu_fit_amd.surrogate import Surrogate

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]), ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipDeviceSynchronize", []), ("hipStreamCreate", [C.POINTER(C.c_void_p)]),
                           ("hipStreamDestroy", [C.c_void_p]), ("hipStreamSynchronize", [C.c_void_p]),
                           ("hipSetDevice", [C.c_int])):
            fn = getattr(self.lib, name)
            fn.argtypes = args
            fn.restype = C.c_int

    def check(self, rc, what):
        assert rc == 0, f"{what}: HIP error {rc}"

    def upload(self, a: np.ndarray) -> C.c_void_p:
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), a.nbytes), "hipMalloc")
        self.check(self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return p

    def alloc(self, nbytes: int) -> C.c_void_p:
        p = C.c_void_p()
        self.check(self.lib.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p

    def download(self, p: C.c_void_p, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype)
        self.check(self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H), "hipMemcpy D2H")
        return out


def test_forward_noalloc_matches_forward(hip_lib):
    lib = hip_lib   # arbo_sur_forward_noalloc bound by surrogate.bind (device pointers as void*)
    widths = [6, 256, 256, 256, 256, 260]
    sur = Surrogate(widths, max_batch=4096, seed=3, lib=lib)
    hip = Hip()
    hip.check(hip.lib.hipSetDevice(0), "hipSetDevice")
    rng = np.random.default_rng(8)
    stream = C.c_void_p()
    hip.check(hip.lib.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    print("[noalloc] 6-256^4-260, 4,096 rows then 77 rows", flush=True)
    for n, on_stream in ((4096, False), (77, True)):
        x = rng.normal(size=(n, widths[0])).astype(np.float32)
        ref = sur.forward(x)
        xd = hip.upload(x)
        od = hip.upload(np.full((n, widths[-1]), np.nan, np.float32))
        rc = lib.arbo_sur_forward_noalloc(sur._h, n, xd, od, stream if on_stream else None)
        assert rc == 0, lib.arbo_sur_last_error(sur._h)
        hip.check(hip.lib.hipStreamSynchronize(stream) if on_stream else hip.lib.hipDeviceSynchronize(), "sync")
        got = hip.download(od, (n, widths[-1]))
        assert np.array_equal(got, ref), f"{n} rows: max difference {np.nanmax(np.abs(got - ref))}"
        hip.lib.hipFree(xd)
        hip.lib.hipFree(od)
    # beyond the scratch: 4,097 rows of 256-wide hidden layers
    big = hip.alloc(4097 * widths[0] * 4)
    out = hip.alloc(4097 * widths[-1] * 4)
    assert lib.arbo_sur_forward_noalloc(sur._h, 4097, big, out, None) == abi.E_CAPACITY
    hip.lib.hipFree(big)
    hip.lib.hipFree(out)
    hip.lib.hipStreamDestroy(stream)
    sur.close()
