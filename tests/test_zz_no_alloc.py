"""pbf_mlp_forward_backward_noalloc (device pointers, no allocation, a fixed
scratch array per GPU) against pbf_mlp_forward_backward (host arrays, the
library allocates), called through the C ABI: two calls in a row on one
stream, the second with fewer samples and another network, so that anything
the first call left in the scratch would show.  Device memory for the
inputs and outputs comes from the HIP runtime that libpbf itself uses."""
import ctypes as C

import numpy as np
import pytest

import mlp_oracle as mo
from pbf_host import load_library, mlp_blob_size, mlp_forward_backward, mlp_init

def _hip():
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


H2D, D2H = 1, 2


class Dev:
    """One device buffer of float32, filled from / read into numpy."""

    def __init__(self, hip, count):
        self.hip, self.count, self.ptr = hip, int(count), C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(self.count, 1) * 4) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32).reshape(-1)
        assert a.size == self.count
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0
        return self

    def get(self):
        out = np.zeros(self.count, np.float32)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _blob(seed, n_feat_scale=1.0):
    b = mlp_init(seed)
    rng = np.random.default_rng(seed)
    b[mo.OFF_FMU:mo.OFF_FSIG] = rng.normal(0, 0.1, 16)
    b[mo.OFF_FSIG:mo.OFF_TMU] = rng.uniform(0.5, 2.0, 16) * n_feat_scale
    b[mo.OFF_TMU], b[mo.OFF_TSIG] = -0.2, 0.7
    b[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.1, 64)
    return b


@pytest.mark.gpu
@pytest.mark.parametrize("first,second", [(5000, 1237)])
def test_noalloc_forward_backward_equals_allocating_entry_point(first, second):
    lib = load_library()
    hip = _hip()
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    for call, (n, seed) in enumerate([(first, 3), (second, 11)]):
        print(f"[noalloc] call {call}: n={n} seed={seed}", flush=True)
        rng = np.random.default_rng(seed)
        blob = _blob(seed)
        f = rng.normal(0, 1, (n, 16)).astype(np.float32)
        t = rng.normal(-0.1, 0.5, n).astype(np.float32)
        y_ref, g_ref, l_ref = mlp_forward_backward(blob, f, t)
        bufs = {"blob": Dev(hip, mlp_blob_size()).put(blob), "f": Dev(hip, 16 * n).put(f),
                "t": Dev(hip, n).put(t), "y": Dev(hip, n), "g": Dev(hip, mo.NPARAM), "l": Dev(hip, 1)}
        rc = lib.pbf_mlp_forward_backward_noalloc(bufs["blob"].ptr, bufs["f"].ptr, bufs["t"].ptr, n, bufs["y"].ptr,
                                                  bufs["g"].ptr, bufs["l"].ptr, stream)
        assert rc == 0, lib.pbf_last_error(None)
        assert hip.hipStreamSynchronize(stream) == 0
        y, g, l = bufs["y"].get(), bufs["g"].get(), float(bufs["l"].get()[0])
        for b in bufs.values():
            b.free()
        # the same kernels: outputs bit-identical; gradient and loss are sums
        # of float atomics, whose order varies from run to run
        assert np.array_equal(y, y_ref), call
        assert np.abs(g - g_ref).max() <= 1e-5 * np.abs(g_ref).max(), call
        assert l == pytest.approx(l_ref, rel=1e-5), call
        # and the CPU network (bf16 rounding where the GPU rounds) agrees
        y_cpu, g_cpu, l_cpu = mo.forward_backward(blob, f, t)
        assert np.abs(y - y_cpu).max() <= 2e-3 * max(1.0, np.abs(y_cpu).max()), call
    hip.hipStreamDestroy(stream)


def test_noalloc_rejects_bad_sizes():
    # argument checks come before any HIP call, so this runs without a GPU
    lib = load_library()
    one = C.c_void_p(16)  # never dereferenced: the call fails before any launch
    assert lib.pbf_mlp_forward_backward_noalloc(one, one, one, 0, None, None, None, None) == -1
    assert b"n must be" in lib.pbf_last_error(None)
    assert lib.pbf_mlp_forward_backward_noalloc(None, one, one, 10, None, None, None, None) == -1
    assert lib.pbf_mlp_forward_backward_noalloc(one, one, one, 2 ** 31 + 1, None, None, None, None) == -2
    assert b"PBF_MLP_NOALLOC_MAX_SAMPLES" in lib.pbf_last_error(None)
