"""The allocation-free MLP gradient (rp3dg_mlp_gradient_noalloc: device inputs and outputs, partial
gradients in a fixed device array) against the allocating rp3dg_mlp_gradient on the same samples: the
gradient and the loss are identical, on two calls in a row, the second with fewer samples (so stale
partials of the first call would show), and sample counts outside [1, 4096] are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_SAMPLES = 4096   # RP3DG_MLP_NOALLOC_MAX_SAMPLES


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the step library already loaded."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(1, a.nbytes)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # hipMemcpyHostToDevice
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # hipMemcpyDeviceToHost
        return out

    def free(self, *ps):
        for p in ps:
            self.L.hipFree(p)


def test_noalloc_gradient_matches_allocating_gradient():
    from reactphysics3d_amd import lib
    L = lib.load()
    hip = _Hip()
    mlp = lib.Mlp(seed=3)
    rng = np.random.default_rng(17)
    mlp.set_normalization(rng.standard_normal(24).astype(np.float32), np.full(24, 0.7, np.float32),
                          np.array([0.5, 0.2, 0.2], np.float32))
    for n in (3000, 700):
        print(f"noalloc gradient: {n} samples", flush=True)
        F = rng.standard_normal((n, 24)).astype(np.float32)
        T = rng.standard_normal((n, 3)).astype(np.float32)
        g_ref, loss_ref = mlp.gradient(F, T)
        dF, dT = hip.upload(F), hip.upload(T)
        dG = hip.upload(np.full(20099, np.nan, np.float32))
        dL = hip.upload(np.full(1, np.nan, np.float32))
        assert hip.L.hipDeviceSynchronize() == 0
        rc = L.rp3dg_mlp_gradient_noalloc(mlp.h, dF, dT, n, dG, dL)
        assert rc == 0, lib.last_error()
        g = hip.download(dG, np.zeros(20099, np.float32))
        loss = hip.download(dL, np.zeros(1, np.float32))[0]
        hip.free(dF, dT, dG, dL)
        assert np.array_equal(g, np.asarray(g_ref, np.float32)), f"{n} samples: gradients differ"
        assert loss == np.float32(loss_ref), f"{n} samples: loss {loss} vs {loss_ref}"
    # sizes beyond the fixed scratch array, and empty calls, are refused before anything runs
    dummy = hip.upload(np.zeros(4, np.float32))
    for n in (MAX_SAMPLES + 1, 0, -5):
        rc = L.rp3dg_mlp_gradient_noalloc(mlp.h, dummy, dummy, n, dummy, dummy)
        assert rc == -1, n
        assert "4096" in lib.last_error()
    hip.free(dummy)
    mlp.close()
