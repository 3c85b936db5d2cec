"""sqg_hsub_apply_batch_noalloc (the H v entry point that allocates nothing,
its per-call buffers in the library's static scratch array) against
sqg_hsub_apply_batch on the same nodes and vectors, through the C-ABI: FP64
bit for bit, mixed precision bit for bit (the same kernels on the same FP32
copies), FP32-Lanczos contexts bit for bit (H v in FP64 on the context's
FP64 operator data), on two calls in a row, the second with fewer nodes; and the
SQG_E_OVERFLOW answer when the scratch array is too small.  Device inputs
and outputs are allocated through the same HIP runtime the library uses
(/opt/rocm/lib/libamdhip64.so), so that one runtime owns every device
buffer of the process."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import helpers as H
from sqdft_amd import _lib as L
from sqdft_amd.configs import smooth_veff
from sqdft_amd.sq import SpectralQuadrature

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    return _hip


class DevArray:
    """A device copy of a float64 numpy array (hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.shape, self.nbytes = a.shape, a.nbytes
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(a.nbytes, 8))) == 0
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self):
        out = np.empty(self.shape)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def noalloc(sq, veff_d, nodes, v_d, hv_d):
    nd = np.ascontiguousarray(np.asarray(nodes, dtype=np.int32))
    return sq._lib.sqg_hsub_apply_batch_noalloc(sq._ctx, veff_d.p, L.iptr(nd), nd.size, v_d.p, hv_d.p)


@pytest.mark.parametrize("prec", ["fp64", "mixed", "mixed_lanczos"])
@pytest.mark.parametrize("case", ["small_lih", "test_al"])
def test_noalloc_matches_hsub(prec, case):
    S, _, _ = getattr(H, case)()
    veff = smooth_veff(S)
    sq = SpectralQuadrature(S, precision=prec)
    M3 = sq.M ** 3
    rng = np.random.Generator(np.random.PCG64(61))
    veff_d = DevArray(veff)
    for nb in (min(S.nn, 37), 5):
        say(f"noalloc {case} {prec} nb={nb}")
        nodes = np.sort(rng.choice(S.nn, size=nb, replace=False)).astype(np.int32)
        v = rng.normal(size=(nb, M3))
        ref = sq.HsubTimesVec(veff, nodes, v).reshape(nb, M3)
        v_d, hv_d = DevArray(v), DevArray(np.full((nb, M3), np.nan))
        code = noalloc(sq, veff_d, nodes, v_d, hv_d)
        assert code == 0, L.STATUS.get(code, code)
        got = hv_d.get()
        v_d.free()
        hv_d.free()
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
    veff_d.free()
    sq.close()


def test_noalloc_rejects_sizes_beyond_the_scratch():
    S, _, _ = H.test_al()
    sq = SpectralQuadrature(S, precision="fp64")
    M3 = sq.M ** 3
    per = M3 * 8 + 512 * 4
    nb = (64 << 20) // per + 8   # more nodes than SQG_NOALLOC_SCRATCH_BYTES holds
    nodes = np.arange(nb, dtype=np.int32) % S.nn
    veff_d = DevArray(np.zeros(S.nn))
    v_d, hv_d = DevArray(np.zeros((nb, M3))), DevArray(np.zeros((nb, M3)))
    assert noalloc(sq, veff_d, nodes, v_d, hv_d) == -11   # SQG_E_OVERFLOW
    for d in (veff_d, v_d, hv_d):
        d.free()
    sq.close()
