"""fmmhip_direct_noalloc (the allocation-free direct!) against fmmhip_direct through the C ABI.

Bar: bit-identical results (both entry points run the same kernels and the same source-chunk
choice; only the workspace of the chunk sums differs: the library's static per-device array here,
memory the library allocates there). Two calls in a row, the second with a smaller input, so the
second reuses the static workspace that the first left filled; a call whose chunk sums would not fit
the workspace is rejected with FMMHIP_EINVAL before any work."""
import numpy as np
import pytest

import fastmultipole_amd as F
from fastmultipole_amd import _lib, workloads as W

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers through the HIP runtime libfmmhip itself links (torch bundles another HIP
    runtime, which cannot share the device with it in one process)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p.value)
        assert self.hip.hipMemcpy(p.value, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p.value

    def get(self, ptr, shape):
        out = np.zeros(shape)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _direct(ctx, kernel, txyz, sbuf, noalloc):
    D = _Dev()
    try:
        t = D.put(txyz)
        s = D.put(sbuf)
        init = np.zeros((len(txyz), 16))
        init[:, :3] = txyz
        out = D.put(init)
        if noalloc:
            st = ctx.lib.fmmhip_direct_noalloc(ctx._h, kernel, len(txyz), t, len(sbuf), s, 1, 1, 1, _lib.FP64, out)
        else:
            st = ctx.lib.fmmhip_direct(ctx._h, kernel, len(txyz), t, len(sbuf), s, 1, 1, 1, _lib.FP64, out, 1)
        ctx.check(st)
        return D.get(out, (len(txyz), 16))
    finally:
        D.free()


@pytest.mark.parametrize("kernel,gen", [(_lib.POINT_VORTEX, lambda n, s: W.vortex_ring(n, seed=s)),
                                        (_lib.POINT_SOURCE, lambda n, s: W.uniform_sources(n, seed=s))],
                         ids=["vortex", "source"])
def test_direct_noalloc_matches_direct(kernel, gen):
    ctx = F.default_context()
    # few targets against many sources (the split form: chunk sums in the workspace), then a smaller call
    for nt, ns, seed in ((3000, 200000, 1), (700, 50000, 2)):
        print(f"[gpu] direct_noalloc kernel {kernel}: {nt} targets x {ns} sources", flush=True)
        sbuf = gen(ns, seed)
        txyz = np.ascontiguousarray(np.random.default_rng(seed).random((nt, 3)) * 2 - 1)
        want = _direct(ctx, kernel, txyz, sbuf, noalloc=False)
        got = _direct(ctx, kernel, txyz, sbuf, noalloc=True)
        np.testing.assert_array_equal(got, want)
        assert np.abs(want[:, 4:7]).max() > 0


def test_direct_noalloc_no_split_and_rejects():
    """Many targets (one workgroup per target block, no workspace) match too; bad arguments fail with
    FMMHIP_EINVAL without touching the output."""
    ctx = F.default_context()
    sbuf = W.uniform_sources(300000, seed=5)
    txyz = np.ascontiguousarray(sbuf[:, :3])
    print("[gpu] direct_noalloc 300000 x 300000", flush=True)
    want = _direct(ctx, _lib.POINT_SOURCE, txyz, sbuf, noalloc=False)
    got = _direct(ctx, _lib.POINT_SOURCE, txyz, sbuf, noalloc=True)
    np.testing.assert_array_equal(got, want)
    st = ctx.lib.fmmhip_direct_noalloc(ctx._h, 99, 1, None, 1, None, 1, 1, 1, _lib.FP64, None)
    assert st == _lib.FMMHIP_EINVAL
