"""The no-allocation commutator entry (g12_h2_commutator_noalloc) against the
allocating one (g12_h2_commutator), through the C ABI.

Inputs and outputs are device buffers that the test allocates with the HIP
runtime the library itself is linked against (/opt/rocm's libamdhip64; a
second runtime in the process, such as torch's bundled one, would not share
its device state).  Two calls in a row on one stream, the second with a
smaller input, each equal bit for bit to the allocating entry (both run the
same tile kernels); sizes above the static scratch's limit are rejected.
"""
import ctypes
import os

import numpy as np
import pytest

from tests.helpers import rand_c

pytestmark = pytest.mark.gpu

g = pytest.importorskip("g12kba_amd")


def hip_runtime():
    """The libamdhip64 already mapped by libg12kba.so (same runtime instance)."""
    for line in open("/proc/self/maps"):
        path = line.split()[-1] if "/" in line else ""
        if os.path.basename(path).startswith("libamdhip64.so") and "/torch/" not in path:
            return ctypes.CDLL(path)
    return ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")


H = hip_runtime()
vp = ctypes.c_void_p
H.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
H.hipFree.argtypes = [vp]
H.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
H.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
H.hipStreamDestroy.argtypes = [vp]
H.hipStreamSynchronize.argtypes = [vp]
H2D, D2H = 1, 2


class Dev:
    """One device buffer of complex128 elements."""

    def __init__(self, n):
        self.nbytes = 16 * n
        self.ptr = vp()
        assert H.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.nbytes == self.nbytes
        assert H.hipMemcpy(self.ptr, a.ctypes.data_as(vp), self.nbytes, H2D) == 0

    def get(self, shape):
        a = np.empty(shape, np.complex128)
        assert H.hipMemcpy(a.ctypes.data_as(vp), self.ptr, self.nbytes, D2H) == 0
        return a

    def free(self):
        H.hipFree(self.ptr)


def noalloc(G2, hx, hy, Ns, Nb, precision, stream):
    N = Ns * Nb
    bufs = [Dev(N ** 4), Dev(N * N), Dev(N * N), Dev(N ** 4)]
    dG, dx, dy, dout = bufs
    try:
        dG.put(G2)
        dx.put(hx)
        dy.put(hy)
        dout.put(np.full((N * N, N * N), np.nan + 1j * np.nan))  # every element must be written
        rc = g.lib.g12_h2_commutator_noalloc(dG.ptr, dx.ptr, dy.ptr, dout.ptr, Ns, Nb, g.PRECISIONS[precision],
                                             stream)
        assert rc == 0
        assert H.hipStreamSynchronize(stream) == 0
        return dout.get((N * N, N * N))
    finally:
        for b in bufs:
            b.free()


@pytest.mark.parametrize("precision", ["fp64", "fp32_3m", "bf16x3"])
def test_noalloc_commutator_matches_allocating_entry(precision):
    rng = np.random.default_rng({"fp64": 1, "fp32_3m": 2, "bf16x3": 3}[precision])
    stream = vp()
    assert H.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        # first the largest size the scratch holds, then a smaller one: the
        # second call must not see the first call's leftovers
        for Ns, Nb in ((8, 3), (5, 2)):
            N = Ns * Nb
            assert N <= 24
            G2 = rand_c(rng, N * N, N * N)
            hx, hy = rand_c(rng, N, N), rand_c(rng, N, N)
            print(f"noalloc commutator {precision} Ns={Ns} Nb={Nb}", flush=True)
            got = noalloc(G2, hx, hy, Ns, Nb, precision, stream)
            want = g.h2xyG2xyxy_commutator(G2, hx, hy, Ns, Nb, precision)
            assert np.isfinite(got).all()
            assert np.array_equal(got, want), np.abs(got - want).max()
    finally:
        H.hipStreamDestroy(stream)


def test_noalloc_commutator_rejects_sizes_beyond_its_scratch():
    z = vp(1)  # never dereferenced: the size check comes first
    assert g.lib.g12_h2_commutator_noalloc(z, z, z, z, 25, 1, 0, None) == -4
    assert g.lib.g12_h2_commutator_noalloc(z, z, z, z, 5, 5, 0, None) == -4
