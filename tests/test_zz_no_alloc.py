"""igc_graph_build_device_noalloc (every temporary in the library's fixed
scratch array, the CSR in caller buffers, no allocation) against
igc_graph_build_device, the entry point it varies, through the C ABI: equal
row_ptr and col on two calls in a row, the second with a smaller set (the
array keeps the first call's contents), plus the oracle on the smaller set
and the refusal of a too-small col buffer.

The caller's device buffers come from the HIP runtime libigc.so itself links
(hipMalloc through ctypes), so the test needs no second runtime in the process."""
import ctypes

import numpy as np
import pytest

import indexgen_amd as igc
from indexgen_amd import workloads as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    def __init__(self):
        igc.lib()  # loads the runtime libigc.so links
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.held = []

    def alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(int(nbytes), 1)) == 0
        self.held.append(p.value)
        return p.value

    def put(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p: int, count: int, dtype) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free_all(self):
        for p in self.held:
            self.L.hipFree(p)
        self.held = []


def test_noalloc_matches_device_build_twice():
    ctx = igc.Context(0)
    hip = Hip()
    try:
        rng = np.random.default_rng(99)
        big = W.filtered_uniform(rng, 20000, 12)
        small = W.filtered_uniform(rng, 3000, 10)
        for syms, ed in ((big, 4), (small, 3)):
            N, n = syms.shape
            d = hip.put(syms)
            ref = ctx.device_csr_to_host(ctx.graph_build_device(d, N, n, ed))
            cap = int(ref.col.size) + 1000
            print(f"[noalloc] N={N} n={n} editDist={ed} edges={ref.n_edges}", flush=True)
            rp = hip.put(np.full(N + 1, 0xDEADBEEF, dtype=np.uint64))
            col = hip.put(np.full(cap, 0xFFFFFFF9, dtype=np.uint32))
            E = ctx.graph_build_device_noalloc(d, N, n, ed, rp, col, cap)
            assert E == ref.n_edges
            got_rp, got_col = hip.get(rp, N + 1, np.uint64), hip.get(col, cap, np.uint32)
            assert np.array_equal(got_rp, ref.row_ptr)
            assert np.array_equal(got_col[:2 * E], ref.col)
            assert (got_col[2 * E:] == np.uint32(0xFFFFFFF9)).all()  # nothing written past the CSR
        rp_o, col_o = O.graph_csr(small, 3)
        assert np.array_equal(got_rp, rp_o) and np.array_equal(got_col[:2 * E], col_o)
        # a col buffer too small: refused with the edge count
        with pytest.raises(igc.IGCError) as e:
            ctx.graph_build_device_noalloc(d, small.shape[0], small.shape[1], 3, rp, col, 2 * E - 1)
        assert e.value.code == igc.IGC_ERR_INVALID and "col needs" in str(e.value)
    finally:
        hip.free_all()
        ctx.close()
