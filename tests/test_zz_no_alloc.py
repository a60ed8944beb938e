"""peps_gemm_noalloc against peps_gemm through the C ABI.

The no-allocation GEMM takes device buffers the caller owns and keeps its
split-K partial products in the library's fixed per-device array.  Three
calls in a row, each smaller than the one before: a real GEMM whose split-K
partials would exceed the array (the split count is lowered), a real one that
fits it unchanged, and a complex one.  Each is held to the host-buffer entry
point on the same inputs to 1e-13 relative (the summation order of a lowered
split count differs; FP64 throughout).  Then the error paths that the host
catches before any launch, and K = 0."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def hip():
    # the HIP runtime the library itself links (torch ships another copy,
    # whose device pointers this runtime would not know)
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(max(nbytes, 8))) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.asfortranarray(a)
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype=dtype, order="F")
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _operands(rng, M, N, K, cplx, a_kfast, b_kfast):
    def mk(s):
        x = rng.standard_normal(s)
        return x + 1j * rng.standard_normal(s) if cplx else x
    A = mk((K, M) if a_kfast else (M, K))
    B = mk((K, N) if b_kfast else (N, K))
    return np.asfortranarray(A), np.asfortranarray(B)


def _noalloc(hip, ctx, A, B, M, N, K, cplx, a_kfast, conjA, b_kfast, conjB):
    from lgt_amd import peps
    dt = np.complex128 if cplx else np.float64
    esz = 16 if cplx else 8
    da, db, dc = DevBuf(hip, A.nbytes), DevBuf(hip, B.nbytes), DevBuf(hip, M * N * esz)
    try:
        da.put(A)
        db.put(B)
        peps.gemm_noalloc(M, N, K, da.p.value, db.p.value, dc.p.value, is_complex=cplx, a_kfast=a_kfast,
                          conjA=conjA, b_kfast=b_kfast, conjB=conjB, ctx=ctx)
        ctx.sync()
        return dc.get((M, N), dt)
    finally:
        for b in (da, db, dc):
            b.free()


# (M, N, K, complex, a_kfast, conjA, b_kfast, conjB), each smaller than the last
CASES = [
    (1024, 1536, 4096, False, False, False, True, False),   # partials 75 MiB at 6 splits: lowered to 3
    (300, 200, 4096, False, True, False, False, False),     # 16 splits, 7.3 MiB: as peps_gemm
    (64, 48, 2048, True, False, True, True, False),         # complex, conj(A), 8 splits
]


def test_gemm_noalloc_matches_gemm(ctx, hip):
    from lgt_amd import peps
    rng = np.random.default_rng(5)
    for M, N, K, cplx, ak, cA, bk, cB in CASES:
        print(f"gemm_noalloc case M={M} N={N} K={K} complex={cplx}", flush=True)
        A, B = _operands(rng, M, N, K, cplx, ak, bk)
        got = _noalloc(hip, ctx, A, B, M, N, K, cplx, ak, cA, bk, cB)
        ref = peps.gemm(A, B, a_kfast=ak, conjA=cA, b_kfast=bk, conjB=cB, ctx=ctx)
        assert np.all(np.isfinite(got))
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_gemm_noalloc_k_zero_and_rejections(ctx, hip):
    from lgt_amd import peps
    M, N = 40, 30
    dc = DevBuf(hip, M * N * 8)
    da, db = DevBuf(hip, M * 8), DevBuf(hip, N * 8)
    try:
        dc.put(np.full((M, N), 7.0))
        peps.gemm_noalloc(M, N, 0, da.p.value, db.p.value, dc.p.value, ctx=ctx)
        ctx.sync()
        assert np.all(dc.get((M, N), np.float64) == 0.0)
        host = np.zeros((M, N))
        with pytest.raises(peps.PepsError) as e:  # a host pointer for C
            peps.gemm_noalloc(M, N, 1, da.p.value, db.p.value, host.ctypes.data, ctx=ctx)
        assert "device memory" in str(e.value)
        with pytest.raises(peps.PepsError) as e:  # A holds M*1 doubles, K=2 needs 2*M
            peps.gemm_noalloc(M, N, 2, da.p.value, db.p.value, dc.p.value, ctx=ctx)
        assert "past the end" in str(e.value)
    finally:
        for b in (da, db, dc):
            b.free()
