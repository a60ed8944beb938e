"""MFMA GEMM of every precision tier against numpy (lane layouts, tile
shapes, split-K, operand layouts, conjugation), through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(7, 5, 3), (64, 64, 16), (130, 129, 40), (200, 144, 37), (48, 40, 5000), (16, 300, 2048)]
TOL = {"fp64": 1e-13, "fp32": 3e-6, "bf16": 2e-2}


@pytest.mark.parametrize("tier", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("cplx", [False, True])
def test_gemm_layouts(ctx, tier, cplx):
    from lgt_amd import peps
    rng = np.random.default_rng(0)
    ctx.set_tier(tier)
    try:
        for (M, N, K) in SHAPES:
            for ak in (False, True):
                for bk in (False, True):
                    for cA, cB in ([(False, False), (True, False), (False, True)] if cplx else [(False, False)]):
                        A = rng.standard_normal((M, K))
                        B = rng.standard_normal((K, N))
                        if cplx:
                            A = A + 1j * rng.standard_normal((M, K))
                            B = B + 1j * rng.standard_normal((K, N))
                        opA = np.conj(A) if cA else A
                        opB = np.conj(B) if cB else B
                        ref = opA @ opB
                        As = A.T.copy() if ak else A
                        Bs = B if bk else B.T.copy()
                        out = peps.gemm(As, Bs, a_kfast=ak, conjA=cA, b_kfast=bk, conjB=cB, ctx=ctx)
                        scale = np.sqrt(K) * np.abs(A).max() * np.abs(B).max()
                        err = np.abs(out - ref).max() / scale
                        assert err < TOL[tier], (tier, cplx, M, N, K, ak, bk, cA, cB, err)
    finally:
        ctx.set_tier("fp64")


def test_dl_traced_matches_oracle(ctx):
    from lgt_amd import peps
    from oracle import peps_oracle as O
    rng = np.random.default_rng(1)
    for shp, cp in [((8, 3, 2, 4, 1), False), ((18, 2, 3, 2, 2), True), ((2, 5, 1, 1, 4), False)]:
        T = rng.standard_normal(shp) + (1j * rng.standard_normal(shp) if cp else 0)
        E = peps.dl_traced(T, ctx=ctx)
        np.testing.assert_allclose(E, O.dl_traced(T), atol=1e-13 * np.abs(T).max() ** 2 * shp[0])


@pytest.mark.parametrize("cplx", [False, True])
def test_jacobi_eigh(ctx, cplx):
    """The ALS eigensolver against numpy, including rank-deficient inputs."""
    from lgt_amd import peps
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 7, 16, 33, 64, 130, 288):
        nb = 3 if n < 100 else 1
        B = rng.standard_normal((nb, n, n)) + (1j * rng.standard_normal((nb, n, n)) if cplx else 0)
        H = B @ np.conj(np.transpose(B, (0, 2, 1)))
        if n > 4:
            H[0] = H[0][:, :] @ np.diag(np.r_[np.ones(n - 2), 1e-12, 0.0]) @ H[0]
        H = 0.5 * (H + np.conj(np.transpose(H, (0, 2, 1))))
        w, V = peps.eigh(H, ctx=ctx)
        for b in range(nb):
            scale = np.abs(H[b]).max()
            rec = (V[b] * w[b]) @ np.conj(V[b].T)
            assert np.abs(rec - H[b]).max() < 1e-13 * n * scale
            assert np.abs(np.conj(V[b].T) @ V[b] - np.eye(n)).max() < 1e-13 * n
            np.testing.assert_allclose(np.sort(w[b]), np.linalg.eigvalsh(H[b]), atol=1e-13 * n * scale)


@pytest.mark.parametrize("cplx", [False, True])
def test_svd(ctx, cplx):
    """Thin SVD (QR + one-sided Jacobi) against numpy, including the
    exact-zero columns and rank deficiencies of Gauss-masked blocks."""
    from lgt_amd import peps
    rng = np.random.default_rng(8)
    for (m, n) in [(1, 1), (1, 7), (7, 1), (3, 3), (8, 24), (24, 8), (16, 32), (40, 40), (130, 70),
                   (70, 200), (300, 288)]:
        A = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
        if min(m, n) > 3:
            A[:, 1] = 0
            A[2, :] = 0
            A[:, -1] = 3.0 * A[:, 0]
        U, S, Vt = peps.svd(A, ctx=ctx)
        Sr = np.linalg.svd(A, compute_uv=False)
        scale = max(Sr.max(), 1e-300)
        np.testing.assert_allclose(S, Sr, atol=1e-13 * scale * max(m, n) ** 0.5)
        assert np.all(np.diff(S) <= 1e-12 * scale)
        rec = (U * S) @ Vt
        assert np.abs(rec - A).max() < 1e-13 * scale * max(m, n)
        nz = S > 1e-10 * scale
        Uk, Vk = U[:, nz], Vt[nz, :]
        assert np.abs(np.conj(Uk.T) @ Uk - np.eye(nz.sum())).max() < 1e-12
        assert np.abs(Vk @ np.conj(Vk.T) - np.eye(nz.sum())).max() < 1e-12
