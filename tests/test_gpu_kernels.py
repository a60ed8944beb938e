"""MFMA GEMM of every precision tier against numpy (lane layouts, tile
shapes, split-K, operand layouts, conjugation), through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(7, 5, 3), (64, 64, 16), (130, 129, 40), (200, 144, 37), (48, 40, 5000), (16, 300, 2048),
          # large real outputs (the tiled FP32 kernel): ragged edges, K not a
          # multiple of the k-tile, split-K, more tile rows than one group
          (300, 260, 70), (257, 1000, 33), (512, 384, 4096), (1100, 300, 5),
          # both extents >= 512: operands cast to FP32 before the tiled kernel
          # (ragged, K not a multiple of the k-tile; split-K)
          (640, 520, 77), (1030, 700, 1500)]
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


@pytest.mark.parametrize("tier", ["fp64", "fp32"])
def test_gemm_accumulate(ctx, tier):
    """C = op(A) op(B) + beta C for the accumulating contractions, including
    the split-K path (beta applied once, in the reduction)."""
    from lgt_amd import peps
    rng = np.random.default_rng(4)
    ctx.set_tier(tier)
    try:
        for (M, N, K, beta) in [(300, 260, 70, 1.0), (512, 384, 4096, 0.5), (40, 30, 20, -2.0),
                                (1100, 300, 64, 1.0), (600, 530, 300, -0.5), (700, 640, 2048, 1.0)]:
            for ak in (False, True):
                for bk in (False, True):
                    A = rng.standard_normal((M, K))
                    B = rng.standard_normal((K, N))
                    C0 = rng.standard_normal((M, N))
                    out = peps.gemm_acc(A.T.copy() if ak else A, B if bk else B.T.copy(), C0, beta,
                                        a_kfast=ak, b_kfast=bk, ctx=ctx)
                    ref = A @ B + beta * C0
                    scale = np.sqrt(K) * np.abs(A).max() * np.abs(B).max() + abs(beta) * np.abs(C0).max()
                    assert np.abs(out - ref).max() / scale < TOL[tier], (tier, M, N, K, beta, ak, bk)
    finally:
        ctx.set_tier("fp64")


@pytest.mark.parametrize("cplx", [False, True])
def test_permute(ctx, cplx):
    """Every permutation path: pure reshape, 2-D/batched transpose, the
    batched tiled transpose (output-fastest dim not input-contiguous, both
    extents >= 16, ragged tiles) and the general gather."""
    from lgt_amd import peps
    rng = np.random.default_rng(6)
    cases = [((5, 7), (1, 0)), ((40, 33), (1, 0)), ((17, 20, 3), (1, 0, 2)),
             ((18, 5, 40), (2, 0, 1)), ((36, 6, 6, 50), (3, 1, 0, 2)), ((20, 3, 17, 2, 19), (4, 2, 0, 3, 1)),
             ((6, 6, 6, 6, 6, 6), (5, 3, 1, 0, 2, 4)), ((1, 48, 1, 33), (3, 0, 1, 2)), ((33, 1, 70), (2, 1, 0)),
             ((3, 4, 5), (0, 1, 2))]
    for shp, perm in cases:
        T = rng.standard_normal(shp) + (1j * rng.standard_normal(shp) if cplx else 0)
        out = peps.permute(T, perm, ctx=ctx)
        np.testing.assert_array_equal(out, np.transpose(T, perm))


def test_dl_traced_matches_oracle(ctx):
    from lgt_amd import peps
    from oracle import peps_oracle as O
    rng = np.random.default_rng(1)
    for shp, cp in [((8, 3, 2, 4, 1), False), ((18, 2, 3, 2, 2), True), ((2, 5, 1, 1, 4), False)]:
        T = rng.standard_normal(shp) + (1j * rng.standard_normal(shp) if cp else 0)
        E = peps.dl_traced(T, ctx=ctx)
        np.testing.assert_allclose(E, O.dl_traced(T), atol=1e-13 * np.abs(T).max() ** 2 * shp[0])


@pytest.mark.parametrize("cplx", [False, True])
def test_eigh(ctx, cplx):
    """The ALS eigensolver (real: the project's one-launch tridiagonal
    reduction and reflector kernels around a divide-and-conquer tridiagonal
    solve; complex: rocSOLVER heevd) against numpy, including rank-deficient
    inputs."""
    from lgt_amd import peps
    rng = np.random.default_rng(4)
    # every variant of the register-resident reduction and its size limits
    # (64, 96, 144, 208, 256, 288), the global-memory kernel above them
    for n in (1, 2, 3, 7, 16, 33, 64, 65, 96, 97, 130, 144, 145, 208, 209, 250, 256, 257, 288, 289, 300):
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
    # graded spectrum down to 1e-14 with an exact null space (the ALS normal
    # matrices of Gauss-masked bonds look like this)
    for n in (5, 33, 128):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0))
        w0 = np.logspace(0, -14, n)
        w0[: n // 5] = 0.0
        H = (Q * w0) @ np.conj(Q.T)
        H = 0.5 * (H + np.conj(H.T))
        w, V = peps.eigh(H, ctx=ctx)
        assert np.abs((V * w) @ np.conj(V.T) - H).max() < 1e-13 * n
        assert np.abs(np.conj(V.T) @ V - np.eye(n)).max() < 1e-13 * n
        np.testing.assert_allclose(np.sort(w), np.linalg.eigvalsh(H), atol=1e-13 * n)


def test_qr_blocked_backend():
    """The project's blocked Householder QR (PEPS_HIP_QR=blocked at context
    creation; rocSOLVER geqrf/orgqr is the default for real panels wider than
    64 columns): the same R rows as LAPACK, Q orthonormal, Q R = A, across
    several panels, a ragged last panel and m < n."""
    import os
    from lgt_amd import peps
    os.environ["PEPS_HIP_QR"] = "blocked"
    try:
        c2 = peps.Context(0)
    finally:
        del os.environ["PEPS_HIP_QR"]
    rng = np.random.default_rng(13)
    for (m, n) in [(300, 130), (2000, 200), (130, 300), (700, 700)]:
        A = rng.standard_normal((m, n))
        Q, R = peps.qr(A, ctx=c2)
        Qn, Rn = np.linalg.qr(A)
        k = min(m, n)
        assert np.abs(Q @ R - A).max() < 1e-13 * np.abs(A).max() * max(m, n) ** 0.5
        assert np.abs(Q.T @ Q - np.eye(k)).max() < 1e-13 * max(m, n) ** 0.5
        assert np.abs(R - Rn).max() < 1e-11 * np.abs(A).max() * max(m, n) ** 0.5
    c2.close()


@pytest.mark.parametrize("cplx", [False, True])
def test_qr(ctx, cplx):
    """Thin QR against numpy's LAPACK QR: the same R (LAPACK's sign
    convention, so no sign freedom) and Q to rounding, Q orthonormal, Q R = A.
    Up to 64 columns the project's panel kernel; wider, rocSOLVER (real) or
    its unblocked fallback (complex): ragged shapes, wide m < n, exact zero
    and repeated columns."""
    from lgt_amd import peps
    rng = np.random.default_rng(12)
    for (m, n) in [(40, 30), (200, 64), (300, 130), (2000, 65), (5184, 288), (130, 300), (700, 700)]:
        A = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
        A[:, 3] = 0.0
        A[:, 70 % n] = 2.0 * A[:, 1]
        Q, R = peps.qr(A, ctx=ctx)
        k = min(m, n)
        Qn, Rn = np.linalg.qr(A)
        scale = np.abs(A).max()
        assert np.abs(Q @ R - A).max() < 1e-13 * scale * max(m, n) ** 0.5, (m, n)
        assert np.abs(np.conj(Q.T) @ Q - np.eye(k)).max() < 1e-13 * max(m, n) ** 0.5, (m, n)
        assert np.abs(np.tril(R, -1)).max() == 0.0
        # with LAPACK's sign convention the rows of R before the repeated
        # column are unique (the exactly zero column gives tau = 0 in both);
        # the repeated column leaves a rounding-level residual whose
        # reflector, and every row after it, is arbitrary
        cut = min(70 % n, k)
        assert np.abs(R[:cut] - Rn[:cut]).max() < 1e-11 * scale * max(m, n) ** 0.5, (m, n, np.abs(R - Rn).max())


@pytest.mark.parametrize("cplx", [False, True])
def test_svd(ctx, cplx):
    """Thin SVD (QR + one-sided Jacobi) against numpy, including the
    exact-zero columns and rank deficiencies of Gauss-masked blocks."""
    from lgt_amd import peps
    rng = np.random.default_rng(8)
    for (m, n) in [(1, 1), (1, 7), (7, 1), (3, 3), (8, 24), (24, 8), (16, 32), (40, 40), (130, 70),
                   (70, 200), (160, 150), (150, 161), (171, 171), (300, 288), (2000, 60)]:
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
    # a 4x3 sector block with two zero rows whose column inner products fall
    # to ~1e-160 during the sweeps (their squares underflow), and columns
    # scaled to 1e-120 .. 1e-200
    blocks = [np.array([[-7.6510e-02, -4.6293e-03, -1.0015e+00], [0, 0, 0], [0, 0, 0],
                        [6.7905e-03, 4.1073e-04, 8.8862e-02]])]
    for sc in (1e-120, 1e-160, 1e-200):
        A = rng.standard_normal((6, 4)) + (1j * rng.standard_normal((6, 4)) if cplx else 0)
        A[:, 2] *= sc
        A[:, 0] = A[:, 1] * 1e-9 + A[:, 0] * sc
        blocks.append(A)
    for A in blocks:
        U, S, Vt = peps.svd(A, ctx=ctx)
        Sr = np.linalg.svd(A, compute_uv=False)
        np.testing.assert_allclose(S, Sr, atol=1e-14 * Sr.max())
        assert np.abs((U * S) @ Vt - A).max() < 1e-14 * Sr.max()
        nz = S > 0
        Vk = Vt[nz, :]
        assert np.abs(Vk @ np.conj(Vk.T) - np.eye(nz.sum())).max() < 1e-13
    # the all-zero blocks of Gauss-masked matrices: S = 0, identity vectors
    for (m, n) in [(4, 29), (29, 4), (1, 1)]:
        A = np.zeros((m, n), complex if cplx else float)
        U, S, Vt = peps.svd(A, ctx=ctx)
        k = min(m, n)
        assert S.shape == (k,) and np.all(S == 0)
        assert np.abs(np.conj(U.T) @ U - np.eye(k)).max() < 1e-15
        assert np.abs(Vt @ np.conj(Vt.T) - np.eye(k)).max() < 1e-15


@pytest.mark.parametrize("n,scale,garbage", [(300, 1.0, False), (300, 1e-120, False), (300, 1.0, True),
                                             (512, 1.0, True), (700, 3e90, True)])
def test_eigh_large_and_lower_triangle_only(ctx, n, scale, garbage):
    """Real sizes that take the 16- and 32-row register variants of the
    reflector kernel (n = 300, 512, 700), badly scaled entries (the
    tridiagonal reduction rescales by a power of two), and the LAPACK
    convention that only the lower triangle is read: garbage in the strict
    upper triangle must not change the result."""
    from lgt_amd import peps
    rng = np.random.default_rng(8)
    B = rng.standard_normal((n, n))
    H = (B + B.T) * scale
    G = np.tril(H) + (np.triu(rng.standard_normal((n, n)) * 1e3 * scale, 1) if garbage else np.triu(H, 1))
    w, V = peps.eigh(G[None], ctx=ctx)
    w, V = w[0], V[0]
    ref = np.linalg.eigvalsh(H)
    assert np.abs(w - ref).max() < 1e-13 * n * np.abs(ref).max()
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-13 * n
    assert np.abs(H @ V - V * w).max() < 1e-13 * n * np.abs(ref).max()


def test_dl_open_matches_oracle(ctx):
    """_dl_open on the GPU (the open-physical-index double layer) against the
    oracle, real and complex, with unit legs."""
    from lgt_amd import peps
    from oracle import peps_oracle as O
    rng = np.random.default_rng(2)
    for shp, cp in [((8, 3, 2, 4, 1), False), ((18, 2, 3, 2, 2), True), ((2, 1, 1, 1, 1), False)]:
        T = rng.standard_normal(shp) + (1j * rng.standard_normal(shp) if cp else 0)
        E = peps.dl_open(T, ctx=ctx)
        np.testing.assert_allclose(E, O.dl_open(T), rtol=0, atol=1e-15 * np.abs(T).max() ** 2)


@pytest.mark.parametrize("noise", [1e-4, 1.0])
def test_als_init_leading_triplets_at_full_size(ctx, noise):
    """The ALS initialisation at the C3 bond shape (theta 48 x 8 x 8 x 48,
    charge sectors of 192 rows) takes only the leading D_max triplets of each
    sector (randomized subspace iteration, residual-checked).  With a
    dominant rank-6 part (noise 1e-4, as a gated bond looks) or pure noise
    (1.0), the cut at D_max falls inside a flat noise spectrum that the
    randomized iteration cannot certify, and the leading triplets come from
    the eigenvectors of the Jordan-Wielandt matrix [[0, B], [B^T, 0]]
    instead (no full Jacobi SVD either way).  The result equals the oracle's
    full-SVD initialisation: same bond fluxes, X Y to 1e-10, the same
    initial cost, and after 2 ALS sweeps the same cost to 1e-8."""
    from lgt_amd import peps
    from oracle import peps_oracle as O
    ctx.stats_reset()
    rng = np.random.default_rng(11)
    nA, dp, D = 48, 8, 6
    eR = np.array([1, -1] * (dp // 2))
    X = rng.standard_normal((nA, dp, D)) * (eR[None, :, None] == np.array([1, -1, 1, -1, 1, -1])[None, None, :])
    Y = rng.standard_normal((nA, dp, D))
    th = np.einsum("apk,bqk->apqb", X, Y) + noise * rng.standard_normal((nA, dp, dp, nA))
    M = rng.standard_normal((nA * nA, nA * nA + 3))
    env = (M @ M.T).reshape(nA, nA, nA, nA, order="F")
    for n_als in (0, 2):
        Xg, Yg, bg, ig = peps.als_truncate(th, env, D, eR, n_als=n_als, ctx=ctx, return_info=True)
        Xo, Yo, bo, io = O.als_truncate(th, env, D, eR, n_als=n_als, return_info=True)
        assert np.array_equal(np.asarray(bg), np.asarray(bo))
        assert abs(ig["cost0"] - io["cost0"]) <= 1e-10 * abs(io["cost0"])
        if n_als == 0:
            XYg = np.einsum("apk,bqk->apqb", Xg, Yg)
            XYo = np.real(np.einsum("apk,bqk->apqb", Xo, Yo))
            assert np.abs(XYg - XYo).max() <= 1e-10 * np.abs(XYo).max()
        else:
            assert abs(ig["cost1"] - io["cost1"]) <= 1e-8 * abs(io["cost1"])
    st = {k: v["launches"] for k, v in ctx.stats().items() if k.startswith("svd_leading")}
    print(f"leading-triplet paths (noise {noise}): {st}", flush=True)
    assert st.get("svd_leading[fallback]", 0) == 0
    assert st.get("svd_leading[jordan-wielandt]", 0) + st.get("svd_leading[iterations]", 0) > 0


def _blocky_psd(rng, sizes, coupling):
    """A symmetric PSD matrix that is block diagonal after a row permutation:
    each block Q diag(w) Q^T with w log-uniform in [1e-5, 1] times 1, 0.1 or
    0.01, or exactly 0: none near the solve's 1e-10 floor, and the kept part
    conditioned to <= 1e7.  Couplings of relative size `coupling` join every
    pair of rows."""
    n = sum(sizes)
    N = np.zeros((n, n))
    o = 0
    for b, k in enumerate(sizes):
        Q, _ = np.linalg.qr(rng.standard_normal((k, k)))
        w = 10.0 ** rng.uniform(-5, 0, k) * 10.0 ** (-b % 3)
        w[rng.random(k) < 0.4] = 0.0
        N[o:o + k, o:o + k] = (Q * w) @ Q.T
        o += k
    E = rng.standard_normal((n, n)) * coupling * np.abs(N).max()
    N = N + 0.5 * (E + E.T)
    p = rng.permutation(n)
    return N[np.ix_(p, p)]


@pytest.mark.parametrize("case", ["split", "coupled", "dense"])
def test_solve_herm_sector_blocks(ctx, case):
    """_solve_herm (finite_peps_fullupdate_v4.jl:37-45) through peps_solve_herm
    against the oracle's dense eigen-floored solve.  A matrix whose couplings
    between row groups total far below eps ||N||_F is solved block by block
    (the library's counter shows it); couplings above that budget, or no
    block structure, keep the dense solve.  The spectra stay clear of the
    floor, so both agree to 1e-8 relative."""
    from lgt_amd import peps
    from oracle import peps_oracle as O
    rng = np.random.default_rng({"split": 3, "coupled": 4, "dense": 5}[case])
    if case == "dense":
        N = _blocky_psd(rng, [140], 0.0)
    else:
        N = _blocky_psd(rng, [70, 45, 30, 9] + [1] * 6, 1e-20 if case == "split" else 1e-12)
    S = rng.standard_normal((N.shape[0], 8))
    before = ctx.stats().get("als_solve[blocks]", {"launches": 0})["launches"]
    X = peps.solve_herm(N, S, 1e-10, ctx=ctx)
    after = ctx.stats().get("als_solve[blocks]", {"launches": 0})["launches"]
    Xr = O.solve_herm(N.astype(complex), S.astype(complex), 1e-10)
    assert (after - before) == (1 if case == "split" else 0)
    assert np.linalg.norm(X - Xr.real) <= 1e-8 * np.linalg.norm(Xr)


def test_timing_modes():
    """peps_set_timing: mode 2 times only the tiers' MFMA GEMM classes (the
    bench's default, the event pairs of every class cost ~4% of a step), mode
    1 every class, other modes are refused."""
    from lgt_amd import peps
    rng = np.random.default_rng(5)
    A, B = rng.standard_normal((256, 192)), rng.standard_normal((192, 320))
    H = rng.standard_normal((64, 64))
    H = (H + H.T)[None]
    ctx = peps.Context(0)
    try:
        for gemm_only in (True, False):
            ctx.stats_reset()
            ctx.set_timing(True, gemm_only=gemm_only)
            C_ = peps.gemm(A, B, ctx=ctx)
            peps.eigh(H, ctx=ctx)
            st = ctx.stats()
            ctx.set_timing(False)
            np.testing.assert_allclose(C_, A @ B, rtol=0, atol=1e-12 * np.abs(A @ B).max())
            assert st["gemm_f64_tiled"]["ms"] > 0 and st["gemm_f64_tiled"]["launches"] == 1
            timed_other = [k for k, v in st.items() if v["ms"] > 0 and k != "gemm_f64_tiled"]
            assert bool(timed_other) == (not gemm_only), timed_other
        with pytest.raises(peps.PepsError):
            ctx.check(peps.lib().peps_set_timing(ctx.h, 3))
    finally:
        ctx.close()
