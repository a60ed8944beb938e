"""Pins for the CPU oracle: each check is an independent route to the same
number (brute-force loops written from the reference's index formulas, dense
state vectors, library factorisations, invariants), chosen so that a dropped
term, a wrong index order or a transposed operand in the oracle fails.
"""
import itertools

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import peps_oracle as O

G, T_HOP, M = 1.0, 1.0, 0.5


def _su_state(nx, ny, gd, D, steps, seed, cplx=False):
    gp = O.init_gauge_peps_random(nx, ny, gd, O.staggered_charges(nx, ny), D, seed, complex_entries=cplx)
    for _ in range(steps):
        O.trotter_step_gauge(gp, G, T_HOP, M, 0.05, D)
    return gp


@pytest.fixture(scope="module")
def st_gd2():
    return _su_state(3, 3, 2, 2, 4, 11)


@pytest.fixture(scope="module")
def st_gd3():
    return _su_state(2, 3, 3, 2, 3, 5, cplx=True)


# ── model ────────────────────────────────────────────────────────────────────

@pytest.mark.parametrize("gd", [2, 3, 4])
def test_site_idx_decode_bijection(gd):
    for ix, iy in [(1, 1), (3, 2), (2, 3), (3, 3)]:
        d, dR, dU = O.site_dims(ix, iy, 3, 3, gd)
        seen = set()
        for p in range(1, d + 1):
            nf, lR, lU = O.decode_phys(p, dR, dU, gd)
            assert O.site_idx(nf, lR, lU, dR, dU, gd) == p
            seen.add((nf, lR, lU))
        assert len(seen) == d


def test_reference_labels_for_odd_gd():
    # gd = 3 is the reference's dg = 1: labels are 2E with E in {-1,0,1}
    assert list(O.link_labels(3)) == [-2, 0, 2]
    assert O.boundary_label(3) == 0 and O.boundary_label(2) == 1
    # decode_phys of the reference: p=1 -> (0, -dg, -dg)
    assert O.decode_phys(1, 3, 3, 3) == (0, -2, -2)


@pytest.mark.parametrize("gd", [2, 3])
def test_hamiltonians_hermitian_and_gate(gd):
    H = O.H_merged_h_site(1, 2, 3, 3, gd, G, T_HOP, M)
    assert np.allclose(H, H.conj().T, atol=0)
    w, V = np.linalg.eigh(H)
    Gt = O.gate(H, 0.05)
    np.testing.assert_allclose(Gt, (V * np.exp(-0.05 * w)) @ V.conj().T, atol=1e-13)
    # hopping moves one fermion across the bond and raises the link by one unit:
    # every non-zero matrix element conserves n_L + n_R
    dL, dRL, dUL = O.site_dims(1, 2, 3, 3, gd)
    dR, dRR, dUR = O.site_dims(2, 2, 3, 3, gd)
    Hh = O.H_hop_h_site(1, 2, 3, 3, gd, T_HOP)
    for i, j in zip(*np.nonzero(Hh)):
        nL1 = O.decode_phys(i // dR + 1, dRL, dUL, gd)[0]
        nR1 = O.decode_phys(i % dR + 1, dRR, dUR, gd)[0]
        nL2 = O.decode_phys(j // dR + 1, dRL, dUL, gd)[0]
        nR2 = O.decode_phys(j % dR + 1, dRR, dUR, gd)[0]
        assert nL1 + nR1 == nL2 + nR2


def test_onsite_weights_sum_to_one():
    nx, ny = 4, 3
    tot = np.zeros((nx, ny))
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            tot[ix - 1, iy - 1] += O.onsite_weight(ix, iy, nx, ny)
            tot[ix, iy - 1] += O.onsite_weight(ix + 1, iy, nx, ny)
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            tot[ix - 1, iy - 1] += O.onsite_weight(ix, iy, nx, ny)
            tot[ix - 1, iy] += O.onsite_weight(ix, iy + 1, nx, ny)
    np.testing.assert_allclose(tot, 1.0)


# ── contractions against the reference's loop formulas ─────────────────────

def test_dl_traced_matches_reference_loop():
    rng = np.random.default_rng(0)
    T = rng.standard_normal((3, 2, 3, 2, 1)) + 1j * rng.standard_normal((3, 2, 3, 2, 1))
    dp, Dl, Dr, Du, Dd = T.shape
    E = np.zeros((Dl * Dl, Dr * Dr, Du * Du, Dd * Dd), complex)
    for d2, d1, u2, u1, r2, r1, l2, l1 in itertools.product(*[range(x) for x in (Dd, Dd, Du, Du, Dr, Dr, Dl, Dl)]):
        v = sum(T[p, l1, r1, u1, d1] * np.conj(T[p, l2, r2, u2, d2]) for p in range(dp))
        E[l1 * Dl + l2, r1 * Dr + r2, u1 * Du + u2, d1 * Dd + d2] = v
    np.testing.assert_allclose(O.dl_traced(T), E, atol=1e-14)


def test_absorb_rows_match_reference_loops():
    rng = np.random.default_rng(1)
    a, b, L2, R2, U2, D2 = 2, 3, 4, 1, 4, 9
    M = rng.standard_normal((a, U2, b)) + 1j * rng.standard_normal((a, U2, b))
    E = rng.standard_normal((L2, R2, U2, D2)) + 1j * rng.standard_normal((L2, R2, U2, D2))
    out = O.absorb_row_from_top([M], [E])[0]
    ref = np.zeros((a * L2, D2, b * R2), complex)
    for rr, dd, ll, bb, aa in itertools.product(range(R2), range(D2), range(L2), range(b), range(a)):
        ref[aa * L2 + ll, dd, bb * R2 + rr] = sum(M[aa, uu, bb] * E[ll, rr, uu, dd] for uu in range(U2))
    np.testing.assert_allclose(out, ref, atol=1e-13)
    Mb = rng.standard_normal((a, D2, b))
    out = O.absorb_row_from_bottom([Mb], [E])[0]
    ref = np.zeros((a * L2, U2, b * R2), complex)
    for rr, uu, ll, bb, aa in itertools.product(range(R2), range(U2), range(L2), range(b), range(a)):
        ref[aa * L2 + ll, uu, bb * R2 + rr] = sum(Mb[aa, dd, bb] * E[ll, rr, uu, dd] for dd in range(D2))
    np.testing.assert_allclose(out, ref, atol=1e-13)


def _mps_dense(mps):
    v = mps[0]
    v = v.reshape(v.shape[0] * v.shape[1], v.shape[2], order="F")
    for A in mps[1:]:
        v = (v @ A.reshape(A.shape[0], -1, order="F")).reshape(-1, A.shape[2], order="F")
    return v[:, 0]


def _seq_truncate_dense(psi, dims, chi, cut=1e-14):
    """Right-to-left sequential SVD truncation of a dense state (independent route)."""
    n = len(dims)
    psi = psi.copy()
    for i in range(n - 1, 0, -1):
        left = int(np.prod(dims[:i]))
        Mx = psi.reshape(left, -1, order="F")
        U, S, Vt = np.linalg.svd(Mx, full_matrices=False)
        k = max(min(chi, int(np.sum(S > cut)), len(S)), 1)
        psi = ((U[:, :k] * S[:k]) @ Vt[:k]).reshape(-1, order="F")
    return psi


def test_compress_mps_exact_and_truncated():
    rng = np.random.default_rng(3)
    dims = [3, 4, 3, 4, 2]
    bonds = [1, 3, 5, 6, 2, 1]
    mps = [rng.standard_normal((bonds[i], dims[i], bonds[i + 1])) for i in range(5)]
    psi = _mps_dense(mps)
    exact = O.compress_mps([m.copy() for m in mps], 100)
    np.testing.assert_allclose(_mps_dense(exact), psi, rtol=1e-12, atol=1e-12 * np.abs(psi).max())
    for chi in (2, 3):
        tr = O.compress_mps([m.copy() for m in mps], chi)
        assert max(m.shape[2] for m in tr) <= chi
        ref = _seq_truncate_dense(psi, dims, chi)
        np.testing.assert_allclose(_mps_dense(tr), ref, atol=1e-12 * np.abs(psi).max())


# ── boundary MPS, environments, energy against the dense state ─────────────

@pytest.mark.parametrize("name", ["st_gd2", "st_gd3"])
def test_norm_env_energy_against_dense(name, request):
    gp = request.getfixturevalue(name)
    psi = O.dense_psi(gp)
    nrm = np.vdot(psi, psi).real
    assert abs(O.peps_norm2(gp, chi=4096) - nrm) < 1e-12 * nrm
    for iy in range(1, gp.ny + 1):
        for ix in range(1, gp.nx):
            E = O.bond_environment_h(gp, ix, iy, 4096)
            th0 = np.einsum("apk,bqk->apqb", E["RA"], E["RB"])
            v = np.einsum("abcd,apqb,cpqd->", E["env"], th0, th0.conj())
            assert abs(v - nrm) < 1e-11 * nrm
            env = E["env"]
            np.testing.assert_allclose(env, np.conj(np.transpose(env, (2, 3, 0, 1))), atol=1e-12 * np.abs(env).max())
    for iy in range(1, gp.ny):
        for ix in range(1, gp.nx + 1):
            E = O.bond_environment_v(gp, ix, iy, 4096)
            th0 = np.einsum("apk,bqk->apqb", E["RA"], E["RB"])
            v = np.einsum("abcd,apqb,cpqd->", E["env"], th0, th0.conj())
            assert abs(v - nrm) < 1e-11 * nrm
    Ed = O.dense_energy(gp, G, T_HOP, M)
    Eb = O.compute_energy_v4(gp, G, T_HOP, M, chi=4096)
    assert abs(Ed - Eb) < 1e-11 * max(1.0, abs(Ed))


def test_su_preserves_gauge_sector_and_filling(st_gd2):
    assert O.forbidden_norm(st_gd2) < 1e-12
    # staggered charges: sum n_f = number of odd sites
    nodd = sum(1 for ix in range(1, 4) for iy in range(1, 4) if (ix + iy) % 2 == 1)
    psi = O.dense_psi(st_gd2)
    dims = O.site_phys_dims(st_gd2)
    T = psi.reshape(dims[::-1])
    prob = np.abs(T) ** 2
    prob /= prob.sum()
    ntot = np.zeros_like(prob)
    for k in range(len(dims)):
        ax = len(dims) - 1 - k
        ix, iy = k % 3 + 1, k // 3 + 1
        nf, _, _ = O.phys_labels(ix, iy, 3, 3, 2)
        shape = [1] * len(dims)
        shape[ax] = dims[k]
        ntot = ntot + nf.reshape(shape)
    assert abs((prob * ntot).sum() - nodd) < 1e-12


# ── truncations against library routines ────────────────────────────────────

def test_charge_svd_block_spectrum():
    rng = np.random.default_rng(5)
    rq = np.array([0, 1, 0, 2, 1, 0, 2, 1])
    cq = np.array([1, 0, 2, 0, 1, 2, 0])
    Th = rng.standard_normal((8, 7))
    Th[rq[:, None] != cq[None, :]] = 0
    U, S, Vt, q = O.charge_svd(Th, rq, cq, 4)
    allS = np.sort(np.linalg.svd(Th, compute_uv=False))[::-1]
    np.testing.assert_allclose(S, allS[:4], rtol=1e-13)
    # reconstruction = best rank-4 approximation (non-degenerate spectrum)
    Uf, Sf, Vf = np.linalg.svd(Th)
    np.testing.assert_allclose((U * S) @ Vt, (Uf[:, :4] * Sf[:4]) @ Vf[:4], atol=1e-12)
    for j in range(len(S)):
        assert set(rq[np.abs(U[:, j]) > 1e-12]) == {q[j]}


def test_solve_herm_matches_pinvh_with_floor():
    rng = np.random.default_rng(6)
    A = rng.standard_normal((12, 5))
    N = A @ A.T  # rank 5
    S = rng.standard_normal((12, 3))
    x = O.solve_herm(N, S, 1e-10)
    w = np.linalg.eigvalsh(N)
    xp = sla.pinvh(N, atol=1e-10 * w.max()) @ S
    np.testing.assert_allclose(x, xp, atol=1e-10)
    B = rng.standard_normal((6, 6))
    N2 = B @ B.T + 6 * np.eye(6)
    np.testing.assert_allclose(O.solve_herm(N2, S[:6], 1e-10), np.linalg.solve(N2, S[:6]), rtol=1e-12)


def test_als_identity_env_is_truncated_svd():
    """With env = δ⊗δ the ALS init (charge SVD) is the optimum of the cost."""
    rng = np.random.default_rng(7)
    nA, dp, nB = 3, 2, 3
    eR = [-1, 1]
    th = rng.standard_normal((nA, dp, dp, nB))
    env = np.einsum("ac,bd->abcd", np.eye(nA), np.eye(nB)).astype(complex)
    X, Y, bf, info = O.als_truncate(th, env, 2, eR, n_als=3, return_info=True)
    XY = np.einsum("apk,bqk->apqb", X, Y)
    # cost equals the squared Frobenius distance under the identity env
    assert abs(O.als_cost(th, env, X, Y) - np.sum(np.abs(th - XY) ** 2)) < 1e-12
    thm = O.jr(O.pd(th, 1, 2, 4, 3), nA * dp, nB * dp)
    rows = np.array([eR[r // nA] for r in range(nA * dp)])
    cols = O.col_charge_from_blocks(thm, rows)
    U, S, Vt, _ = O.charge_svd(thm, rows, cols, 2)
    assert info["cost1"] <= info["cost0"] + 1e-12
    assert abs(info["cost0"] - np.sum(np.abs(thm - (U * S) @ Vt) ** 2)) < 1e-10


def test_als_positive_cost_rule():
    """env = -δ⊗δ makes every cost -||θ - XY||² <= 0.  The floored solve then
    zeroes X (no eigenvalue lies above the floor), so the sweeps end at
    cost1 = -||θ||² < cost0: the reference's rule keeps that refinement, the
    positive-cost rule must keep the SVD initialisation instead."""
    rng = np.random.default_rng(17)
    nA, dp, nB = 3, 2, 3
    eR = [-1, 1]
    th = rng.standard_normal((nA, dp, dp, nB))
    env = -np.einsum("ac,bd->abcd", np.eye(nA), np.eye(nB)).astype(complex)
    X0, Y0, _, i0 = O.als_truncate(th, env, 2, eR, n_als=0, return_info=True)
    assert i0["cost0"] < 0 and not i0["accepted"]
    X, Y, _, info = O.als_truncate(th, env, 2, eR, n_als=2, return_info=True)
    assert info["accepted"]
    assert abs(info["cost1"] + np.sum(th ** 2)) < 1e-10 and info["cost1"] < info["cost0"]
    assert np.abs(X).max() == 0.0
    Xp, Yp, _, ip = O.als_truncate(th, env, 2, eR, n_als=2, return_info=True, positive_cost=True)
    assert not ip["accepted"]
    np.testing.assert_array_equal(Xp, X0)
    np.testing.assert_array_equal(Yp, Y0)


def test_lattice_init_matches_oracle_init():
    from lgt_amd import lattice
    nx, ny = 3, 4
    for gd, D, cp in [(2, 2, False), (3, 2, True), (2, 3, False)]:
        g = O.staggered_charges(nx, ny)
        a = O.init_gauge_peps_random(nx, ny, gd, g, D, 42, complex_entries=cp)
        b = lattice.random_gauge_peps_arrays(nx, ny, gd, g, D, 42, complex_entries=cp)
        for ix in range(nx):
            for iy in range(ny):
                np.testing.assert_allclose(a.tensors[ix][iy], b["tensors"][ix][iy], rtol=0, atol=1e-15)
        assert all(np.array_equal(a.qh[i][j], b["qh"][i][j]) for i in range(nx - 1) for j in range(ny))
        np.testing.assert_array_equal(lattice.staggered_charges(nx, ny), g)


def test_dl_open_matches_reference_loop_and_traces_to_dl_traced():
    """_dl_open (finite_peps_full_update.jl:97-109): the reference's explicit
    loop (il = (l1-1) Dl + l2 in 1-based form, i.e. the bra copy fast), and
    its physical trace is _dl_traced."""
    rng = np.random.default_rng(3)
    T = rng.standard_normal((3, 2, 3, 1, 2)) + 1j * rng.standard_normal((3, 2, 3, 1, 2))
    dp, Dl, Dr, Du, Dd = T.shape
    E = np.zeros((dp, dp, Dl * Dl, Dr * Dr, Du * Du, Dd * Dd), complex)
    for d2 in range(Dd):
        for d1 in range(Dd):
            for u2 in range(Du):
                for u1 in range(Du):
                    for r2 in range(Dr):
                        for r1 in range(Dr):
                            for l2 in range(Dl):
                                for l1 in range(Dl):
                                    il, ir, iu, idd = l1 * Dl + l2, r1 * Dr + r2, u1 * Du + u2, d1 * Dd + d2
                                    for p1 in range(dp):
                                        for p2 in range(dp):
                                            E[p1, p2, il, ir, iu, idd] = T[p1, l1, r1, u1, d1] * np.conj(
                                                T[p2, l2, r2, u2, d2])
    np.testing.assert_allclose(O.dl_open(T), E, atol=1e-14)
    np.testing.assert_allclose(np.einsum("ppabcd->abcd", O.dl_open(T)), O.dl_traced(T), atol=1e-13)
