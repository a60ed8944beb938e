"""Pins for the oracle's dense simple update (finite_peps_ground_state.jl
update_bond_h!/v!, update_plaquette!), the plaquette gate split
(u1_lgt_hamiltonian.jl decompose_plaquette_LShape) and the simple-update
observables.  Every expected value comes from a different route than the
oracle's code: matrix elements enumerated from the operators' definitions,
dense state vectors with the gate applied leg by leg, numpy SVDs of the dense
two-site block, and product states where the weighted expectation is exact.
"""
import itertools

import numpy as np
import pytest

from oracle import peps_oracle as O
from tests._util import random_peps as _random_peps

G, T_HOP, M = 1.0, 1.0, 0.5


def _site_axes(psi, gp):
    """Dense psi -> tensor with one (nf, R, U) triple of axes per site, sites in
    the order (1,1), (2,1), ..., (nx,ny)."""
    n = gp.nx * gp.ny
    T = psi.reshape(O.site_phys_dims(gp)[::-1])  # axes: site n-1 ... site 0
    T = np.transpose(T, list(range(n))[::-1])     # axes: site 0 ... site n-1
    shp = []
    for iy in range(1, gp.ny + 1):
        for ix in range(1, gp.nx + 1):
            _, dR, dU = O.site_dims(ix, iy, gp.nx, gp.ny, gp.gd)
            shp += [2, dR, dU]  # p = iU + dU*(iR + dR*nf): C order (nf, R, U)
    return T.reshape(shp)


def _from_site_axes(T, gp):
    n = gp.nx * gp.ny
    T = T.reshape(O.site_phys_dims(gp))
    return np.transpose(T, list(range(n))[::-1]).reshape(-1)


def _apply_links(psi, gp, ix, iy, G):
    """Apply a gd^4 x gd^4 gate to the links A.R, A.U, B.U, C.R of the L-shape
    at (ix, iy), taking the gate's flat index as column-major over
    (A.R, A.U, B.U, C.R) (the axis order update_plaquette! assigns)."""
    gd, nx = gp.gd, gp.nx
    T = _site_axes(psi, gp)
    s = lambda jx, jy: 3 * ((jy - 1) * nx + (jx - 1))
    axes = [s(ix, iy) + 1, s(ix, iy) + 2, s(ix + 1, iy) + 2, s(ix, iy + 1) + 1]
    G8 = O.jr(G, *([gd] * 8))  # [a1..a4 (out), b1..b4 (in)]
    out = np.tensordot(G8, T, axes=([4, 5, 6, 7], axes))
    rest = [k for k in range(T.ndim) if k not in axes]
    out = np.transpose(out, np.argsort(axes + rest))
    return _from_site_axes(out, gp)


# ── plaquette operator and its split ─────────────────────────────────────────

@pytest.mark.parametrize("gd", [2, 3])
def test_plaquette_hamiltonian_matrix_elements(gd):
    """-(1/2g^2)(W + W^H) with W = Ud_b U_l Ud_r U_t, enumerated state by state:
    W |b,l,r,t> = |b-1, l+1, r-1, t+1> (indices ascending in E), flat index
    with b slowest (Julia kron order)."""
    g = 1.3
    n = gd ** 4
    W = np.zeros((n, n))
    flat = lambda b, l, r, t: t + gd * (r + gd * (l + gd * b))
    for b, l, r, t in itertools.product(range(gd), repeat=4):
        b2, l2, r2, t2 = b - 1, l + 1, r - 1, t + 1
        if min(b2, r2) >= 0 and max(l2, t2) < gd:
            W[flat(b2, l2, r2, t2), flat(b, l, r, t)] = 1.0
    H = O.H_plaquette_gauge(gd, g)
    np.testing.assert_allclose(H, -(W + W.T) / (2 * g ** 2), atol=1e-15)


@pytest.mark.parametrize("gd,tau", [(2, 0.05), (3, 0.1), (3, 0.0)])
def test_plaquette_split_reconstructs_gate(gd, tau):
    Gp = O.gate(O.H_plaquette_gauge(gd, G), tau)
    GA, GB, GC = O.decompose_plaquette_LShape(Gp, gd, cutoff=O.SVD_CUT)
    assert GA.shape[:4] == (gd,) * 4 and GB.shape[:2] == (gd, gd) and GC.shape[:2] == (gd, gd)
    assert GB.shape[2] <= gd ** 2 and GC.shape[2] <= gd ** 2
    # the reference's own check (reconstruct_plaquette_LShape) to 1e-12
    assert np.abs(O.reconstruct_plaquette_LShape(GA, GB, GC) - Gp).max() < 1e-12
    if tau == 0.0:  # identity gate: rank one in both splits
        assert GB.shape[2] == 1 and GC.shape[2] == 1


def test_plaquette_split_of_a_generic_gate():
    rng = np.random.default_rng(3)
    gd = 2
    Gr = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    GA, GB, GC = O.decompose_plaquette_LShape(Gr, gd)
    assert GB.shape[2] == 4 and GC.shape[2] == 4
    assert np.abs(O.reconstruct_plaquette_LShape(GA, GB, GC) - Gr).max() < 1e-12


# ── dense bond update ────────────────────────────────────────────────────────

@pytest.mark.parametrize("orient,D,Dtr,cplx", [("h", 3, 8, False), ("h", 4, 3, True), ("v", 3, 8, True),
                                               ("v", 4, 2, False)])
def test_dense_su_two_site_closed_form(orient, D, Dtr, cplx):
    """On a two-site lattice psi is the gated block M itself; the update keeps
    L = U sqrt(S), R = sqrt(S) V^H and lambda = S/||S||, so the new dense
    state is U_k S_k^2 V_k^H / ||S_k|| of the SVD of M = G psi."""
    nx, ny = (2, 1) if orient == "h" else (1, 2)
    gd = 3
    gp = _random_peps(nx, ny, gd, D, 7, cplx)
    dims = O.site_phys_dims(gp)
    rng = np.random.default_rng(11)
    n = dims[0] * dims[1]
    Gt = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    Mg = O.apply_two_site(O.dense_psi(gp), dims, 0, 1, Gt).reshape(dims[0], dims[1], order="F")
    U, S, Vh = np.linalg.svd(Mg)
    k = min(Dtr, int(np.sum(S > 1e-12)))
    expect = (U[:, :k] * (S[:k] ** 2 / np.linalg.norm(S[:k]))) @ Vh[:k]
    (O.update_bond_h if orient == "h" else O.update_bond_v)(gp, 1, 1, Gt, Dtr)
    got = O.dense_psi(gp).reshape(dims[0], dims[1], order="F")
    assert np.abs(got - expect).max() < 1e-11 * np.abs(expect).max()
    lam = gp.LH(1, 1) if orient == "h" else gp.LV(1, 1)
    np.testing.assert_allclose(lam, S[:k] / np.linalg.norm(S[:k]), rtol=1e-12)


# ── plaquette update ─────────────────────────────────────────────────────────

@pytest.mark.parametrize("nx,ny,ix,iy,gd,cplx", [(2, 2, 1, 1, 2, False), (3, 2, 2, 1, 2, True),
                                                (3, 2, 1, 1, 3, False)])
def test_plaquette_enlargement_applies_gate_on_links(nx, ny, ix, iy, gd, cplx):
    """Before recompression, the enlarged network's dense state is the gate
    applied on the four plaquette links of the old dense state.  The
    reference contracts the tensors with the gate's row (output) indices, so
    the operator applied is the transpose G^T (equal to G for the real
    symmetric plaquette gate); a generic complex gate makes a reversed link
    order or a missing transpose fail."""
    gp = _random_peps(nx, ny, gd, 2, 5, cplx)
    rng = np.random.default_rng(9)
    n = gd ** 4
    Gr = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    GA, GB, GC = O.decompose_plaquette_LShape(Gr, gd)
    expect = _apply_links(O.dense_psi(gp), gp, ix, iy, Gr.T)
    O.update_plaquette(gp, ix, iy, GA, GB, GC, 64, recompress=False)
    got = O.dense_psi(gp)
    assert np.abs(got - expect).max() < 1e-11 * np.abs(expect).max()


def test_dense_trotter_step_runs_and_normalises():
    gp = _random_peps(3, 3, 2, 2, 1)
    O.trotter_step_dense(gp, G, T_HOP, M, 0.05, 3)
    for iy in range(1, 4):
        for ix in range(1, 3):
            lam = gp.LH(ix, iy)
            assert len(lam) <= 3 and abs(np.linalg.norm(lam) - 1) < 1e-12
    assert all(np.isfinite(gp.T(ix, iy)).all() for ix in range(1, 4) for iy in range(1, 4))


# ── simple-update observables ────────────────────────────────────────────────

def test_su_observables_exact_on_product_states():
    """For a D=1 product state the weighted expectations are exact: on-site
    terms are <phi|H|phi>, plaquettes <psi|H_plaq|psi> with H_plaq applied on
    the four links, and the hopping term of total_energy is <v|H_hop|v> with
    v = conj(phi_R) (x) phi_L (its rho is flattened left-site-fastest)."""
    nx, ny, gd = 3, 3, 2
    gp = _random_peps(nx, ny, gd, 1, 21, cplx=True)
    psi = O.dense_psi(gp)
    nrm = np.vdot(psi, psi).real
    phi = lambda jx, jy: gp.T(jx, jy).reshape(-1)
    E_os = 0.0
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            H = O.H_onsite_site(ix, iy, nx, ny, gd, G, M)
            f = phi(ix, iy)
            e = np.vdot(f, H @ f).real / np.vdot(f, f).real
            assert abs(O.expect_site(gp, ix, iy, H) - e) < 1e-13
            E_os += e
    E_hop = 0.0
    for (ix, iy, jx, jy, H) in ([(ix, iy, ix + 1, iy, O.H_hop_h_site(ix, iy, nx, ny, gd, T_HOP))
                                 for iy in range(1, ny + 1) for ix in range(1, nx)]
                                + [(ix, iy, ix, iy + 1, O.H_hop_v_site(ix, iy, nx, ny, gd, T_HOP))
                                   for iy in range(1, ny) for ix in range(1, nx + 1)]):
        v = np.kron(phi(jx, jy).conj(), phi(ix, iy))
        E_hop += np.vdot(v, H @ v).real / np.vdot(v, v).real
    E_pl = 0.0
    Hp = O.H_plaquette_gauge(gd, G)
    for iy in range(1, ny):
        for ix in range(1, nx):
            e = np.vdot(psi, _apply_links(psi, gp, ix, iy, Hp)).real / nrm
            assert abs(O.plaquette_energy(gp, ix, iy, G) - e) < 1e-13
            E_pl += e
    assert abs(O.total_energy(gp, G, T_HOP, M) - (E_os + E_hop + E_pl)) < 1e-12


def test_expect_site_weights_bonds_by_lambda():
    """On a two-site lattice after a dense update, site L is U sqrt(S) and the
    bond weight S/||S||, so expect_site(L) = tr(O U S^3 U^H) / tr(S^3): the
    power 3 fails if the bond were weighted by sqrt(lambda) or lambda^2."""
    gp = _random_peps(2, 1, 3, 3, 4)
    dims = O.site_phys_dims(gp)
    n = dims[0] * dims[1]
    Gt = O.gate(O.H_merged_h_site(1, 1, 2, 1, 3, G, T_HOP, M), 0.1)
    Mg = O.apply_two_site(O.dense_psi(gp), dims, 0, 1, Gt).reshape(dims[0], dims[1], order="F")
    U, S, _ = np.linalg.svd(Mg)
    O.update_bond_h(gp, 1, 1, Gt, 8)
    rng = np.random.default_rng(2)
    Op = rng.standard_normal((dims[0], dims[0]))
    Op = Op + Op.T
    k = len(gp.LH(1, 1))
    Uk, Sk = U[:, :k], S[:k]
    expect = np.trace(Op @ (Uk * Sk ** 3) @ Uk.conj().T).real / np.sum(Sk ** 3)
    assert abs(O.expect_site(gp, 1, 1, Op) - expect) < 1e-12 * max(1, abs(expect))


@pytest.mark.parametrize("variant", ["dense", "v2"])
def test_su_variants_cut_and_peel(variant):
    """Two-site closed form with a second singular value between 1e-12 and
    1e-10: update_bond_h! (cut 1e-12) keeps it, full_update_bond_h! (cut
    1e-10) drops it; both give U_k S_k^2 V_k^H / ||S_k||, times the peel of
    the six missing legs (weight 1 divided by 1 + 1e-15, resp. 1 + 1e-10)."""
    gp = _random_peps(2, 1, 3, 2, 13)
    gp.lh[0][0] = np.array([1.0, 4e-11])
    dims = O.site_phys_dims(gp)
    Mg = O.dense_psi(gp).reshape(dims[0], dims[1], order="F")
    U, S, Vh = np.linalg.svd(Mg)
    assert len(S) == 2 and 1e-12 < S[1] < 1e-10
    k, peel = (2, 1e-15) if variant == "dense" else (1, 1e-10)
    expect = (U[:, :k] * (S[:k] ** 2 / np.linalg.norm(S[:k]))) @ Vh[:k] / (1 + peel) ** 6
    Id = np.eye(dims[0] * dims[1])
    (O.update_bond_h if variant == "dense" else O.full_update_bond_h_v2)(gp, 1, 1, Id, 4)
    got = O.dense_psi(gp).reshape(dims[0], dims[1], order="F")
    assert len(gp.LH(1, 1)) == k
    assert np.abs(got - expect).max() < 1e-12 * np.abs(expect).max()


def test_v2_peel_regulariser():
    """full_update_bond_v! divides the transverse weights out with
    1/(sqrt(lambda) + 1e-10): on a 2x2 lattice with the identity gate, the
    left legs of the updated pair carry sqrt(l)/(sqrt(l) + 1e-10) relative to
    the dense update's sqrt(l)/(sqrt(l) + 1e-15)."""
    ga = _random_peps(2, 2, 2, 2, 6)
    ga.lh[0][0] = np.array([1.0, 1e-16])
    gb = ga.copy()
    Id = np.eye(4 * 8)
    O.update_bond_v(ga, 1, 1, Id, 8)
    O.full_update_bond_v_v2(gb, 1, 1, Id, 8)
    sq = np.sqrt(ga.LH(1, 1))
    # the right leg (r) of the lower site (1,1) carries lambda_h(1,1)
    ratio = (sq + 1e-15) / (sq + 1e-10)
    Ta, Tb = ga.T(1, 1), gb.T(1, 1)
    for r in range(2):
        na, nb = np.linalg.norm(Ta[:, :, r]), np.linalg.norm(Tb[:, :, r])
        assert abs(nb / na - ratio[r]) < 1e-9 * ratio[r]


def test_bond_entropy_closed_forms():
    assert O.bond_entropy([1.0]) == 0.0
    assert abs(O.bond_entropy(np.ones(5) * 0.3) - np.log(5)) < 1e-14
    p = np.array([0.7, 0.2, 0.1])
    assert abs(O.bond_entropy(np.sqrt(p) * 3.0) + np.sum(p * np.log(p))) < 1e-14
    assert O.bond_entropy([1e-20, 0.0]) == 0.0


def test_measurements_on_basis_product_states():
    """A D=1 product of Gauss-law basis states: <n_f> and <E_R^2 + E_U^2> are
    the labels themselves, and every bond entropy is 0."""
    nx, ny, gd = 3, 2, 3
    gp = _random_peps(nx, ny, gd, 1, 2)
    rng = np.random.default_rng(1)
    nf_exp, e2_exp = np.zeros((nx, ny)), np.zeros((nx, ny))
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d, dR, dU = O.site_dims(ix, iy, nx, ny, gd)
            p = int(rng.integers(1, d + 1))
            T = np.zeros((d, 1, 1, 1, 1), dtype=np.complex128)
            T[p - 1] = 1.0
            gp.setT(ix, iy, T)
            nf, lR, lU = O.decode_phys(p, dR, dU, gd)
            nf_exp[ix - 1, iy - 1] = nf
            e2_exp[ix - 1, iy - 1] = (lR / 2) ** 2 * (ix < nx) + (lU / 2) ** 2 * (iy < ny)
    m = O.measure_all_finite(gp, 0.5)
    np.testing.assert_allclose(m["nf_grid"], nf_exp, atol=1e-14)
    np.testing.assert_allclose(m["E2_grid"], e2_exp, atol=1e-14)
    assert abs(m["nf_mean"] - nf_exp.mean()) < 1e-14 and m["S_mean"] == 0.0 and m["t"] == 0.5
    even = [nf_exp[ix - 1, iy - 1] for iy in range(1, ny + 1) for ix in range(1, nx + 1) if (ix + iy) % 2 == 0]
    assert abs(m["nf_even"] - np.mean(even)) < 1e-14


def _wilson_matrix(gd):
    """W = Ud (x) U (x) Ud (x) U on the links (A.R, A.U, B.U, C.R), flat index
    column-major over them (A.R fastest), built factor by factor."""
    U = O.op_U_gauge(gd)
    Ud = U.conj().T
    W8 = np.einsum("ae,bf,cg,dh->abcdefgh", Ud, U, Ud, U)
    return O.jr(W8, gd ** 4, gd ** 4)


@pytest.mark.parametrize("nx,ny,gd,D,cplx", [(2, 2, 2, 2, False), (3, 2, 2, 2, True), (3, 3, 2, 2, False),
                                            (2, 3, 3, 2, True)])
def test_plaquette_expect_against_dense(nx, ny, gd, D, cplx):
    """Boundary-MPS <W> at exact chi equals <psi|W|psi>/<psi|psi> of the dense
    state with W applied link by link (an independent embedding), for every
    plaquette; also -(1/g^2) Re<W> matches plaquette_energy on a D=1 state."""
    gp = _random_peps(nx, ny, gd, D, 12, cplx)
    psi = O.dense_psi(gp)
    nrm = np.vdot(psi, psi)
    Wm = _wilson_matrix(gd)
    for iy in range(1, ny):
        for ix in range(1, nx):
            expect = np.vdot(psi, _apply_links(psi, gp, ix, iy, Wm)) / nrm
            got = O.plaquette_expect(gp, ix, iy, chi=4096)
            assert abs(got - expect) < 1e-12 * max(1.0, abs(expect)), (ix, iy, got, expect)
    g1 = _random_peps(nx, ny, gd, 1, 3, cplx)
    for iy in range(1, ny):
        for ix in range(1, nx):
            e_su = O.plaquette_energy(g1, ix, iy, 1.0)
            assert abs(-O.plaquette_expect(g1, ix, iy).real - e_su) < 1e-12
