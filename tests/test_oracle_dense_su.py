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
