"""CPU FP64 oracle for the finite-PEPS contraction engine.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``lgt_amd``) never calls it.

It restates, function by function, the Julia reference (NoaFeldman/LGT) in
numpy/complex128, keeping the reference's column-major index conventions:
every ``reshape`` of the reference is ``jr`` (``order='F'``) here and every
``permutedims`` is ``pd`` with the same 1-based permutation.  Each function
cites the reference file:line it follows.

Generalisation beyond the reference (recorded in the design notes):
  * the link dimension is ``gd`` (the reference has ``gd = 2dg+1``); flux
    labels are stored as integers ``2E``.  For odd ``gd`` this is the
    reference's model with every label doubled.
  * a missing boundary link carries the label ``b0 = 0`` for odd ``gd`` (as
    in the reference) and ``b0 = +1`` (E = +1/2) for even ``gd``, so that the
    Gauss law stays satisfiable on 3-link edge sites of the spin-1/2 model.
  * Gauss law in label units: lR + lU - fl - fd == 2 (n_f + q).

Pins (the reference ships no golden vectors and its Julia code cannot run
here): each function is checked in tests/test_oracle.py,
tests/test_oracle_dense_su.py and tests/test_oracle_mlp.py against the
reference's own loop formulas, dense brute force on tiny lattices, library
routines (numpy/scipy SVD, eigh, pinvh), closed forms, finite differences or
invariants; the positive-cost accept rule (no reference counterpart) by
test_als_positive_cost_rule.  Parity unpinned:
``trotter_step_v4`` / ``_recanonicalise`` beyond one bond — their multi-step
trajectories depend on LAPACK phase and tie choices that only the Julia
reference could fix; they are pinned only through the per-bond pieces
(environment, charge SVD, ALS, solve) and the invariants (Gauss sector,
half filling, environment Hermiticity).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

FU_REG = 1e-10      # finite_peps_full_update.jl:336
GI_REG = 1e-12      # finite_peps_gauge_invariant.jl:224
V4_ALS_REG = 1e-10  # finite_peps_fullupdate_v4.jl:30
SVD_CUT = 1e-12     # finite_peps_ground_state.jl:57
CMPS_CUT = 1e-14    # finite_peps_boundary_mps.jl:93


def jr(a, *shape):
    """Julia ``reshape`` (column-major)."""
    return np.reshape(a, shape, order="F")


def pd(a, *perm):
    """Julia ``permutedims`` with a 1-based permutation."""
    return np.ascontiguousarray(np.transpose(a, [p - 1 for p in perm]))


# ════════════════════════════════════════════════════════════════════════════
#  Model (u1_lgt_hamiltonian.jl)
# ════════════════════════════════════════════════════════════════════════════

def link_labels(gd):
    """Flux labels 2E of one link, index order (gauge_idx, u1_lgt_hamiltonian.jl:95)."""
    return np.array([2 * i - (gd - 1) for i in range(gd)], dtype=np.int64)


def boundary_label(gd):
    """Label carried by a missing boundary link (0 in the reference, decode_phys :64)."""
    return 0 if gd % 2 == 1 else 1


def site_dims(ix, iy, nx, ny, gd):
    """u1_lgt_hamiltonian.jl:605-611."""
    d_gR = gd if ix < nx else 1
    d_gU = gd if iy < ny else 1
    return 2 * d_gR * d_gU, d_gR, d_gU


def site_idx(nf, lR, lU, d_gR, d_gU, gd):
    """u1_lgt_hamiltonian.jl:614-618 (labels 2E)."""
    iR = 1 if d_gR == 1 else (lR + gd - 1) // 2 + 1
    iU = 1 if d_gU == 1 else (lU + gd - 1) // 2 + 1
    return nf * d_gR * d_gU + (iR - 1) * d_gU + iU


def decode_phys(p, d_gR, d_gU, gd):
    """finite_peps_gauge_invariant.jl:58-67 (labels 2E, missing -> b0)."""
    p0 = p - 1
    nf = p0 // (d_gR * d_gU)
    rem = p0 % (d_gR * d_gU)
    iR1 = rem // d_gU
    iU0 = rem % d_gU
    b0 = boundary_label(gd)
    lR = b0 if d_gR == 1 else 2 * iR1 - (gd - 1)
    lU = b0 if d_gU == 1 else 2 * iU0 - (gd - 1)
    return nf, lR, lU


def _I(n):
    return np.eye(n, dtype=np.complex128)


def op_E(gd):
    """u1_lgt_hamiltonian.jl:110-117 (E = label/2)."""
    return np.diag(link_labels(gd) / 2.0).astype(np.complex128)


def op_E2(gd):
    E = op_E(gd)
    return E @ E


def op_U_gauge(gd):
    """u1_lgt_hamiltonian.jl:123-130: U|e> = |e+1>."""
    U = np.zeros((gd, gd), dtype=np.complex128)
    for i in range(gd - 1):
        U[i + 1, i] = 1.0
    return U


def op_c():
    """u1_lgt_hamiltonian.jl:138-144."""
    c = np.zeros((2, 2), dtype=np.complex128)
    c[0, 1] = 1.0
    return c


def op_nf():
    """u1_lgt_hamiltonian.jl:150-155."""
    n = np.zeros((2, 2), dtype=np.complex128)
    n[1, 1] = 1.0
    return n


def H_onsite_site(ix, iy, nx, ny, gd, g, m, mu=0.0):
    """u1_lgt_hamiltonian.jl:658-676."""
    d_phys, d_gR, d_gU = site_dims(ix, iy, nx, ny, gd)
    H = np.zeros((d_phys, d_phys), dtype=np.complex128)
    if ix < nx:
        H += (g ** 2 / 2) * np.kron(np.kron(_I(2), op_E2(gd)), _I(d_gU))
    if iy < ny:
        H += (g ** 2 / 2) * np.kron(np.kron(_I(2), _I(d_gR)), op_E2(gd))
    stag = 1 if (ix + iy) % 2 == 0 else -1
    H += m * stag * np.kron(np.kron(op_nf(), _I(d_gR)), _I(d_gU))
    if mu != 0.0:
        H += mu * np.kron(np.kron(op_nf(), _I(d_gR)), _I(d_gU))
    return 0.5 * (H + H.conj().T)


def H_hop_h_site(ix, iy, nx, ny, gd, t):
    """u1_lgt_hamiltonian.jl:689-701."""
    assert ix < nx
    _, d_gR_L, d_gU_L = site_dims(ix, iy, nx, ny, gd)
    _, d_gR_R, d_gU_R = site_dims(ix + 1, iy, nx, ny, gd)
    cdag_U_L = np.kron(np.kron(op_c().conj().T, op_U_gauge(gd)), _I(d_gU_L))
    c_R = np.kron(np.kron(op_c(), _I(d_gR_R)), _I(d_gU_R))
    hop = np.kron(cdag_U_L, c_R)
    H = -t * (hop + hop.conj().T)
    return 0.5 * (H + H.conj().T)


def H_hop_v_site(ix, iy, nx, ny, gd, t):
    """u1_lgt_hamiltonian.jl:789-801."""
    assert iy < ny
    _, d_gR_D, d_gU_D = site_dims(ix, iy, nx, ny, gd)
    _, d_gR_U, d_gU_U = site_dims(ix, iy + 1, nx, ny, gd)
    cdag_U_D = np.kron(np.kron(op_c().conj().T, _I(d_gR_D)), op_U_gauge(gd))
    c_U = np.kron(np.kron(op_c(), _I(d_gR_U)), _I(d_gU_U))
    hop = np.kron(cdag_U_D, c_U)
    H = -t * (hop + hop.conj().T)
    return 0.5 * (H + H.conj().T)


def onsite_weight(ix, iy, nx, ny):
    """u1_lgt_hamiltonian.jl:815-818."""
    nb = int(ix > 1) + int(ix < nx) + int(iy > 1) + int(iy < ny)
    return 1.0 / nb


def H_merged_h_site(ix, iy, nx, ny, gd, g, t, m, mu=0.0):
    """u1_lgt_hamiltonian.jl:829-841."""
    dL = site_dims(ix, iy, nx, ny, gd)[0]
    dR = site_dims(ix + 1, iy, nx, ny, gd)[0]
    HL = H_onsite_site(ix, iy, nx, ny, gd, g, m, mu)
    HR = H_onsite_site(ix + 1, iy, nx, ny, gd, g, m, mu)
    aL = onsite_weight(ix, iy, nx, ny)
    aR = onsite_weight(ix + 1, iy, nx, ny)
    return (H_hop_h_site(ix, iy, nx, ny, gd, t) + aL * np.kron(HL, _I(dR))
            + aR * np.kron(_I(dL), HR))


def H_merged_v_site(ix, iy, nx, ny, gd, g, t, m, mu=0.0):
    """u1_lgt_hamiltonian.jl:851-863."""
    dD = site_dims(ix, iy, nx, ny, gd)[0]
    dU = site_dims(ix, iy + 1, nx, ny, gd)[0]
    HD = H_onsite_site(ix, iy, nx, ny, gd, g, m, mu)
    HU = H_onsite_site(ix, iy + 1, nx, ny, gd, g, m, mu)
    aD = onsite_weight(ix, iy, nx, ny)
    aU = onsite_weight(ix, iy + 1, nx, ny)
    return (H_hop_v_site(ix, iy, nx, ny, gd, t) + aD * np.kron(HD, _I(dU))
            + aU * np.kron(_I(dD), HU))


def gate(H, tau):
    """lgt_gate, u1_lgt_hamiltonian.jl:396: exp(-tau H) (Pade, scaling & squaring)."""
    return sla.expm(-tau * H)


# ════════════════════════════════════════════════════════════════════════════
#  State (finite_peps_ground_state.jl:88-103, finite_peps_gauge_invariant.jl:82-97)
# ════════════════════════════════════════════════════════════════════════════

class GaugePEPS:
    """GaugePEPS: dense masked tensors (p,l,r,u,d) + weights + flux labels.

    Index convention: ``T(ix, iy)`` etc. take the reference's 1-based
    coordinates.
    """

    def __init__(self, nx, ny, gd, gch):
        self.nx, self.ny, self.gd = nx, ny, gd
        self.gch = np.asarray(gch, dtype=np.int64).reshape(nx, ny)
        self.tensors = [[None] * ny for _ in range(nx)]
        self.lh = [[None] * ny for _ in range(nx - 1)]
        self.lv = [[None] * (ny - 1) for _ in range(nx)]
        self.qh = [[None] * ny for _ in range(nx - 1)]
        self.qv = [[None] * (ny - 1) for _ in range(nx)]

    def T(self, ix, iy):
        return self.tensors[ix - 1][iy - 1]

    def setT(self, ix, iy, A):
        self.tensors[ix - 1][iy - 1] = A

    def LH(self, ix, iy):
        return self.lh[ix - 1][iy - 1]

    def LV(self, ix, iy):
        return self.lv[ix - 1][iy - 1]

    def copy(self):
        g = GaugePEPS(self.nx, self.ny, self.gd, self.gch)
        g.tensors = [[t.copy() for t in col] for col in self.tensors]
        g.lh = [[x.copy() for x in col] for col in self.lh]
        g.lv = [[x.copy() for x in col] for col in self.lv]
        g.qh = [[x.copy() for x in col] for col in self.qh]
        g.qv = [[x.copy() for x in col] for col in self.qv]
        return g


def vdims(gp, ix, iy):
    """finite_peps_ground_state.jl:97-103."""
    D_l = len(gp.LH(ix - 1, iy)) if ix > 1 else 1
    D_r = len(gp.LH(ix, iy)) if ix < gp.nx else 1
    D_u = len(gp.LV(ix, iy)) if iy < gp.ny else 1
    D_d = len(gp.LV(ix, iy - 1)) if iy > 1 else 1
    return D_l, D_r, D_u, D_d


def _wt_l(gp, ix, iy):
    return gp.LH(ix - 1, iy) if ix > 1 else np.ones(1)


def _wt_r(gp, ix, iy):
    return gp.LH(ix, iy) if ix < gp.nx else np.ones(1)


def _wt_u(gp, ix, iy):
    return gp.LV(ix, iy) if iy < gp.ny else np.ones(1)


def _wt_d(gp, ix, iy):
    return gp.LV(ix, iy - 1) if iy > 1 else np.ones(1)


def _flux_vec(gp, ix, iy):
    """Flux labels of the (l, r, u, d) legs of site (ix,iy) (gauge_invariant.jl:94-97)."""
    b0 = boundary_label(gp.gd)
    fl = gp.qh[ix - 2][iy - 1] if ix > 1 else np.array([b0])
    fr = gp.qh[ix - 1][iy - 1] if ix < gp.nx else np.array([b0])
    fu = gp.qv[ix - 1][iy - 1] if iy < gp.ny else np.array([b0])
    fd = gp.qv[ix - 1][iy - 2] if iy > 1 else np.array([b0])
    return (np.asarray(fl), np.asarray(fr), np.asarray(fu), np.asarray(fd))


def phys_labels(ix, iy, nx, ny, gd):
    d, dR, dU = site_dims(ix, iy, nx, ny, gd)
    nf = np.zeros(d, np.int64)
    lR = np.zeros(d, np.int64)
    lU = np.zeros(d, np.int64)
    for p in range(1, d + 1):
        nf[p - 1], lR[p - 1], lU[p - 1] = decode_phys(p, dR, dU, gd)
    return nf, lR, lU


def gauss_mask(gp, ix, iy):
    """finite_peps_gauge_invariant.jl:109-136."""
    nf, lR, lU = phys_labels(ix, iy, gp.nx, gp.ny, gp.gd)
    fl, fr, fu, fd = _flux_vec(gp, ix, iy)
    q = gp.gch[ix - 1, iy - 1]
    P = nf[:, None, None, None, None]
    R = lR[:, None, None, None, None]
    U = lU[:, None, None, None, None]
    L_ = fl[None, :, None, None, None]
    Rr = fr[None, None, :, None, None]
    Uu = fu[None, None, None, :, None]
    Dd = fd[None, None, None, None, :]
    return (Rr == R) & (Uu == U) & (R + U - L_ - Dd == 2 * (P + q))


def apply_mask(gp, ix, iy):
    """finite_peps_gauge_invariant.jl:139-146."""
    T = gp.T(ix, iy)
    T[~gauss_mask(gp, ix, iy)] = 0


def forbidden_norm(gp):
    """finite_peps_gauge_invariant.jl:154-164."""
    tot = 0.0
    for iy in range(1, gp.ny + 1):
        for ix in range(1, gp.nx + 1):
            T = gp.T(ix, iy)
            tot += np.sum(np.abs(T[~gauss_mask(gp, ix, iy)]) ** 2)
    return np.sqrt(tot)


def default_bond_labels(gd, D):
    """Flux labels of a fresh bond of dimension D (cycle through the link labels)."""
    labs = link_labels(gd)
    order = sorted(labs, key=lambda x: (abs(x), x))
    return np.array([order[i % gd] for i in range(D)], dtype=np.int64)


def staggered_charges(nx, ny):
    """g = -1 on odd ix+iy (gs_benchmark_v4.jl:46-52, finite_peps_boundary_mps.jl:263-266)."""
    g = np.zeros((nx, ny), np.int64)
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            if (ix + iy) % 2 == 1:
                g[ix - 1, iy - 1] = -1
    return g


def init_gauge_peps_random(nx, ny, gd, gch, D, seed, complex_entries=False):
    """Seeded random gauge-invariant PEPS at bond dimension D.

    Analogue of ``init_gauge_peps`` (finite_peps_gauge_invariant.jl:180-218)
    with the configs' initial state: i.i.d. N(0,1) entries from
    ``numpy.random.default_rng(seed)`` (PCG64) on Gauss-allowed entries only,
    each site normalised, all weights 1, bond labels from
    ``default_bond_labels``.  Entries are drawn site by site, iy outer, ix
    inner, over the full (p,l,r,u,d) array in column-major order.
    """
    gp = GaugePEPS(nx, ny, gd, gch)
    rng = np.random.default_rng(seed)
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            gp.lh[ix - 1][iy - 1] = np.ones(D)
            gp.qh[ix - 1][iy - 1] = default_bond_labels(gd, D)
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            gp.lv[ix - 1][iy - 1] = np.ones(D)
            gp.qv[ix - 1][iy - 1] = default_bond_labels(gd, D)
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d = site_dims(ix, iy, nx, ny, gd)[0]
            shp = (d,) + vdims(gp, ix, iy)
            n = int(np.prod(shp))
            v = rng.standard_normal(n)
            if complex_entries:
                v = v + 1j * rng.standard_normal(n)
            T = jr(v.astype(np.complex128), *shp).copy()
            gp.setT(ix, iy, T)
            apply_mask(gp, ix, iy)
            nrm = np.sqrt(np.sum(np.abs(T) ** 2))
            if nrm > 1e-15:
                T /= nrm
    return gp


def init_gauge_peps(nx, ny, gd, gch, noise=0.0, seed=None):
    """finite_peps_gauge_invariant.jl:180-218: every link at the flux label
    b0 (zero for odd gd, the reference's state; +1, E = +1/2, for even gd,
    which has no zero label), n_f = -g on every site."""
    b0 = boundary_label(gd)
    gp = GaugePEPS(nx, ny, gd, gch)
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            gp.lh[ix - 1][iy - 1] = np.array([1.0])
            gp.qh[ix - 1][iy - 1] = np.array([b0], np.int64)
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            gp.lv[ix - 1][iy - 1] = np.array([1.0])
            gp.qv[ix - 1][iy - 1] = np.array([b0], np.int64)
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d, dR, dU = site_dims(ix, iy, nx, ny, gd)
            arr = np.zeros((d, 1, 1, 1, 1), np.complex128)
            nf0 = -int(gp.gch[ix - 1, iy - 1])
            assert 0 <= nf0 <= 1
            arr[site_idx(nf0, b0, b0, dR, dU, gd) - 1, 0, 0, 0, 0] = 1.0
            gp.setT(ix, iy, arr)
    if noise > 0:
        rng = np.random.default_rng(seed)
        for iy in range(1, ny + 1):
            for ix in range(1, nx + 1):
                T = gp.T(ix, iy)
                T += noise * (rng.standard_normal(T.shape) + 1j * rng.standard_normal(T.shape)) / np.sqrt(2)
                apply_mask(gp, ix, iy)
                nrm = np.sqrt(np.sum(np.abs(T) ** 2))
                if nrm > 1e-15:
                    T /= nrm
    return gp


# ════════════════════════════════════════════════════════════════════════════
#  Charge-resolved SVD and the v3 simple update (finite_peps_gauge_invariant.jl)
# ════════════════════════════════════════════════════════════════════════════

def charge_svd(Th, row_charge, col_charge, D_max, reg=GI_REG):
    """finite_peps_gauge_invariant.jl:236-281."""
    nrow, ncol = Th.shape
    row_charge = np.asarray(row_charge)
    col_charge = np.asarray(col_charge)
    sectors = sorted(set(row_charge.tolist()) & set(col_charge.tolist()))
    u_short, v_short, rows_of, cols_of, svals, charges = [], [], [], [], [], []
    for s in sectors:
        Rs = np.nonzero(row_charge == s)[0]
        Cs = np.nonzero(col_charge == s)[0]
        if len(Rs) == 0 or len(Cs) == 0:
            continue
        U, S, Vt = np.linalg.svd(Th[np.ix_(Rs, Cs)], full_matrices=False)
        kmax = min(D_max, len(S))
        for k in range(kmax):
            if S[k] <= reg:
                continue
            u_short.append(U[:, k]); v_short.append(Vt[k, :])
            rows_of.append(Rs); cols_of.append(Cs)
            svals.append(S[k]); charges.append(s)
    if not svals:
        raise RuntimeError("charge_svd: no singular values above reg in any sector")
    svals = np.array(svals)
    keep = np.argsort(-svals, kind="stable")[:min(D_max, len(svals))]
    Dk = len(keep)
    U = np.zeros((nrow, Dk), np.complex128)
    Vt = np.zeros((Dk, ncol), np.complex128)
    S = np.zeros(Dk)
    bq = np.zeros(Dk, np.int64)
    for j, idx in enumerate(keep):
        U[rows_of[idx], j] = u_short[idx]
        Vt[j, cols_of[idx]] = v_short[idx]
        S[j] = svals[idx]
        bq[j] = charges[idx]
    return U, S, Vt, bq


def simple_update_bond_h(gp, ix, iy, G, D_max):
    """finite_peps_gauge_invariant.jl:296-404."""
    nxv, nyv, gd = gp.nx, gp.ny, gp.gd
    apply_mask(gp, ix, iy); apply_mask(gp, ix + 1, iy)
    T_L, T_R = gp.T(ix, iy), gp.T(ix + 1, iy)
    dp_L, dp_R = T_L.shape[0], T_R.shape[0]
    _, dgRL, dgUL = site_dims(ix, iy, nxv, nyv, gd)
    _, dgRR, dgUR = site_dims(ix + 1, iy, nxv, nyv, gd)
    lam = gp.LH(ix, iy)
    bond_dim = len(lam)
    l_l = gp.LH(ix - 1, iy) if ix > 1 else np.ones(1)
    l_r = gp.LH(ix + 1, iy) if ix + 1 < nxv else np.ones(1)
    l_uL = gp.LV(ix, iy) if iy < nyv else np.ones(1)
    l_dL = gp.LV(ix, iy - 1) if iy > 1 else np.ones(1)
    l_uR = gp.LV(ix + 1, iy) if iy < nyv else np.ones(1)
    l_dR = gp.LV(ix + 1, iy - 1) if iy > 1 else np.ones(1)
    sqb = np.sqrt(lam)
    sq_l, sq_r = np.sqrt(l_l), np.sqrt(l_r)
    sq_uL, sq_dL, sq_uR, sq_dR = map(np.sqrt, (l_uL, l_dL, l_uR, l_dR))
    L_w = T_L * sq_l[None, :, None, None, None] * sqb[None, None, :, None, None] \
        * sq_uL[None, None, None, :, None] * sq_dL[None, None, None, None, :]
    R_w = T_R * sqb[None, :, None, None, None] * sq_r[None, None, :, None, None] \
        * sq_uR[None, None, None, :, None] * sq_dR[None, None, None, None, :]
    szL, szR = L_w.shape, R_w.shape
    L_mat = jr(pd(L_w, 1, 2, 4, 5, 3), dp_L * szL[1] * szL[3] * szL[4], bond_dim)
    R_mat = jr(pd(R_w, 2, 1, 3, 4, 5), bond_dim, dp_R * szR[2] * szR[3] * szR[4])
    Th = jr(L_mat @ R_mat, dp_L, szL[1], szL[3], szL[4], dp_R, szR[2], szR[3], szR[4])
    sz = Th.shape
    Tpp = jr(pd(Th, 5, 1, 2, 3, 4, 6, 7, 8), dp_R * dp_L, sz[1] * sz[2] * sz[3] * sz[5] * sz[6] * sz[7])
    Tnew = pd(jr(G @ Tpp, dp_R, dp_L, sz[1], sz[2], sz[3], sz[5], sz[6], sz[7]), 2, 3, 4, 5, 1, 6, 7, 8)
    Tm = jr(Tnew, dp_L * sz[1] * sz[2] * sz[3], dp_R * sz[5] * sz[6] * sz[7]).copy()
    nrow, ncol = Tm.shape
    row_charge = np.array([decode_phys(r0 % dp_L + 1, dgRL, dgUL, gd)[1] for r0 in range(nrow)])
    b0 = boundary_label(gd)
    q = gp.gch[ix, iy - 1]
    col_charge = np.zeros(ncol, np.int64)
    for c0 in range(ncol):
        pR = c0 % dp_R + 1
        rest = c0 // dp_R
        dR0 = (rest // sz[5]) // sz[6]
        f_dR = gp.qv[ix][iy - 2][dR0] if iy > 1 else b0
        nfR, eRR, eUR = decode_phys(pR, dgRR, dgUR, gd)
        col_charge[c0] = eRR + eUR - f_dR - 2 * nfR - 2 * q
    Tm[row_charge[:, None] != col_charge[None, :]] = 0
    U, S, Vt, newq = charge_svd(Tm, row_charge, col_charge, D_max)
    Dk = len(S)
    lam_new = S / max(np.linalg.norm(S), 1e-15)
    sqS = np.sqrt(S)
    L_new = pd(jr(U * sqS[None, :], dp_L, sz[1], sz[2], sz[3], Dk), 1, 2, 5, 3, 4)
    R_new = pd(jr(sqS[:, None] * Vt, Dk, dp_R, sz[5], sz[6], sz[7]), 2, 1, 3, 4, 5)
    inv = lambda s: 1.0 / (s + FU_REG)
    L_new *= inv(sq_l)[None, :, None, None, None] * inv(sq_uL)[None, None, None, :, None] \
        * inv(sq_dL)[None, None, None, None, :]
    R_new *= inv(sq_r)[None, None, :, None, None] * inv(sq_uR)[None, None, None, :, None] \
        * inv(sq_dR)[None, None, None, None, :]
    gp.setT(ix, iy, L_new); gp.setT(ix + 1, iy, R_new)
    gp.lh[ix - 1][iy - 1] = lam_new
    gp.qh[ix - 1][iy - 1] = newq
    apply_mask(gp, ix, iy); apply_mask(gp, ix + 1, iy)


def simple_update_bond_v(gp, ix, iy, G, D_max):
    """finite_peps_gauge_invariant.jl:412-516."""
    nxv, nyv, gd = gp.nx, gp.ny, gp.gd
    apply_mask(gp, ix, iy); apply_mask(gp, ix, iy + 1)
    T_D, T_U = gp.T(ix, iy), gp.T(ix, iy + 1)
    dp_D, dp_U = T_D.shape[0], T_U.shape[0]
    _, dgRD, dgUD = site_dims(ix, iy, nxv, nyv, gd)
    _, dgRU, dgUU = site_dims(ix, iy + 1, nxv, nyv, gd)
    lam = gp.LV(ix, iy)
    bond_dim = len(lam)
    l_lD = gp.LH(ix - 1, iy) if ix > 1 else np.ones(1)
    l_rD = gp.LH(ix, iy) if ix < nxv else np.ones(1)
    l_dD = gp.LV(ix, iy - 1) if iy > 1 else np.ones(1)
    l_lU = gp.LH(ix - 1, iy + 1) if ix > 1 else np.ones(1)
    l_rU = gp.LH(ix, iy + 1) if ix < nxv else np.ones(1)
    l_uU = gp.LV(ix, iy + 1) if iy + 1 < nyv else np.ones(1)
    sqb = np.sqrt(lam)
    sq_lD, sq_rD, sq_dD, sq_lU, sq_rU, sq_uU = map(np.sqrt, (l_lD, l_rD, l_dD, l_lU, l_rU, l_uU))
    D_w = T_D * sq_lD[None, :, None, None, None] * sq_rD[None, None, :, None, None] \
        * sqb[None, None, None, :, None] * sq_dD[None, None, None, None, :]
    U_w = T_U * sq_lU[None, :, None, None, None] * sq_rU[None, None, :, None, None] \
        * sq_uU[None, None, None, :, None] * sqb[None, None, None, None, :]
    szD, szU = D_w.shape, U_w.shape
    D_mat = jr(pd(D_w, 1, 2, 3, 5, 4), dp_D * szD[1] * szD[2] * szD[4], bond_dim)
    U_mat = jr(pd(U_w, 5, 1, 2, 3, 4), bond_dim, dp_U * szU[1] * szU[2] * szU[3])
    Th = jr(D_mat @ U_mat, dp_D, szD[1], szD[2], szD[4], dp_U, szU[1], szU[2], szU[3])
    sz = Th.shape
    Tpp = jr(pd(Th, 5, 1, 2, 3, 4, 6, 7, 8), dp_U * dp_D, sz[1] * sz[2] * sz[3] * sz[5] * sz[6] * sz[7])
    Tnew = pd(jr(G @ Tpp, dp_U, dp_D, sz[1], sz[2], sz[3], sz[5], sz[6], sz[7]), 2, 3, 4, 5, 1, 6, 7, 8)
    Tm = jr(Tnew, dp_D * sz[1] * sz[2] * sz[3], dp_U * sz[5] * sz[6] * sz[7]).copy()
    nrow, ncol = Tm.shape
    row_charge = np.array([decode_phys(r0 % dp_D + 1, dgRD, dgUD, gd)[2] for r0 in range(nrow)])
    b0 = boundary_label(gd)
    q = gp.gch[ix - 1, iy]
    col_charge = np.zeros(ncol, np.int64)
    for c0 in range(ncol):
        pU = c0 % dp_U + 1
        rest = c0 // dp_U
        lU0 = rest % sz[5]
        f_lU = gp.qh[ix - 2][iy][lU0] if ix > 1 else b0
        nfU, eRU, eUU = decode_phys(pU, dgRU, dgUU, gd)
        col_charge[c0] = eRU + eUU - f_lU - 2 * nfU - 2 * q
    Tm[row_charge[:, None] != col_charge[None, :]] = 0
    U, S, Vt, newq = charge_svd(Tm, row_charge, col_charge, D_max)
    Dk = len(S)
    lam_new = S / max(np.linalg.norm(S), 1e-15)
    sqS = np.sqrt(S)
    D_new = pd(jr(U * sqS[None, :], dp_D, sz[1], sz[2], sz[3], Dk), 1, 2, 3, 5, 4)
    U_new = pd(jr(sqS[:, None] * Vt, Dk, dp_U, sz[5], sz[6], sz[7]), 2, 3, 4, 5, 1)
    inv = lambda s: 1.0 / (s + FU_REG)
    D_new *= inv(sq_lD)[None, :, None, None, None] * inv(sq_rD)[None, None, :, None, None] \
        * inv(sq_dD)[None, None, None, None, :]
    U_new *= inv(sq_lU)[None, :, None, None, None] * inv(sq_rU)[None, None, :, None, None] \
        * inv(sq_uU)[None, None, None, :, None]
    gp.setT(ix, iy, D_new); gp.setT(ix, iy + 1, U_new)
    gp.lv[ix - 1][iy - 1] = lam_new
    gp.qv[ix - 1][iy - 1] = newq
    apply_mask(gp, ix, iy); apply_mask(gp, ix, iy + 1)


def trotter_step_gauge(gp, g, t_hop, m, tau, D_max):
    """finite_peps_gauge_invariant.jl:523-539."""
    nxv, nyv, gd = gp.nx, gp.ny, gp.gd
    for iy in range(1, nyv + 1):
        for ix in range(1, nxv):
            simple_update_bond_h(gp, ix, iy, gate(H_merged_h_site(ix, iy, nxv, nyv, gd, g, t_hop, m), tau / 2), D_max)
    for iy in range(1, nyv):
        for ix in range(1, nxv + 1):
            simple_update_bond_v(gp, ix, iy, gate(H_merged_v_site(ix, iy, nxv, nyv, gd, g, t_hop, m), tau), D_max)
    for iy in range(1, nyv + 1):
        for ix in range(nxv - 1, 0, -1):
            simple_update_bond_h(gp, ix, iy, gate(H_merged_h_site(ix, iy, nxv, nyv, gd, g, t_hop, m), tau / 2), D_max)


# ════════════════════════════════════════════════════════════════════════════
#  Boundary MPS (finite_peps_boundary_mps.jl)
# ════════════════════════════════════════════════════════════════════════════

def sqrt_absorbed_site(gp, ix, iy):
    """finite_peps_boundary_mps.jl:41-53."""
    T = gp.T(ix, iy)
    sl, sr = np.sqrt(_wt_l(gp, ix, iy)), np.sqrt(_wt_r(gp, ix, iy))
    su, sd = np.sqrt(_wt_u(gp, ix, iy)), np.sqrt(_wt_d(gp, ix, iy))
    return T * sl[None, :, None, None, None] * sr[None, None, :, None, None] \
        * su[None, None, None, :, None] * sd[None, None, None, None, :]


def dl_traced(T):
    """_dl_traced, finite_peps_full_update.jl:79-91 (bra is the fast sub-index)."""
    dp, Dl, Dr, Du, Dd = T.shape
    E8 = np.einsum("pabcd,pefgh->eafbgchd", T, T.conj(), optimize=True)
    return jr(E8, Dl * Dl, Dr * Dr, Du * Du, Dd * Dd)


def dl_open(T):
    """_dl_open, finite_peps_full_update.jl:97-109: E[p1,p2,(l1,l2),...] =
    T[p1,l1,..] conj(T[p2,l2,..]) with l2 (bra) the fast sub-index."""
    dp, Dl, Dr, Du, Dd = T.shape
    E10 = np.einsum("pabcd,qefgh->pqeafbgchd", T, T.conj(), optimize=True)
    return jr(E10, dp, dp, Dl * Dl, Dr * Dr, Du * Du, Dd * Dd)


def dl_traced_site(gp, ix, iy):
    """finite_peps_boundary_mps.jl:56-57."""
    return dl_traced(sqrt_absorbed_site(gp, ix, iy))


def compress_mps(mps, chi, cutoff=CMPS_CUT, relative=False):
    """finite_peps_boundary_mps.jl:70-105 (in place, returns the list).

    relative=True (the library's PEPS_STEP_RENORM, no reference counterpart)
    keeps singular values above cutoff * S[0] of each bond instead of above
    the absolute cutoff."""
    n = len(mps)
    if n == 1:
        return mps
    for i in range(n - 1):
        Dl, dp, Dr = mps[i].shape
        Q, R = np.linalg.qr(jr(mps[i], Dl * dp, Dr))
        k = Q.shape[1]
        mps[i] = jr(Q, Dl, dp, k)
        Dl2, dp2, Dr2 = mps[i + 1].shape
        mps[i + 1] = jr(R @ jr(mps[i + 1], Dl2, dp2 * Dr2), k, dp2, Dr2)
    for i in range(n - 1, 0, -1):
        Dl, dp, Dr = mps[i].shape
        U, S, Vt = np.linalg.svd(jr(mps[i], Dl, dp * Dr), full_matrices=False)
        thr = cutoff * S[0] if (relative and len(S)) else cutoff
        k = max(min(chi, int(np.sum(S > thr)), len(S)), 1)
        mps[i] = jr(Vt[:k, :], k, dp, Dr)
        Dl2, dp2, Dr2 = mps[i - 1].shape
        mps[i - 1] = jr(jr(mps[i - 1], Dl2 * dp2, Dr2) @ (U[:, :k] * S[None, :k]), Dl2, dp2, k)
    return mps


def row_dl(gp, iy):
    """finite_peps_boundary_mps.jl:112."""
    return [dl_traced_site(gp, ix, iy) for ix in range(1, gp.nx + 1)]


def row_as_mps_down(row):
    """finite_peps_boundary_mps.jl:115-119."""
    return [jr(pd(E, 1, 4, 2, 3), E.shape[0], E.shape[3], E.shape[1] * E.shape[2]) for E in row]


def row_as_mps_up(row):
    """finite_peps_boundary_mps.jl:196-200."""
    return [jr(pd(E, 1, 3, 2, 4), E.shape[0], E.shape[2], E.shape[1] * E.shape[3]) for E in row]


def absorb_row_from_top(mps, row):
    """finite_peps_boundary_mps.jl:128-148: newM[(a,l),d,(b,r)], l fast."""
    out = []
    for M, E in zip(mps, row):
        a, u, b = M.shape
        L2, R2, U2, D2 = E.shape
        assert U2 == u
        T = np.einsum("aub,lrud->ladrb", M, E, optimize=True)
        out.append(jr(T, L2 * a, D2, R2 * b))
    return out


def absorb_row_from_bottom(mps, row):
    """finite_peps_boundary_mps.jl:156-176."""
    out = []
    for M, E in zip(mps, row):
        a, d, b = M.shape
        L2, R2, U2, D2 = E.shape
        assert D2 == d
        T = np.einsum("adb,lrud->laurb", M, E, optimize=True)
        out.append(jr(T, L2 * a, U2, R2 * b))
    return out


def _trivial_mps(n):
    return [np.ones((1, 1, 1), np.complex128) for _ in range(n)]


def normalise_mps(mps, canonical):
    """The rescaling of PEPS_STEP_RENORM (no reference counterpart): a
    compressed boundary is right-canonical (every tensor but the first is a
    right isometry, compress_mps's right-to-left sweep), so its norm is
    ||mps[0]|| and mps[0] alone is divided by it; a raw row MPS has every
    tensor divided by its own norm.  Boundaries enter the full update and the
    energy only through ratios, so the scale never changes a result."""
    for i in range(1 if canonical else len(mps)):
        n = np.linalg.norm(mps[i])
        if np.isfinite(n) and n > 0:
            mps[i] = mps[i] / n
    return mps


def boundary_top(gp, stop_row, chi, renorm=False):
    """finite_peps_boundary_mps.jl:183-193.  renorm: relative compression cut
    and normalisation after every absorption (PEPS_STEP_RENORM)."""
    if stop_row >= gp.ny:
        return _trivial_mps(gp.nx)
    mps = row_as_mps_down(row_dl(gp, gp.ny))
    if renorm:
        normalise_mps(mps, False)
    for r in range(gp.ny - 1, stop_row, -1):
        mps = absorb_row_from_top(mps, row_dl(gp, r))
        compress_mps(mps, chi, relative=renorm)
        if renorm:
            normalise_mps(mps, True)
    return mps


def boundary_bottom(gp, stop_row, chi, renorm=False):
    """finite_peps_boundary_mps.jl:207-216 (renorm as boundary_top)."""
    if stop_row <= 1:
        return _trivial_mps(gp.nx)
    mps = row_as_mps_up(row_dl(gp, 1))
    if renorm:
        normalise_mps(mps, False)
    for r in range(2, stop_row):
        mps = absorb_row_from_bottom(mps, row_dl(gp, r))
        compress_mps(mps, chi, relative=renorm)
        if renorm:
            normalise_mps(mps, True)
    return mps


def peps_norm2(gp, chi=64):
    """finite_peps_boundary_mps.jl:228-245."""
    mps = row_as_mps_down(row_dl(gp, gp.ny))
    for r in range(gp.ny - 1, 0, -1):
        mps = absorb_row_from_top(mps, row_dl(gp, r))
        compress_mps(mps, chi)
    acc = np.ones((1, 1), np.complex128)
    for M in mps:
        a, p, b = M.shape
        assert p == 1
        acc = acc @ jr(M, a, b)
    return acc[0, 0]


# ── bond environments ────────────────────────────────────────────────────────

def split_dl(E, Dl, Dr, Du, Dd):
    """finite_peps_boundary_mps.jl:308-310 (bra, ket) pairs."""
    return jr(E, Dl, Dl, Dr, Dr, Du, Du, Dd, Dd)


def split_mps(M, DX):
    """finite_peps_boundary_mps.jl:313-314 (vL, x_bra, x_ket, vR)."""
    return jr(M, M.shape[0], DX, DX, M.shape[2])


def _qr(M):
    Q, R = np.linalg.qr(M)
    return Q, R


def qr_reduce_h_left(A):
    """finite_peps_boundary_mps.jl:325-334."""
    dp, Dl, Db, Du, Dd = A.shape
    Q, R = _qr(jr(pd(A, 2, 4, 5, 1, 3), Dl * Du * Dd, dp * Db))
    a = Q.shape[1]
    return jr(Q, Dl, Du, Dd, a), jr(R, a, dp, Db)


def qr_reduce_h_right(B):
    """finite_peps_boundary_mps.jl:342-351."""
    dp, Db, Dr, Du, Dd = B.shape
    Q, R = _qr(jr(pd(B, 3, 4, 5, 1, 2), Dr * Du * Dd, dp * Db))
    b = Q.shape[1]
    return jr(Q, Dr, Du, Dd, b), jr(R, b, dp, Db)


def qr_reduce_v_down(A):
    """finite_peps_boundary_mps.jl:448-453."""
    dp, Dl, Dr, Db, Dd = A.shape
    Q, R = _qr(jr(pd(A, 2, 3, 5, 1, 4), Dl * Dr * Dd, dp * Db))
    a = Q.shape[1]
    return jr(Q, Dl, Dr, Dd, a), jr(R, a, dp, Db)


def qr_reduce_v_up(B):
    """finite_peps_boundary_mps.jl:457-462."""
    dp, Dl, Dr, Du, Db = B.shape
    Q, R = _qr(jr(pd(B, 2, 3, 4, 1, 5), Dl * Dr * Du, dp * Db))
    b = Q.shape[1]
    return jr(Q, Dl, Dr, Du, b), jr(R, b, dp, Db)


def _isqrt(n):
    r = int(round(np.sqrt(n)))
    assert r * r == n
    return r


def _dims4(E):
    return tuple(_isqrt(s) for s in E.shape)


def spectator_left(Lenv, topj, botj, Ej):
    """finite_peps_boundary_mps.jl:357-365."""
    Dl, Dr, Du, Dd = _dims4(Ej)
    E8 = split_dl(Ej, Dl, Dr, Du, Dd)
    tp = split_mps(topj, Du)
    bt = split_mps(botj, Dd)
    # the four-tensor sum taken pairwise in a fixed order (chi^3 D^4 + chi^2
    # D^8 + chi^3 D^4 MACs; numpy's greedy path can build a chi^2 D^6
    # intermediate here)
    x = np.einsum("ABkb,AvuC->BkbvuC", Lenv, tp, optimize=True)
    x = np.einsum("BkbvuC,bkxyvued->BCxyed", x, E8, optimize=True)
    return np.einsum("BCxyed,BedF->CFyx", x, bt, optimize=True)


def spectator_right(Renv, topj, botj, Ej):
    """finite_peps_boundary_mps.jl:369-377."""
    Dl, Dr, Du, Dd = _dims4(Ej)
    E8 = split_dl(Ej, Dl, Dr, Du, Dd)
    tp = split_mps(topj, Du)
    bt = split_mps(botj, Dd)
    x = np.einsum("CFyx,AvuC->FyxAvu", Renv, tp, optimize=True)
    x = np.einsum("FyxAvu,bkxyvued->FAbked", x, E8, optimize=True)
    return np.einsum("FAbked,BedF->ABkb", x, bt, optimize=True)


def bond_environment_h(gp, ix, iy, chi=64, top=None, bot=None, renorm=False, Lenv=None, Renv=None):
    """finite_peps_boundary_mps.jl:389-436 (renorm: boundaries as boundary_top;
    Lenv / Renv: the strip environments left of ix and right of ix+1 when
    already known, otherwise built from the boundaries as the reference does)."""
    nxv = gp.nx
    if top is None:
        top = boundary_top(gp, iy, chi, renorm)
    if bot is None:
        bot = boundary_bottom(gp, iy, chi, renorm)
    row = row_dl(gp, iy)
    A = sqrt_absorbed_site(gp, ix, iy)
    B = sqrt_absorbed_site(gp, ix + 1, iy)
    QA, RA = qr_reduce_h_left(A)
    QB, RB = qr_reduce_h_right(B)
    DuA, DdA = A.shape[3], A.shape[4]
    DuB, DdB = B.shape[3], B.shape[4]
    if Lenv is None:
        Lenv = np.ones((1, 1, 1, 1), np.complex128)
        for j in range(1, ix):
            Lenv = spectator_left(Lenv, top[j - 1], bot[j - 1], row[j - 1])
    if Renv is None:
        Renv = np.ones((1, 1, 1, 1), np.complex128)
        for j in range(nxv, ix + 1, -1):
            Renv = spectator_right(Renv, top[j - 1], bot[j - 1], row[j - 1])
    tpA, btA = split_mps(top[ix - 1], DuA), split_mps(bot[ix - 1], DdA)
    # GA[a,c,C,F] = Lenv[A,B,k,b] tpA[A,v,u,C] btA[B,e,d,F] QA[k,u,d,a] QA*[b,v,e,c],
    # pairwise in a fixed order
    x = np.einsum("ABkb,AvuC->BkbvuC", Lenv, tpA, optimize=True)
    x = np.einsum("BkbvuC,kuda->BbvCda", x, QA, optimize=True)
    x = np.einsum("BbvCda,BedF->bvCaeF", x, btA, optimize=True)
    GA = np.einsum("bvCaeF,bvec->acCF", x, QA.conj(), optimize=True)
    tpB, btB = split_mps(top[ix], DuB), split_mps(bot[ix], DdB)
    # GB[a,c,A,B] = Renv[C,F,k,b] tpB[A,v,u,C] btB[B,e,d,F] QB[k,u,d,a] QB*[b,v,e,c]
    x = np.einsum("CFkb,AvuC->FkbAvu", Renv, tpB, optimize=True)
    x = np.einsum("FkbAvu,kuda->FbAvda", x, QB, optimize=True)
    x = np.einsum("FbAvda,BedF->bAvaBe", x, btB, optimize=True)
    GB = np.einsum("bAvaBe,bvec->acAB", x, QB.conj(), optimize=True)
    env = np.einsum("acTB,bdTB->abcd", GA, GB, optimize=True)
    return dict(env=env, RA=RA, RB=RB, QA=QA, QB=QB)


def spectator_left_v(Lenv, topj, botj, Eiy, Eiy1):
    """finite_peps_boundary_mps.jl:468-482."""
    Dl, Dr, Du, Dd = _dims4(Eiy)
    Dl1, Dr1, Du1, Dd1 = _dims4(Eiy1)
    e0 = split_dl(Eiy, Dl, Dr, Du, Dd)
    e1 = split_dl(Eiy1, Dl1, Dr1, Du1, Dd1)
    tp = split_mps(topj, Du1)
    bt = split_mps(botj, Dd)
    # Lenv[tL,bL,liyk,liyb,liy1k,liy1b]; the five-tensor sum taken pairwise in a
    # fixed order (numpy's greedy path can build a huge intermediate here)
    x = np.einsum("ABkbmn,BedF->AkbmnedF", Lenv, bt, optimize=True)
    x = np.einsum("AkbmnedF,bkxyvued->AmnFxyvu", x, e0, optimize=True)
    x = np.einsum("AmnFxyvu,nmzwsrvu->AFxyzwsr", x, e1, optimize=True)
    return np.einsum("AFxyzwsr,AsrC->CFyxwz", x, tp, optimize=True)


def spectator_right_v(Renv, topj, botj, Eiy, Eiy1):
    """finite_peps_boundary_mps.jl:485-499."""
    Dl, Dr, Du, Dd = _dims4(Eiy)
    Dl1, Dr1, Du1, Dd1 = _dims4(Eiy1)
    e0 = split_dl(Eiy, Dl, Dr, Du, Dd)
    e1 = split_dl(Eiy1, Dl1, Dr1, Du1, Dd1)
    tp = split_mps(topj, Du1)
    bt = split_mps(botj, Dd)
    x = np.einsum("CFyxwz,BedF->CyxwzBed", Renv, bt, optimize=True)
    x = np.einsum("CyxwzBed,bkxyvued->CwzBbkvu", x, e0, optimize=True)
    x = np.einsum("CwzBbkvu,nmzwsrvu->CBbknmsr", x, e1, optimize=True)
    return np.einsum("CBbknmsr,AsrC->ABkbmn", x, tp, optimize=True)


def bond_environment_v(gp, ix, iy, chi=64, top=None, bot=None, renorm=False):
    """finite_peps_boundary_mps.jl:507-547 (same contraction, optimal pairwise
    order; renorm: boundaries as boundary_top)."""
    nxv = gp.nx
    if top is None:
        top = boundary_top(gp, iy + 1, chi, renorm)
    if bot is None:
        bot = boundary_bottom(gp, iy, chi, renorm)
    row0, row1 = row_dl(gp, iy), row_dl(gp, iy + 1)
    A = sqrt_absorbed_site(gp, ix, iy)
    B = sqrt_absorbed_site(gp, ix, iy + 1)
    QA, RA = qr_reduce_v_down(A)
    QB, RB = qr_reduce_v_up(B)
    DdA, DuB = A.shape[4], B.shape[3]
    Lenv = np.ones((1,) * 6, np.complex128)
    for j in range(1, ix):
        Lenv = spectator_left_v(Lenv, top[j - 1], bot[j - 1], row0[j - 1], row1[j - 1])
    Renv = np.ones((1,) * 6, np.complex128)
    for j in range(nxv, ix, -1):
        Renv = spectator_right_v(Renv, top[j - 1], bot[j - 1], row0[j - 1], row1[j - 1])
    tp = split_mps(top[ix - 1], DuB)
    bt = split_mps(bot[ix - 1], DdA)
    # Lenv[tL,bL,lk,lb,l1k,l1b] tp[tL,ut2,ut1,tR] Renv[tR,bR,rk,rb,r1k,r1b]
    # bt[bL,d2,d1,bR] QA[lk,rk,d1,aA] QA*[lb,rb,d2,aAp] QB[l1k,r1k,ut1,bB] QB*[l1b,r1b,ut2,bBp]
    env = np.einsum("ABkbmn,AvuC,CFxyzw,BedF,kxda,byec,mzuf,nwvg->afcg",
                    Lenv, tp, Renv, bt, QA, QA.conj(), QB, QB.conj(), optimize=True)
    return dict(env=env, RA=RA, RB=RB, QA=QA, QB=QB)


def dl_with_op(T, Op):
    """Double layer with Op on the ket's physical leg, layout of dl_traced:
    E = sum_{p,q} Op[p,q] T[q,...] conj(T[p,...]) (bra fast)."""
    TO = np.einsum("pq,qabcd->pabcd", Op, T, optimize=True)
    dp, Dl, Dr, Du, Dd = T.shape
    E8 = np.einsum("pabcd,pefgh->eafbgchd", TO, T.conj(), optimize=True)
    return jr(E8, Dl * Dl, Dr * Dr, Du * Du, Dd * Dd)


def plaquette_expect(gp, ix, iy, chi=64):
    """<W> = <psi|Ud_b U_l Ud_r U_t|psi> / <psi|psi> on the plaquette with lower-left
    site (ix, iy), from boundary MPS: the two-row strip (rows iy, iy+1) between
    boundary_top(iy+1) and boundary_bottom(iy), W's site factors on the ket
    layer of A, B, C, closed column by column with spectator_left_v.  No
    reference function: the boundary-MPS counterpart of plaquette_energy
    (finite_peps_ground_state.jl:589-684), which uses simple-update weights."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    assert 1 <= ix < nx and 1 <= iy < ny
    top, bot = boundary_top(gp, iy + 1, chi), boundary_bottom(gp, iy, chi)
    row0, row1 = row_dl(gp, iy), row_dl(gp, iy + 1)
    WA, WB, WC = wilson_site_factors(nx, ny, gd, ix, iy)
    w0, w1 = list(row0), list(row1)
    w0[ix - 1] = dl_with_op(sqrt_absorbed_site(gp, ix, iy), WA)
    w0[ix] = dl_with_op(sqrt_absorbed_site(gp, ix + 1, iy), WB)
    w1[ix - 1] = dl_with_op(sqrt_absorbed_site(gp, ix, iy + 1), WC)

    def close(r0, r1):
        L = np.ones((1,) * 6, np.complex128)
        for j in range(1, nx + 1):
            L = spectator_left_v(L, top[j - 1], bot[j - 1], r0[j - 1], r1[j - 1])
        return L.reshape(-1)[0]

    return close(w0, w1) / close(row0, row1)


# ════════════════════════════════════════════════════════════════════════════
#  Full update v4 (finite_peps_fullupdate_v4.jl)
# ════════════════════════════════════════════════════════════════════════════

def solve_herm(N, S, reg):
    """_solve_herm, finite_peps_fullupdate_v4.jl:37-45."""
    Nh = 0.5 * (N + N.conj().T)
    w, V = np.linalg.eigh(Nh)
    wmax = max(np.max(w), 0.0) if len(w) else 0.0
    floor = reg * max(wmax, 1e-30)
    inv_w = np.where(w > floor, 1.0 / np.where(w > floor, w, 1.0), 0.0)
    return V @ (inv_w[:, None] * (V.conj().T @ S))


def apply_gate(th, G, dpL, dpR):
    """_apply_gate, finite_peps_fullupdate_v4.jl:48-55."""
    nA, _, _, nB = th.shape
    M = jr(pd(th, 3, 2, 1, 4), dpR * dpL, nA * nB)
    return pd(jr(G @ M, dpR, dpL, nA, nB), 3, 2, 1, 4)


def col_charge_from_blocks(M, row_charge):
    """_col_charge_from_blocks, finite_peps_fullupdate_v4.jl:59-71 (first strict max)."""
    a2 = np.abs(M) ** 2
    rbest = np.argmax(a2, axis=0)  # numpy argmax returns the first maximum
    return np.asarray(row_charge)[rbest]


def als_cost(th, env, X, Y):
    """finite_peps_fullupdate_v4.jl:83-92."""
    thperm = pd(th, 1, 4, 2, 3).conj()
    Eth = np.einsum("abcd,apqb->cdpq", env, th, optimize=True)
    thth = np.real(np.sum(Eth * thperm))
    XY = np.einsum("apk,bqk->apqb", X, Y, optimize=True)
    EXY = np.einsum("abcd,apqb->cdpq", env, XY, optimize=True)
    XYth = np.sum(EXY * thperm)
    XYXY = np.real(np.sum(EXY * pd(XY, 1, 4, 2, 3).conj()))
    return thth - 2 * np.real(XYth) + XYXY


def als_truncate(th, env, D_max, eRofpL, n_als=30, reg=V4_ALS_REG, return_info=False, init=None,
                 positive_cost=False, als_tol=0.0):
    """finite_peps_fullupdate_v4.jl:94-154.

    init=(X0, Y0) replaces the charge-SVD initial bond tensors (the SVD still
    fixes the bond fluxes and the mask).  For complex tensors the sweeps are
    not covariant under the phase freedom of the singular vectors (NX pairs
    Y with conj(Y) in the reference's index order), so comparing two
    implementations of the sweeps needs a shared init.

    positive_cost (no reference counterpart) also requires cost1 >= 0 for
    the refinement to be kept; the sweeps always run here.

    als_tol > 0 (no reference counterpart; the reference runs exactly n_als
    sweeps) stops after the first sweep whose cost differs from the previous
    one (the initial bond's before the first sweep) by at most
    als_tol * |previous|; the accept rule is unchanged."""
    nA, dpL, dpR, nB = th.shape
    thm = jr(pd(th, 1, 2, 4, 3), nA * dpL, nB * dpR)
    row_charge = np.array([eRofpL[r // nA] for r in range(nA * dpL)])
    col_charge = col_charge_from_blocks(thm, row_charge)
    U, S, Vt, bflux = charge_svd(thm, row_charge, col_charge, D_max)
    Dp = len(S)
    sq = np.sqrt(S)
    X = jr(U * sq[None, :], nA, dpL, Dp).copy()
    Y = pd(jr(sq[:, None] * Vt, Dp, nB, dpR), 2, 3, 1)
    Xmask = np.broadcast_to((np.asarray(eRofpL)[:, None] == bflux[None, :])[None, :, :], X.shape)
    if init is not None:
        assert init[0].shape == X.shape and init[1].shape == Y.shape
        X, Y = np.array(init[0], dtype=complex), np.array(init[1], dtype=complex)
    X0, Y0 = X.copy(), Y.copy()
    cost0 = als_cost(th, env, X0, Y0)
    prev = cost0
    sweeps = 0
    for _ in range(n_als):
        sweeps += 1
        NX = np.einsum("bqk,abcd,dql->akcl", Y, env, Y.conj(), optimize=True)
        SX = np.einsum("ebcd,epqb,dqk->cpk", env, th, Y.conj(), optimize=True)
        Nm = jr(NX, nA * Dp, nA * Dp)
        for pL in range(dpL):
            s = jr(SX[:, pL, :], nA * Dp, 1)
            X[:, pL, :] = jr(solve_herm(Nm, s, reg), nA, Dp)
        X[~Xmask] = 0
        NY = np.einsum("apk,abcd,cpl->bkdl", X, env, X.conj(), optimize=True)
        SY = np.einsum("aecb,apqe,cpk->bqk", env, th, X.conj(), optimize=True)
        Nm = jr(NY, nB * Dp, nB * Dp)
        for pR in range(dpR):
            s = jr(SY[:, pR, :], nB * Dp, 1)
            Y[:, pR, :] = jr(solve_herm(Nm, s, reg), nB, Dp)
        if als_tol > 0:
            cur = als_cost(th, env, X, Y)
            conv = abs(prev - cur) <= als_tol * max(abs(prev), 1e-300)
            prev = cur
            if conv:
                break
    cost1 = als_cost(th, env, X, Y)
    accepted = bool(np.isfinite(cost1) and cost1 < cost0 and (not positive_cost or cost1 >= 0))
    if accepted:
        out = (X, Y, bflux)
    else:
        out = (X0, Y0, bflux)
    if return_info:
        return out + (dict(cost0=cost0, cost1=cost1, accepted=accepted, sweeps=sweeps),)
    return out


def _isq(gp_lam):
    """1/(sqrt(lam)+FU_REG) on an existing leg, exactly 1 on a missing one (v4 :249-255)."""
    if gp_lam is None:
        return np.ones(1)
    return 1.0 / (np.sqrt(gp_lam) + FU_REG)


def recanonicalise_h(gp, ix, iy, A, B, D_max):
    """_recanonicalise_h!, finite_peps_fullupdate_v4.jl:215-267."""
    gd = gp.gd
    dpL, DlA, _, DuA, DdA = A.shape
    dpR, _, DrB, DuB, DdB = B.shape
    _, dgRL, dgUL = site_dims(ix, iy, gp.nx, gp.ny, gd)
    Th = np.einsum("plkud,qkrvw->pludqrvw", A, B, optimize=True)
    Tm = jr(Th, dpL * DlA * DuA * DdA, dpR * DrB * DuB * DdB)
    row_charge = np.array([decode_phys(r0 % dpL + 1, dgRL, dgUL, gd)[1] for r0 in range(Tm.shape[0])])
    col_charge = col_charge_from_blocks(Tm, row_charge)
    U, S, Vt, newq = charge_svd(Tm, row_charge, col_charge, D_max)
    Dk = len(S)
    lam = S / max(np.linalg.norm(S), 1e-15)
    Lnew = pd(jr(U, dpL, DlA, DuA, DdA, Dk), 1, 2, 5, 3, 4)
    Rnew = pd(jr(Vt, Dk, dpR, DrB, DuB, DdB), 2, 1, 3, 4, 5)
    nx, ny = gp.nx, gp.ny
    il = _isq(gp.LH(ix - 1, iy) if ix > 1 else None)
    iuL = _isq(gp.LV(ix, iy) if iy < ny else None)
    idL = _isq(gp.LV(ix, iy - 1) if iy > 1 else None)
    ir = _isq(gp.LH(ix + 1, iy) if ix + 1 < nx else None)
    iuR = _isq(gp.LV(ix + 1, iy) if iy < ny else None)
    idR = _isq(gp.LV(ix + 1, iy - 1) if iy > 1 else None)
    Lnew = Lnew * il[None, :, None, None, None] * iuL[None, None, None, :, None] * idL[None, None, None, None, :]
    Rnew = Rnew * ir[None, None, :, None, None] * iuR[None, None, None, :, None] * idR[None, None, None, None, :]
    gp.setT(ix, iy, Lnew); gp.setT(ix + 1, iy, Rnew)
    gp.lh[ix - 1][iy - 1] = lam
    gp.qh[ix - 1][iy - 1] = newq


def recanonicalise_v(gp, ix, iy, Dt, Ut, D_max):
    """_recanonicalise_v!, finite_peps_fullupdate_v4.jl:318-367."""
    gd = gp.gd
    dpD, DlD, DrD, _, DdD = Dt.shape
    dpU, DlU, DrU, DuU, _ = Ut.shape
    _, dgRD, dgUD = site_dims(ix, iy, gp.nx, gp.ny, gd)
    Th = np.einsum("plrkd,qxyzk->plrdqxyz", Dt, Ut, optimize=True)
    Tm = jr(Th, dpD * DlD * DrD * DdD, dpU * DlU * DrU * DuU)
    row_charge = np.array([decode_phys(r0 % dpD + 1, dgRD, dgUD, gd)[2] for r0 in range(Tm.shape[0])])
    col_charge = col_charge_from_blocks(Tm, row_charge)
    U, S, Vt, newq = charge_svd(Tm, row_charge, col_charge, D_max)
    Dk = len(S)
    lam = S / max(np.linalg.norm(S), 1e-15)
    Dnew = pd(jr(U, dpD, DlD, DrD, DdD, Dk), 1, 2, 3, 5, 4)
    Unew = pd(jr(Vt, Dk, dpU, DlU, DrU, DuU), 2, 3, 4, 5, 1)
    nx, ny = gp.nx, gp.ny
    ilD = _isq(gp.LH(ix - 1, iy) if ix > 1 else None)
    irD = _isq(gp.LH(ix, iy) if ix < nx else None)
    idD = _isq(gp.LV(ix, iy - 1) if iy > 1 else None)
    ilU = _isq(gp.LH(ix - 1, iy + 1) if ix > 1 else None)
    irU = _isq(gp.LH(ix, iy + 1) if ix < nx else None)
    iuU = _isq(gp.LV(ix, iy + 1) if iy + 1 < ny else None)
    Dnew = Dnew * ilD[None, :, None, None, None] * irD[None, None, :, None, None] * idD[None, None, None, None, :]
    Unew = Unew * ilU[None, :, None, None, None] * irU[None, None, :, None, None] * iuU[None, None, None, :, None]
    gp.setT(ix, iy, Dnew); gp.setT(ix, iy + 1, Unew)
    gp.lv[ix - 1][iy - 1] = lam
    gp.qv[ix - 1][iy - 1] = newq


def full_update_bond_h_v4(gp, ix, iy, G, D_max, chi=64, n_als=30, top=None, bot=None, positive_cost=False,
                          infos=None, renorm=False, als_tol=0.0):
    """finite_peps_fullupdate_v4.jl:165-193 (infos: a list that receives the
    ALS info dict of the bond; renorm: PEPS_STEP_RENORM boundaries)."""
    gd = gp.gd
    apply_mask(gp, ix, iy); apply_mask(gp, ix + 1, iy)
    E = bond_environment_h(gp, ix, iy, chi, top=top, bot=bot, renorm=renorm)
    dpL, dpR = E["RA"].shape[1], E["RB"].shape[1]
    _, dgRL, dgUL = site_dims(ix, iy, gp.nx, gp.ny, gd)
    th = np.einsum("apk,bqk->apqb", E["RA"], E["RB"], optimize=True)
    th = apply_gate(th, G, dpL, dpR)
    eR = [decode_phys(p, dgRL, dgUL, gd)[1] for p in range(1, dpL + 1)]
    X, Y, bflux, info = als_truncate(th, E["env"], D_max, eR, n_als=n_als, return_info=True,
                                     positive_cost=positive_cost, als_tol=als_tol)
    if infos is not None:
        infos.append(info)
    Afull = np.einsum("luda,apr->plrud", E["QA"], X, optimize=True)
    Bfull = np.einsum("rudb,bpl->plrud", E["QB"], Y, optimize=True)
    recanonicalise_h(gp, ix, iy, Afull, Bfull, D_max)
    apply_mask(gp, ix, iy); apply_mask(gp, ix + 1, iy)


def full_update_bond_v_v4(gp, ix, iy, G, D_max, chi=64, n_als=30, top=None, bot=None, positive_cost=False,
                          infos=None, renorm=False, als_tol=0.0):
    """finite_peps_fullupdate_v4.jl:274-299 (infos, renorm as full_update_bond_h_v4)."""
    gd = gp.gd
    apply_mask(gp, ix, iy); apply_mask(gp, ix, iy + 1)
    E = bond_environment_v(gp, ix, iy, chi, top=top, bot=bot, renorm=renorm)
    dpD, dpU = E["RA"].shape[1], E["RB"].shape[1]
    _, dgRD, dgUD = site_dims(ix, iy, gp.nx, gp.ny, gd)
    th = np.einsum("apk,bqk->apqb", E["RA"], E["RB"], optimize=True)
    th = apply_gate(th, G, dpD, dpU)
    eU = [decode_phys(p, dgRD, dgUD, gd)[2] for p in range(1, dpD + 1)]
    X, Y, bflux, info = als_truncate(th, E["env"], D_max, eU, n_als=n_als, return_info=True,
                                     positive_cost=positive_cost, als_tol=als_tol)
    if infos is not None:
        infos.append(info)
    Afull = np.einsum("lrda,apu->plrud", E["QA"], X, optimize=True)
    Bfull = np.einsum("lrub,bpd->plrud", E["QB"], Y, optimize=True)
    recanonicalise_v(gp, ix, iy, Afull, Bfull, D_max)
    apply_mask(gp, ix, iy); apply_mask(gp, ix, iy + 1)


def trotter_step_v4(gp, g, t_hop, m, tau, D_max, chi=64, n_als=30, positive_cost=False, infos=None,
                    renorm=False, als_tol=0.0):
    """trotter_step_v4!, finite_peps_fullupdate_v4.jl:415-432 (positive_cost,
    infos, renorm and als_tol are passed to every bond update)."""
    nxv, nyv, gd = gp.nx, gp.ny, gp.gd
    for iy in range(1, nyv + 1):
        for ix in range(1, nxv):
            G = gate(H_merged_h_site(ix, iy, nxv, nyv, gd, g, t_hop, m), tau / 2)
            full_update_bond_h_v4(gp, ix, iy, G, D_max, chi, n_als, positive_cost=positive_cost, infos=infos,
                                  renorm=renorm, als_tol=als_tol)
    for iy in range(1, nyv):
        for ix in range(1, nxv + 1):
            G = gate(H_merged_v_site(ix, iy, nxv, nyv, gd, g, t_hop, m), tau)
            full_update_bond_v_v4(gp, ix, iy, G, D_max, chi, n_als, positive_cost=positive_cost, infos=infos,
                                  renorm=renorm, als_tol=als_tol)
    for iy in range(1, nyv + 1):
        for ix in range(nxv - 1, 0, -1):
            G = gate(H_merged_h_site(ix, iy, nxv, nyv, gd, g, t_hop, m), tau / 2)
            full_update_bond_h_v4(gp, ix, iy, G, D_max, chi, n_als, positive_cost=positive_cost, infos=infos,
                                  renorm=renorm, als_tol=als_tol)


def tau_stages(n_ite):
    """The tau annealing of ite_ground_state_v4/_v3/_v2
    (finite_peps_fullupdate_v4.jl:447-449, finite_peps_gauge_invariant.jl:
    556-559, finite_peps_full_update.jl:595-608)."""
    s1, s2, s3, s4, s5 = n_ite // 8, n_ite // 8, n_ite // 4, n_ite // 6, n_ite // 6
    return [(0.1, s1), (0.05, s2), (0.02, s3), (0.01, s4), (0.005, s5), (0.001, n_ite - (s1 + s2 + s3 + s4 + s5))]


def ite_ground_state_v4(gp, g, t_hop, m, D_max, n_ite=300, chi=64, n_als=30, **opts):
    """ite_ground_state_v4 (finite_peps_fullupdate_v4.jl:440-467) from the
    initial state gp (init_gauge_peps), stepped in place."""
    for tau, ns in tau_stages(n_ite):
        for _ in range(ns):
            trotter_step_v4(gp, g, t_hop, m, tau, D_max, chi, n_als, **opts)
    return gp


def ite_ground_state_v3(gp, g, t_hop, m, D_max, n_ite=600):
    """ite_ground_state_v3 (finite_peps_gauge_invariant.jl:548-577) from gp."""
    for tau, ns in tau_stages(n_ite):
        for _ in range(ns):
            trotter_step_gauge(gp, g, t_hop, m, tau, D_max)
    return gp


def ite_ground_state_v2(gp, g, t_hop, m, D_max, n_ite=600):
    """ite_ground_state_v2 (finite_peps_full_update.jl:582-629, use_gauss=false) from gp."""
    for tau, ns in tau_stages(n_ite):
        for _ in range(ns):
            trotter_step_full(gp, g, t_hop, m, tau, D_max)
    return gp


def bond_expect(E, h, dpL, dpR):
    """_bond_expect, finite_peps_fullupdate_v4.jl:374-382."""
    th0 = np.einsum("apk,bqk->apqb", E["RA"], E["RB"], optimize=True)
    hth = apply_gate(th0, h, dpL, dpR)
    c = pd(th0, 1, 4, 2, 3).conj()
    Tn = np.einsum("abcd,apqb->cdpq", E["env"], th0, optimize=True)
    Th = np.einsum("abcd,apqb->cdpq", E["env"], hth, optimize=True)
    return np.real(np.sum(Th * c)), np.real(np.sum(Tn * c))


def compute_energy_v4(gp, g, t_hop, m, chi=64, renorm=False):
    """finite_peps_fullupdate_v4.jl:390-409 (total energy; renorm:
    PEPS_STEP_RENORM boundaries)."""
    nxv, nyv, gd = gp.nx, gp.ny, gp.gd
    E_tot = 0.0
    for iy in range(1, nyv + 1):
        top, bot = boundary_top(gp, iy, chi, renorm), boundary_bottom(gp, iy, chi, renorm)
        for ix in range(1, nxv):
            Eb = bond_environment_h(gp, ix, iy, chi, top=top, bot=bot)
            h = H_merged_h_site(ix, iy, nxv, nyv, gd, g, t_hop, m)
            hv, nr = bond_expect(Eb, h, Eb["RA"].shape[1], Eb["RB"].shape[1])
            E_tot += hv / nr
    for iy in range(1, nyv):
        top, bot = boundary_top(gp, iy + 1, chi, renorm), boundary_bottom(gp, iy, chi, renorm)
        for ix in range(1, nxv + 1):
            Eb = bond_environment_v(gp, ix, iy, chi, top=top, bot=bot)
            h = H_merged_v_site(ix, iy, nxv, nyv, gd, g, t_hop, m)
            hv, nr = bond_expect(Eb, h, Eb["RA"].shape[1], Eb["RB"].shape[1])
            E_tot += hv / nr
    return E_tot


def mean_nf(gp):
    """Simple-update-weight <n_f> averaged over sites (finite_peps_quench.jl:134-146)."""
    vals = []
    for iy in range(1, gp.ny + 1):
        for ix in range(1, gp.nx + 1):
            nf, _, _ = phys_labels(ix, iy, gp.nx, gp.ny, gp.gd)
            T = gp.T(ix, iy)
            w = T * _wt_l(gp, ix, iy)[None, :, None, None, None] * _wt_r(gp, ix, iy)[None, None, :, None, None] \
                * _wt_u(gp, ix, iy)[None, None, None, :, None] * _wt_d(gp, ix, iy)[None, None, None, None, :]
            rho = np.sum(np.abs(w) ** 2, axis=(1, 2, 3, 4))
            vals.append(np.sum(rho * nf) / np.sum(rho))
    return float(np.mean(vals))


def bond_entropy(lam):
    """finite_peps_quench.jl:121-127."""
    p = np.asarray(lam) ** 2
    Z = p.sum()
    if Z < 1e-30:
        return 0.0
    p = p / Z
    return float(-np.sum([x * np.log(x) if x > 1e-30 else 0.0 for x in p]))


# ════════════════════════════════════════════════════════════════════════════
#  Dense simple update with plaquettes (finite_peps_ground_state.jl) and the
#  plaquette gate split (u1_lgt_hamiltonian.jl).  A FinitePEPS is held in the
#  GaugePEPS container; its flux labels are not used (nor kept up to date).
# ════════════════════════════════════════════════════════════════════════════

SU_PEEL = 1e-15  # finite_peps_ground_state.jl:264-269, :364-369


def _bond_w(gp, orient, ix, iy, exists):
    """Bond weight vector, or ones(1) for a missing bond."""
    if not exists:
        return np.ones(1)
    return np.asarray(gp.LH(ix, iy) if orient == "h" else gp.LV(ix, iy), dtype=np.float64)


def _scale5(T, wl, wr, wu, wd):
    return T * wl[None, :, None, None, None] * wr[None, None, :, None, None] \
        * wu[None, None, None, :, None] * wd[None, None, None, None, :]


def _su_split(Thn, nrow, ncol, D_trunc, cut=SVD_CUT, norm_floor=0.0):
    """SVD split of the gated two-site tensor, finite_peps_ground_state.jl:241-246
    (finite_peps_full_update.jl:419-424 with cut 1e-10 and norm floor 1e-15)."""
    U, S, Vt = np.linalg.svd(jr(Thn, nrow, ncol), full_matrices=False)
    Dk = max(1, min(D_trunc, int(np.sum(S > cut)), len(S)))
    St = S[:Dk]
    return U[:, :Dk], St, Vt[:Dk], St / max(np.linalg.norm(St), norm_floor), np.sqrt(St)


def update_bond_h(gp, ix, iy, G, D_trunc, cut=SVD_CUT, peel=SU_PEEL, norm_floor=0.0):
    """update_bond_h!, finite_peps_ground_state.jl:175-273: absorb sqrt(lambda)
    on every leg, Theta = L.R, gate on (p_R fast, p_L slow), SVD, keep
    <= D_trunc values above 1e-12, L = U sqrt(S), R = sqrt(S) V^H,
    lambda = S/||S||, divide the transverse weights out with +1e-15.
    (cut, peel, norm_floor) = (1e-10, 1e-10, 1e-15) is full_update_bond_h!."""
    nx, ny = gp.nx, gp.ny
    TL, TR = gp.T(ix, iy), gp.T(ix + 1, iy)
    dpL, dpR = TL.shape[0], TR.shape[0]
    lam = np.asarray(gp.LH(ix, iy), dtype=np.float64)
    bd = len(lam)
    assert TL.shape[2] == bd and TR.shape[1] == bd
    sq_b = np.sqrt(lam)
    sq_l = np.sqrt(_bond_w(gp, "h", ix - 1, iy, ix > 1))
    sq_r = np.sqrt(_bond_w(gp, "h", ix + 1, iy, ix + 1 < nx))
    sq_uL = np.sqrt(_bond_w(gp, "v", ix, iy, iy < ny))
    sq_dL = np.sqrt(_bond_w(gp, "v", ix, iy - 1, iy > 1))
    sq_uR = np.sqrt(_bond_w(gp, "v", ix + 1, iy, iy < ny))
    sq_dR = np.sqrt(_bond_w(gp, "v", ix + 1, iy - 1, iy > 1))
    Lw = _scale5(TL, sq_l, sq_b, sq_uL, sq_dL)
    Rw = _scale5(TR, sq_b, sq_r, sq_uR, sq_dR)
    _, Dl, _, DuL, DdL = Lw.shape
    _, _, Dr, DuR, DdR = Rw.shape
    Lm = jr(pd(Lw, 1, 2, 4, 5, 3), dpL * Dl * DuL * DdL, bd)
    Rm = jr(pd(Rw, 2, 1, 3, 4, 5), bd, dpR * Dr * DuR * DdR)
    Th = jr(Lm @ Rm, dpL, Dl, DuL, DdL, dpR, Dr, DuR, DdR)
    Thpp = jr(pd(Th, 5, 1, 2, 3, 4, 6, 7, 8), dpR * dpL, Dl * DuL * DdL * Dr * DuR * DdR)
    Thn = pd(jr(G @ Thpp, dpR, dpL, Dl, DuL, DdL, Dr, DuR, DdR), 2, 3, 4, 5, 1, 6, 7, 8)
    U, St, Vt, lam_new, sqS = _su_split(Thn, dpL * Dl * DuL * DdL, dpR * Dr * DuR * DdR, D_trunc, cut, norm_floor)
    Dk = len(St)
    Lnew = pd(jr(U * sqS[None, :], dpL, Dl, DuL, DdL, Dk), 1, 2, 5, 3, 4)
    Rnew = pd(jr(sqS[:, None] * Vt, Dk, dpR, Dr, DuR, DdR), 2, 1, 3, 4, 5)
    inv = lambda x: 1.0 / (x + peel)
    gp.setT(ix, iy, _scale5(Lnew, inv(sq_l), np.ones(Dk), inv(sq_uL), inv(sq_dL)))
    gp.setT(ix + 1, iy, _scale5(Rnew, np.ones(Dk), inv(sq_r), inv(sq_uR), inv(sq_dR)))
    gp.lh[ix - 1][iy - 1] = lam_new


def update_bond_v(gp, ix, iy, G, D_trunc, cut=SVD_CUT, peel=SU_PEEL, norm_floor=0.0):
    """update_bond_v!, finite_peps_ground_state.jl:280-382 (lower D=(ix,iy),
    upper U=(ix,iy+1); gate on (p_U fast, p_D slow)).  (cut, peel,
    norm_floor) = (1e-10, 1e-10, 1e-15) is full_update_bond_v!."""
    nx, ny = gp.nx, gp.ny
    TD, TU = gp.T(ix, iy), gp.T(ix, iy + 1)
    dpD, dpU = TD.shape[0], TU.shape[0]
    lam = np.asarray(gp.LV(ix, iy), dtype=np.float64)
    bd = len(lam)
    assert TD.shape[3] == bd and TU.shape[4] == bd
    sq_b = np.sqrt(lam)
    sq_lD = np.sqrt(_bond_w(gp, "h", ix - 1, iy, ix > 1))
    sq_rD = np.sqrt(_bond_w(gp, "h", ix, iy, ix < nx))
    sq_dD = np.sqrt(_bond_w(gp, "v", ix, iy - 1, iy > 1))
    sq_lU = np.sqrt(_bond_w(gp, "h", ix - 1, iy + 1, ix > 1))
    sq_rU = np.sqrt(_bond_w(gp, "h", ix, iy + 1, ix < nx))
    sq_uU = np.sqrt(_bond_w(gp, "v", ix, iy + 1, iy + 1 < ny))
    Dw = _scale5(TD, sq_lD, sq_rD, sq_b, sq_dD)
    Uw = _scale5(TU, sq_lU, sq_rU, sq_uU, sq_b)
    _, DlD, DrD, _, DdD = Dw.shape
    _, DlU, DrU, DuU, _ = Uw.shape
    Dm = jr(pd(Dw, 1, 2, 3, 5, 4), dpD * DlD * DrD * DdD, bd)
    Um = jr(pd(Uw, 5, 1, 2, 3, 4), bd, dpU * DlU * DrU * DuU)
    Th = jr(Dm @ Um, dpD, DlD, DrD, DdD, dpU, DlU, DrU, DuU)
    Thpp = jr(pd(Th, 5, 1, 2, 3, 4, 6, 7, 8), dpU * dpD, DlD * DrD * DdD * DlU * DrU * DuU)
    Thn = pd(jr(G @ Thpp, dpU, dpD, DlD, DrD, DdD, DlU, DrU, DuU), 2, 3, 4, 5, 1, 6, 7, 8)
    U, St, Vt, lam_new, sqS = _su_split(Thn, dpD * DlD * DrD * DdD, dpU * DlU * DrU * DuU, D_trunc, cut, norm_floor)
    Dk = len(St)
    Dnew = pd(jr(U * sqS[None, :], dpD, DlD, DrD, DdD, Dk), 1, 2, 3, 5, 4)
    Unew = pd(jr(sqS[:, None] * Vt, Dk, dpU, DlU, DrU, DuU), 2, 3, 4, 5, 1)
    inv = lambda x: 1.0 / (x + peel)
    gp.setT(ix, iy, _scale5(Dnew, inv(sq_lD), inv(sq_rD), np.ones(Dk), inv(sq_dD)))
    gp.setT(ix, iy + 1, _scale5(Unew, inv(sq_lU), inv(sq_rU), inv(sq_uU), np.ones(Dk)))
    gp.lv[ix - 1][iy - 1] = lam_new


def H_plaquette_gauge(gd, g):
    """u1_lgt_hamiltonian.jl:283-289: -(1/2g^2)(W + W^H), W = Ud (x) U (x) Ud (x) U
    (Julia kron: the last factor is the fastest index)."""
    Ug = op_U_gauge(gd)
    Ud = Ug.conj().T
    W = np.kron(np.kron(np.kron(Ud, Ug), Ud), Ug)
    H = -(1.0 / (2.0 * g ** 2)) * (W + W.conj().T)
    return 0.5 * (H + H.conj().T)


def decompose_plaquette_LShape(G, gd, D_aux_max=64, cutoff=1e-12):
    """u1_lgt_hamiltonian.jl:728-768: two SVDs split the gd^4 x gd^4 gate into
    GA (gd,gd,gd,gd,Da,Db), GB (gd,gd,Da), GC (gd,gd,Db)."""
    G8 = jr(G, *([gd] * 8))
    M1 = jr(pd(G8, 1, 2, 4, 5, 6, 8, 3, 7), gd ** 6, gd ** 2)
    U1, S1, V1 = np.linalg.svd(M1, full_matrices=False)
    Da = max(1, min(D_aux_max, int(np.sum(S1 > cutoff)), len(S1)))
    sq1 = np.sqrt(S1[:Da])
    GAC = U1[:, :Da] * sq1[None, :]
    GB = pd(jr(sq1[:, None] * V1[:Da], Da, gd, gd), 2, 3, 1)
    GAC5 = jr(GAC, gd, gd, gd, gd, gd, gd, Da)
    M2 = jr(pd(GAC5, 1, 2, 4, 5, 7, 3, 6), gd ** 4 * Da, gd ** 2)
    U2, S2, V2 = np.linalg.svd(M2, full_matrices=False)
    Db = max(1, min(D_aux_max, int(np.sum(S2 > cutoff)), len(S2)))
    sq2 = np.sqrt(S2[:Db])
    GA = jr(U2[:, :Db] * sq2[None, :], gd, gd, gd, gd, Da, Db)
    GC = pd(jr(sq2[:, None] * V2[:Db], Db, gd, gd), 2, 3, 1)
    return GA, GB, GC


def reconstruct_plaquette_LShape(GA, GB, GC):
    """u1_lgt_hamiltonian.jl:776-781."""
    gd = GA.shape[0]
    Gr = np.einsum("blBLxy,rRx,tTy->blrtBLRT", GA, GB, GC, optimize=True)
    return jr(Gr, gd ** 4, gd ** 4)


def update_plaquette(gp, ix, iy, GA, GB, GC, D_trunc, recompress=True):
    """update_plaquette!, finite_peps_ground_state.jl:405-483: GA on A's
    (R, U) links, GB on B's U link, GC on C's R link; the A-B and A-C bonds
    grow by Da, Db (lambda repeated) and are recompressed by the dense
    update with identity gates (recompress=False stops before that, for the
    checks of the enlargement).  GA/GB/GC are contracted through their first
    (row) indices, so the operator applied to the links is G^T, as in the
    reference."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    assert ix < nx and iy < ny
    Da, Db = GB.shape[2], GC.shape[2]
    TA, TB, TC = gp.T(ix, iy), gp.T(ix + 1, iy), gp.T(ix, iy + 1)
    dpA, gRA, gUA = site_dims(ix, iy, nx, ny, gd)
    dpB, gRB, gUB = site_dims(ix + 1, iy, nx, ny, gd)
    dpC, gRC, gUC = site_dims(ix, iy + 1, nx, ny, gd)
    assert gRA == gd and gUA == gd and gUB == gd and gRC == gd
    _, DlA, DrA, DuA, DdA = TA.shape
    _, DlB, DrB, DuB, DdB = TB.shape
    _, DlC, DrC, DuC, DdC = TC.shape
    TA3 = pd(jr(TA, gd, gd, 2, DlA, DrA, DuA, DdA), 3, 2, 1, 4, 5, 6, 7)
    TAg = np.einsum("RUSVab,nRUlrud->nSVablrud", GA, TA3, optimize=True)
    TAp = pd(pd(TAg, 1, 2, 3, 6, 7, 4, 8, 5, 9), 3, 2, 1, 4, 5, 6, 7, 8, 9)
    TB3 = pd(jr(TB, gUB, gRB, 2, DlB, DrB, DuB, DdB), 3, 2, 1, 4, 5, 6, 7)
    TBg = np.einsum("UVa,nRUlrud->nRValrud", GB, TB3, optimize=True)
    TBp = pd(pd(TBg, 1, 2, 3, 5, 4, 6, 7, 8), 3, 2, 1, 4, 5, 6, 7, 8)
    TC3 = pd(jr(TC, gUC, gRC, 2, DlC, DrC, DuC, DdC), 3, 2, 1, 4, 5, 6, 7)
    TCg = np.einsum("RSb,nRUlrud->nSUblrud", GC, TC3, optimize=True)
    TCp = pd(pd(TCg, 1, 2, 3, 5, 6, 7, 8, 4), 3, 2, 1, 4, 5, 6, 7, 8)
    gp.setT(ix, iy, jr(TAp, dpA, DlA, DrA * Da, DuA * Db, DdA))
    gp.setT(ix + 1, iy, jr(TBp, dpB, DlB * Da, DrB, DuB, DdB))
    gp.setT(ix, iy + 1, jr(TCp, dpC, DlC, DrC, DuC, DdC * Db))
    gp.lh[ix - 1][iy - 1] = np.tile(np.asarray(gp.LH(ix, iy)), Da)
    gp.lv[ix - 1][iy - 1] = np.tile(np.asarray(gp.LV(ix, iy)), Db)
    if not recompress:
        return
    update_bond_h(gp, ix, iy, np.eye(dpA * dpB, dtype=np.complex128), D_trunc)
    update_bond_v(gp, ix, iy, np.eye(dpA * dpC, dtype=np.complex128), D_trunc)


def trotter_step_dense(gp, g, t_hop, m, tau, D_trunc, mu=0.0):
    """trotter_step!, finite_peps_ground_state.jl:498-541:
    h-half (L->R) -> v-half (B->T) -> plaquettes -> v-half (T->B) -> h-half (R->L);
    mu is the chemical potential of H_merged (:507,514,528,535)."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    GA, GB, GC = decompose_plaquette_LShape(gate(H_plaquette_gauge(gd, g), tau), gd, cutoff=SVD_CUT)
    gh = lambda ix, iy: gate(H_merged_h_site(ix, iy, nx, ny, gd, g, t_hop, m, mu), tau / 2)
    gv = lambda ix, iy: gate(H_merged_v_site(ix, iy, nx, ny, gd, g, t_hop, m, mu), tau / 2)
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            update_bond_h(gp, ix, iy, gh(ix, iy), D_trunc)
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            update_bond_v(gp, ix, iy, gv(ix, iy), D_trunc)
    for iy in range(1, ny):
        for ix in range(1, nx):
            update_plaquette(gp, ix, iy, GA, GB, GC, D_trunc)
    for iy in range(ny - 1, 0, -1):
        for ix in range(nx, 0, -1):
            update_bond_v(gp, ix, iy, gv(ix, iy), D_trunc)
    for iy in range(ny, 0, -1):
        for ix in range(nx - 1, 0, -1):
            update_bond_h(gp, ix, iy, gh(ix, iy), D_trunc)


def realtime_gate(H, dt):
    """exp(-i dt H), the real-time gate of realtime_trotter_step!
    (finite_peps_quench.jl:88-89: exp(-tau H) at tau = i dt), from the
    spectral decomposition of the real symmetric H (Pade on a complex
    324 x 324 matrix costs ~50 times more)."""
    w, V = np.linalg.eigh(np.asarray(H))
    return (V * np.exp(-1j * dt * w)) @ V.conj().T


def realtime_trotter_step(gp, g, t_hop, m, dt, D_trunc):
    """realtime_trotter_step!, finite_peps_quench.jl:81-109: the dense
    update_bond_h!/v! with exp(-i dt/2 H_h) (iy ascending, ix ascending),
    exp(-i dt H_v) (iy ascending, ix ascending) and exp(-i dt/2 H_h) (iy
    ascending, ix descending)."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    gh = lambda ix, iy: realtime_gate(H_merged_h_site(ix, iy, nx, ny, gd, g, t_hop, m), dt / 2)
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            update_bond_h(gp, ix, iy, gh(ix, iy), D_trunc)
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            update_bond_v(gp, ix, iy, realtime_gate(H_merged_v_site(ix, iy, nx, ny, gd, g, t_hop, m), dt), D_trunc)
    for iy in range(1, ny + 1):
        for ix in range(nx - 1, 0, -1):
            update_bond_h(gp, ix, iy, gh(ix, iy), D_trunc)


def apply_site_op(gp, ix, iy, O):
    """apply_site_op!, finite_peps_quench.jl:228-234: T <- O * reshape(T, dp, :)."""
    T = gp.T(ix, iy)
    gp.setT(ix, iy, jr(np.asarray(O) @ jr(T, T.shape[0], -1), *T.shape))


def full_update_bond_h_v2(gp, ix, iy, G, D_trunc):
    """full_update_bond_h!, finite_peps_full_update.jl:346-454 with
    use_gauss=false (its driver's setting): the environment-weighted branch
    was removed upstream, so this is the simple update with cut 1e-10,
    lambda = S/max(||S||, 1e-15) and transverse weights peeled with +1e-10."""
    update_bond_h(gp, ix, iy, G, D_trunc, cut=FU_REG, peel=FU_REG, norm_floor=1e-15)


def full_update_bond_v_v2(gp, ix, iy, G, D_trunc):
    """full_update_bond_v!, finite_peps_full_update.jl:462-543."""
    update_bond_v(gp, ix, iy, G, D_trunc, cut=FU_REG, peel=FU_REG, norm_floor=1e-15)


def trotter_step_full(gp, g, t_hop, m, tau, D_trunc):
    """trotter_step_full!, finite_peps_full_update.jl:550-575: h with tau/2
    (iy ascending, ix ascending), v with tau, h with tau/2 (iy ascending, ix
    descending)."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    gh = lambda ix, iy: gate(H_merged_h_site(ix, iy, nx, ny, gd, g, t_hop, m), tau / 2)
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            full_update_bond_h_v2(gp, ix, iy, gh(ix, iy), D_trunc)
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            full_update_bond_v_v2(gp, ix, iy, gate(H_merged_v_site(ix, iy, nx, ny, gd, g, t_hop, m), tau), D_trunc)
    for iy in range(1, ny + 1):
        for ix in range(nx - 1, 0, -1):
            full_update_bond_h_v2(gp, ix, iy, gh(ix, iy), D_trunc)


def expect_site(gp, ix, iy, O):
    """expect_site, finite_peps_ground_state.jl:549-567: lambda (not its root) on
    every leg, <O> = tr(T^H O T) / tr(T^H T), 0 when the norm is <= 1e-15."""
    nx, ny = gp.nx, gp.ny
    Tw = _scale5(gp.T(ix, iy), _bond_w(gp, "h", ix - 1, iy, ix > 1), _bond_w(gp, "h", ix, iy, ix < nx),
                 _bond_w(gp, "v", ix, iy, iy < ny), _bond_w(gp, "v", ix, iy - 1, iy > 1))
    Tf = jr(Tw, Tw.shape[0], Tw.size // Tw.shape[0])
    num = np.trace(Tf.conj().T @ O @ Tf).real
    den = np.trace(Tf.conj().T @ Tf).real
    return num / den if den > 1e-15 else 0.0


def wilson_site_factors(nx, ny, gd, ix, iy):
    """Single-site factors of W = Ud_bottom U_left Ud_right U_top on the L-shape
    (finite_peps_ground_state.jl:632-641): A=(ix,iy) carries Ud on its R link and
    U on its U link, B=(ix+1,iy) Ud on its U link, C=(ix,iy+1) U on its R link."""
    _, gRB, _ = site_dims(ix + 1, iy, nx, ny, gd)
    _, _, gUC = site_dims(ix, iy + 1, nx, ny, gd)
    Ug = op_U_gauge(gd)
    Ud = Ug.conj().T
    WA = np.kron(np.kron(_I(2), Ud), Ug)
    WB = np.kron(_I(2), Ud) if gRB == 1 else np.kron(np.kron(_I(2), _I(gRB)), Ud)
    WC = np.kron(_I(2), Ug) if gUC == 1 else np.kron(np.kron(_I(2), Ug), _I(gUC))
    return WA, WB, WC


def plaquette_energy(gp, ix, iy, g):
    """plaquette_energy, finite_peps_ground_state.jl:589-684: -(1/g^2) Re<W> on
    the L-shape A=(ix,iy), B=(ix+1,iy), C=(ix,iy+1) with sqrt(lambda) on
    every leg and the external legs closed diagonally."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    assert ix < nx and iy < ny
    _, gRB, _ = site_dims(ix + 1, iy, nx, ny, gd)
    _, _, gUC = site_dims(ix, iy + 1, nx, ny, gd)
    sw = lambda o, jx, jy, e: np.sqrt(_bond_w(gp, o, jx, jy, e))
    sq_AB, sq_AC = np.sqrt(np.asarray(gp.LH(ix, iy))), np.sqrt(np.asarray(gp.LV(ix, iy)))
    TA = _scale5(gp.T(ix, iy), sw("h", ix - 1, iy, ix > 1), sq_AB, sq_AC, sw("v", ix, iy - 1, iy > 1))
    TB = _scale5(gp.T(ix + 1, iy), sq_AB, sw("h", ix + 1, iy, ix + 1 < nx), sw("v", ix + 1, iy, iy < ny),
                 sw("v", ix + 1, iy - 1, iy > 1))
    TC = _scale5(gp.T(ix, iy + 1), sw("h", ix - 1, iy + 1, ix > 1), sw("h", ix, iy + 1, ix < nx),
                 sw("v", ix, iy + 1, iy + 1 < ny), sq_AC)
    WA, WB, WC = wilson_site_factors(nx, ny, gd, ix, iy)
    app = lambda T, O: jr(O @ jr(T, T.shape[0], T.size // T.shape[0]), *T.shape)
    MA = lambda k, b: np.einsum("plrud,plsvd->rsuv", k, b.conj(), optimize=True)
    MB = lambda k, b: np.einsum("plrud,pmrud->lm", k, b.conj(), optimize=True)
    MC = lambda k, b: np.einsum("plrud,plrue->de", k, b.conj(), optimize=True)
    sc = lambda a, b, c: np.einsum("rsuv,rs,uv->", a, b, c, optimize=True)
    num = sc(MA(app(TA, WA), TA), MB(app(TB, WB), TB), MC(app(TC, WC), TC))
    den = sc(MA(TA, TA), MB(TB, TB), MC(TC, TC))
    return -(1.0 / g ** 2) * num.real / den.real if abs(den) > 1e-15 else 0.0


def _su_hop_expect(Lw, Rw, Lperm, Rperm, H):
    """The hopping term of total_energy (finite_peps_ground_state.jl:727-735,
    :778-788): Theta[p1,p2,a,b] = sum_chi L[p1,chi,a] conj(R[p2,chi,b]), flattened
    with p1 fastest, rho = Theta Theta^H, <H> = Re tr(H rho) / Re tr(rho)."""
    dpL, dpR = Lw.shape[0], Rw.shape[0]
    Lb = pd(Lw, *Lperm)
    Rb = pd(Rw, *Rperm)
    bd = Lb.shape[1]
    Lb = jr(Lb, dpL, bd, Lb.size // (dpL * bd))
    Rb = jr(Rb, dpR, bd, Rb.size // (dpR * bd))
    Th = np.einsum("pca,qcb->pqab", Lb, Rb.conj(), optimize=True)
    Tf = jr(Th, dpL * dpR, Th.size // (dpL * dpR))
    rho = Tf @ Tf.conj().T
    den = np.trace(rho).real
    return np.trace(H @ rho).real / den if den > 1e-15 else 0.0


def total_energy(gp, g, t_hop, m):
    """total_energy, finite_peps_ground_state.jl:690-805: on-site terms by
    expect_site, hopping terms by _su_hop_expect (the rho of each bond is
    flattened with the left/lower site fastest while H_hop has the right/upper
    site fastest; kept as written), plaquettes by plaquette_energy."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    E = 0.0
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            E += expect_site(gp, ix, iy, H_onsite_site(ix, iy, nx, ny, gd, g, m))
    sw = lambda o, jx, jy, e: np.sqrt(_bond_w(gp, o, jx, jy, e))
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            sb = np.sqrt(np.asarray(gp.LH(ix, iy)))
            Lw = _scale5(gp.T(ix, iy), sw("h", ix - 1, iy, ix > 1), sb, sw("v", ix, iy, iy < ny),
                         sw("v", ix, iy - 1, iy > 1))
            Rw = _scale5(gp.T(ix + 1, iy), sb, sw("h", ix + 1, iy, ix + 1 < nx), sw("v", ix + 1, iy, iy < ny),
                         sw("v", ix + 1, iy - 1, iy > 1))
            E += _su_hop_expect(Lw, Rw, (1, 3, 2, 4, 5), (1, 2, 3, 4, 5), H_hop_h_site(ix, iy, nx, ny, gd, t_hop))
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            sb = np.sqrt(np.asarray(gp.LV(ix, iy)))
            Dw = _scale5(gp.T(ix, iy), sw("h", ix - 1, iy, ix > 1), sw("h", ix, iy, ix < nx), sb,
                         sw("v", ix, iy - 1, iy > 1))
            Uw = _scale5(gp.T(ix, iy + 1), sw("h", ix - 1, iy + 1, ix > 1), sw("h", ix, iy + 1, ix < nx),
                         sw("v", ix, iy + 1, iy + 1 < ny), sb)
            E += _su_hop_expect(Dw, Uw, (1, 4, 2, 3, 5), (1, 5, 2, 3, 4), H_hop_v_site(ix, iy, nx, ny, gd, t_hop))
    for iy in range(1, ny):
        for ix in range(1, nx):
            E += plaquette_energy(gp, ix, iy, g)
    return E


def measure_nf_grid(gp):
    """measure_nf_grid, finite_peps_quench.jl:135-142: (nx, ny) grid of <n_f>."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    grid = np.zeros((nx, ny))
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            _, dR, dU = site_dims(ix, iy, nx, ny, gd)
            grid[ix - 1, iy - 1] = expect_site(gp, ix, iy, np.kron(np.kron(op_nf(), _I(dR)), _I(dU)))
    return grid


def measure_E2_grid(gp):
    """measure_E2_grid, finite_peps_quench.jl:149-163: <E_R^2 + E_U^2>, existing links only."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    grid = np.zeros((nx, ny))
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d, dR, dU = site_dims(ix, iy, nx, ny, gd)
            Op = np.zeros((d, d), dtype=np.complex128)
            if ix < nx:
                Op += np.kron(np.kron(_I(2), op_E2(gd)), _I(dU))
            if iy < ny:
                Op += np.kron(np.kron(_I(2), _I(dR)), op_E2(gd))
            grid[ix - 1, iy - 1] = expect_site(gp, ix, iy, Op)
    return grid


def measure_all_finite(gp, t=0.0):
    """measure_all_finite, finite_peps_quench.jl:166-194."""
    nx, ny = gp.nx, gp.ny
    nf, E2 = measure_nf_grid(gp), measure_E2_grid(gp)
    ev = [nf[ix - 1, iy - 1] for iy in range(1, ny + 1) for ix in range(1, nx + 1) if (ix + iy) % 2 == 0]
    od = [nf[ix - 1, iy - 1] for iy in range(1, ny + 1) for ix in range(1, nx + 1) if (ix + iy) % 2 == 1]
    Sh = [bond_entropy(gp.LH(ix, iy)) for iy in range(1, ny + 1) for ix in range(1, nx)]
    Sv = [bond_entropy(gp.LV(ix, iy)) for iy in range(1, ny) for ix in range(1, nx + 1)]
    Sh_m = float(np.mean(Sh)) if Sh else 0.0
    Sv_m = float(np.mean(Sv)) if Sv else 0.0
    return dict(t=t, nf_mean=float(np.mean(nf)), nf_even=float(np.mean(ev)) if ev else 0.0,
                nf_odd=float(np.mean(od)) if od else 0.0, E2_mean=float(np.mean(E2)), S_h_mean=Sh_m,
                S_v_mean=Sv_m, S_mean=(Sh_m + Sv_m) / 2, nf_grid=nf, E2_grid=E2)


# ════════════════════════════════════════════════════════════════════════════
#  The small MLP of the learned ALS warm start (no reference counterpart)
# ════════════════════════════════════════════════════════════════════════════

def mlp_unpack(p, n_in, width, n_out):
    """Flat parameters W1 (width x n_in), b1, W2 (width x width), b2, W3
    (n_out x width), b3, each column-major."""
    shapes = [(width, n_in), (width,), (width, width), (width,), (n_out, width), (n_out,)]
    out, off = [], 0
    for shp in shapes:
        n = int(np.prod(shp))
        out.append(np.reshape(p[off:off + n], shp, order="F"))
        off += n
    assert off == p.size
    return out


def mlp_forward(p, X, n_in, width, n_out):
    """Y = W3 relu(W2 relu(W1 X + b1) + b2) + b3, batch along columns."""
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, n_in, width, n_out)
    H1 = np.maximum(W1 @ X + b1[:, None], 0.0)
    H2 = np.maximum(W2 @ H1 + b2[:, None], 0.0)
    return W3 @ H2 + b3[:, None]


def mlp_loss_grads(p, X, T, n_in, width, n_out):
    """loss 1/(2B) ||Y - T||^2 and its gradient (flattened like p), by backprop."""
    W1, b1, W2, b2, W3, b3 = mlp_unpack(p, n_in, width, n_out)
    B = X.shape[1]
    H1 = np.maximum(W1 @ X + b1[:, None], 0.0)
    H2 = np.maximum(W2 @ H1 + b2[:, None], 0.0)
    Y = W3 @ H2 + b3[:, None]
    D3 = (Y - T) / B
    loss = 0.5 * np.sum((Y - T) ** 2) / B
    D2 = (W3.T @ D3) * (H2 > 0)
    D1 = (W2.T @ D2) * (H1 > 0)
    grads = [D1 @ X.T, D1.sum(1), D2 @ H1.T, D2.sum(1), D3 @ H2.T, D3.sum(1)]
    return loss, np.concatenate([np.reshape(g, -1, order="F") for g in grads])


# ════════════════════════════════════════════════════════════════════════════
#  Brute-force references (tiny lattices only): dense |psi>, <psi|psi>, <H>
# ════════════════════════════════════════════════════════════════════════════

def dense_psi(gp):
    """Contract the sqrt-absorbed network into a dense state psi[p(1,1), p(2,1), ...].

    Physical legs are ordered ix fastest, then iy (column-major over sites).
    """
    nx, ny = gp.nx, gp.ny
    # contract row by row: a row tensor R[p_1..p_nx, u_1..u_nx] (down legs closed)
    psi = None
    for iy in range(1, ny + 1):
        row = None
        for ix in range(1, nx + 1):
            A = sqrt_absorbed_site(gp, ix, iy)  # (p,l,r,u,d)
            if row is None:
                # (p, r, u, d) with l dim 1
                row = A[:, 0, :, :, :]
                row = row.reshape(A.shape[0], A.shape[2], A.shape[3], A.shape[4])
                shape_p, shape_u, shape_d = [A.shape[0]], [A.shape[3]], [A.shape[4]]
                row = np.transpose(row, (0, 2, 3, 1))  # (p, u, d, r)
                row = row.reshape(A.shape[0], A.shape[3], A.shape[4], A.shape[2])
            else:
                # row[P, U, Dd, r] * A[p, r, r2, u, d] -> [P,p, U,u, Dd,d, r2]
                P, U, Dd, r = row.shape
                t = np.einsum("PUDr,prsud->PpUuDds", row, A, optimize=True)
                shape_p.append(A.shape[0]); shape_u.append(A.shape[3]); shape_d.append(A.shape[4])
                row = t.reshape(P * A.shape[0], U * A.shape[3], Dd * A.shape[4], A.shape[2], order="F")
        assert row.shape[3] == 1
        row = row[:, :, :, 0]  # (P_row, U_row, D_row)
        if psi is None:
            assert row.shape[2] == 1
            psi = row[:, :, 0]  # (P_all, U)
        else:
            # psi[Pall, U] * row[P, U', Drow] with Drow == U
            psi = np.einsum("Au,Bvu->ABv", psi, row, optimize=True)
            psi = psi.reshape(psi.shape[0] * psi.shape[1], psi.shape[2], order="F")
    assert psi.shape[1] == 1
    return psi[:, 0]


def site_phys_dims(gp):
    return [site_dims(ix, iy, gp.nx, gp.ny, gp.gd)[0] for iy in range(1, gp.ny + 1) for ix in range(1, gp.nx + 1)]


def apply_two_site(psi, dims, i, j, O):
    """Apply O (acting on sites i<j... any order; kron(A_i, B_j) with j fast) to psi."""
    n = len(dims)
    T = psi.reshape(dims[::-1])  # C-order reshape of a column-major flat index
    ai, aj = n - 1 - i, n - 1 - j
    di, dj = dims[i], dims[j]
    O4 = O.reshape(di, dj, di, dj)  # [pi', pj', pi, pj] (row-major kron: j fast)
    out = np.tensordot(O4, T, axes=([2, 3], [ai, aj]))  # (pi', pj', rest...)
    rest = [k for k in range(n) if k not in (ai, aj)]
    perm_src = [ai, aj] + rest
    inv = np.argsort(perm_src)
    out = np.transpose(out, inv)
    return out.reshape(-1)


def dense_energy(gp, g, t_hop, m):
    """<psi|H|psi>/<psi|psi> with H = sum of H_merged over all bonds (brute force)."""
    nx, ny, gd = gp.nx, gp.ny, gp.gd
    psi = dense_psi(gp)
    dims = site_phys_dims(gp)
    idx = lambda ix, iy: (iy - 1) * nx + (ix - 1)
    nrm = np.vdot(psi, psi).real
    E = 0.0
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            h = H_merged_h_site(ix, iy, nx, ny, gd, g, t_hop, m)
            E += np.vdot(psi, apply_two_site(psi, dims, idx(ix, iy), idx(ix + 1, iy), h)).real
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            h = H_merged_v_site(ix, iy, nx, ny, gd, g, t_hop, m)
            E += np.vdot(psi, apply_two_site(psi, dims, idx(ix, iy), idx(ix, iy + 1), h)).real
    return E / nrm
