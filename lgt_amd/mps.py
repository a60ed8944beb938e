"""Snake-MPS / DMRG engine for the U(1) lattice gauge theory (mps_lgt.jl).

The Hamiltonian is assembled as a sum of product-operator terms into a
compressed MPO on the boustrophedon chain of lattice sites (host side, the
SVDs through the library), and its ground state is found by two-site DMRG on
the GPU (`peps.dmrg`: environments, effective Hamiltonian, Lanczos and the
truncated split on device tensors).  Observables are MPO expectations on the
GPU.  Real arithmetic: the Hamiltonian is real symmetric.

Conventions: MPS tensors (Dl, Dr, d), MPO tensors (Dl, Dr, k, b) with k the
bra and b the ket physical index (the reference's CMPS / CMPO); `gd` is the
link dimension (the reference's dg = (gd - 1) / 2: gd = 3 for dg = 1); the
site space is fermion (slowest) x right link x up link (fastest), missing
boundary links dropped (site_dims, u1_lgt_hamiltonian.jl:605-611).
"""
from __future__ import annotations

import numpy as np

from . import lattice, peps


def snake_nodes(nx, ny):
    """snake_nodes (mps_lgt.jl:51-61): (chain, pos), chain[p] = (ix, iy),
    odd rows left to right, even rows right to left; p from 0."""
    chain = []
    for iy in range(1, ny + 1):
        xs = range(1, nx + 1) if iy % 2 == 1 else range(nx, 0, -1)
        chain += [(ix, iy) for ix in xs]
    return chain, {node: p for p, node in enumerate(chain)}


def node_dims(nx, ny, gd):
    """node_dims (mps_lgt.jl:64-67)."""
    chain, _ = snake_nodes(nx, ny)
    return [lattice.site_dims(ix, iy, nx, ny, gd)[0] for (ix, iy) in chain]


def _ops(gd):
    E = np.diag((np.arange(gd) - (gd - 1) / 2.0))          # E = (2i - (gd - 1)) / 2
    U = np.zeros((gd, gd))
    for k in range(gd - 1):
        U[k + 1, k] = 1.0                                   # U|e> = |e + 1>
    c = np.array([[0.0, 1.0], [0.0, 0.0]])
    nf = np.diag([0.0, 1.0])
    return E, U, c, nf


def _embed(ix, iy, nx, ny, gd, f=None, R=None, Up=None):
    """fermion x right-link x up-link operator of site (ix, iy) (identities
    where None; embed_f_site / embed_R_site / embed_U_site)."""
    _, dR, dU = lattice.site_dims(ix, iy, nx, ny, gd)
    return np.kron(np.kron(np.eye(2) if f is None else f, np.eye(dR) if R is None else R),
                   np.eye(dU) if Up is None else Up)


def lgt_terms(nx, ny, gd, g, t, m):
    """lgt_terms (mps_lgt.jl:83-116): (coef, {node: op}) for every on-site
    term (g^2/2)(E_R^2 + E_U^2) + m (-1)^(ix+iy) n_f and every hopping
    -t (c^dag U c + h.c.), horizontal (U on the right link of the left site)
    and vertical (U on the up link of the lower site)."""
    _, pos = snake_nodes(nx, ny)
    E, U, c, nf = _ops(gd)
    terms = []
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            H = m * (1 if (ix + iy) % 2 == 0 else -1) * _embed(ix, iy, nx, ny, gd, f=nf)
            if ix < nx:
                H = H + (g ** 2 / 2) * _embed(ix, iy, nx, ny, gd, R=E @ E)
            if iy < ny:
                H = H + (g ** 2 / 2) * _embed(ix, iy, nx, ny, gd, Up=E @ E)
            terms.append((1.0, {pos[(ix, iy)]: H}))
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            A = _embed(ix, iy, nx, ny, gd, f=c.T, R=U)
            B = _embed(ix + 1, iy, nx, ny, gd, f=c)
            pL, pR = pos[(ix, iy)], pos[(ix + 1, iy)]
            terms.append((-t, {pL: A, pR: B}))
            terms.append((-t, {pL: A.T.copy(), pR: B.T.copy()}))
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            A = _embed(ix, iy, nx, ny, gd, f=c.T, Up=U)
            B = _embed(ix, iy + 1, nx, ny, gd, f=c)
            pD, pU = pos[(ix, iy)], pos[(ix, iy + 1)]
            terms.append((-t, {pD: A, pU: B}))
            terms.append((-t, {pD: A.T.copy(), pU: B.T.copy()}))
    return terms


def gauss_penalty_terms(nx, ny, gd, gch, lam):
    """gauss_penalty_terms (mps_lgt.jl:474-514): Lambda (G_i - g_i)^2 with
    G_i = E_R(i) - E_R(i - x) + E_U(i) - E_U(i - y) - n_f(i), expanded as
    Lambda sum_ab O_a O_b - 2 Lambda g_i sum_a O_a (the constant Lambda g_i^2
    dropped)."""
    _, pos = snake_nodes(nx, ny)
    E, _, _, nf = _ops(gd)
    terms = []
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            pieces = {}

            def add(node, O):
                pieces[node] = pieces[node] + O if node in pieces else O

            if ix < nx:
                add(pos[(ix, iy)], _embed(ix, iy, nx, ny, gd, R=E))
            if ix > 1:
                add(pos[(ix - 1, iy)], -_embed(ix - 1, iy, nx, ny, gd, R=E))
            if iy < ny:
                add(pos[(ix, iy)], _embed(ix, iy, nx, ny, gd, Up=E))
            if iy > 1:
                add(pos[(ix, iy - 1)], -_embed(ix, iy - 1, nx, ny, gd, Up=E))
            add(pos[(ix, iy)], -_embed(ix, iy, nx, ny, gd, f=nf))
            gi = float(gch[ix - 1][iy - 1])
            nodes = list(pieces)
            for a in nodes:
                for b in nodes:
                    terms.append((lam, {a: pieces[a] @ pieces[a]} if a == b else {a: pieces[a], b: pieces[b]}))
            for a in nodes:
                terms.append((-2.0 * lam * gi, {a: pieces[a]}))
    return terms


def term_mpo(dims, term):
    """term_mpo (mps_lgt.jl:119-129): bond-1 MPO, coefficient on the first site."""
    coef, ops = term
    W = []
    for p, d in enumerate(dims):
        O = ops.get(p, np.eye(d))
        if p == 0:
            O = coef * O
        W.append(np.asarray(O, dtype=np.float64).reshape(1, 1, d, d))
    return W


def mpo_axpby(alpha, A, beta, B):
    """mpo_axpby (mps_lgt.jl:132-155): alpha A + beta B as a direct sum."""
    n = len(A)
    C = []
    for p in range(n):
        (la, ra, d, _), (lb, rb, _, _) = A[p].shape, B[p].shape
        if p == 0:
            T = np.zeros((1, ra + rb, d, d))
            T[0, :ra] = alpha * A[p][0]
            T[0, ra:] = beta * B[p][0]
        elif p == n - 1:
            T = np.zeros((la + lb, 1, d, d))
            T[:la, 0] = A[p][:, 0]
            T[la:, 0] = B[p][:, 0]
        else:
            T = np.zeros((la + lb, ra + rb, d, d))
            T[:la, :ra] = A[p]
            T[la:, ra:] = B[p]
        C.append(T)
    return C


def mpo_compress(W, eps=1e-12, Dmax=400, ctx=None):
    """mpo_compress! (mps_lgt.jl:157-181): a left-to-right SVD sweep, then a
    right-to-left one truncating singular values below eps s_1 or beyond Dmax
    (the SVDs through the library)."""
    n = len(W)
    W = [w.copy() for w in W]
    for p in range(n - 1):
        Dl, Dr, d, _ = W[p].shape
        U, S, Vt = peps.svd(np.transpose(W[p], (0, 2, 3, 1)).reshape(Dl * d * d, Dr), ctx=ctx)
        r = len(S)
        W[p] = np.transpose(U.reshape(Dl, d, d, r), (0, 3, 1, 2))
        W[p + 1] = np.einsum("ij,jklm->iklm", S[:, None] * Vt, W[p + 1])
    for p in range(n - 1, 0, -1):
        Dl, Dr, d, _ = W[p].shape
        U, S, Vt = peps.svd(np.transpose(W[p], (0, 2, 3, 1)).reshape(Dl, d * d * Dr), ctx=ctx)
        keep = max(1, min(Dmax, int(np.sum(S > eps * S[0])) if S[0] > 0 else 1))
        W[p] = np.transpose(Vt[:keep].reshape(keep, d, d, Dr), (0, 3, 1, 2))
        US = U[:, :keep] * S[:keep]
        W[p - 1] = np.einsum("iajk,al->iljk", W[p - 1], US)
    return W


def assemble_mpo(dims, terms, eps=1e-12, ctx=None):
    """_assemble_mpo / build_H_mpo (mps_lgt.jl:183-197, 516-524): the terms
    summed one by one, compressed every 8 terms and at the end."""
    H = term_mpo(dims, terms[0])
    for k in range(1, len(terms)):
        H = mpo_axpby(1.0, H, 1.0, term_mpo(dims, terms[k]))
        if (k + 1) % 8 == 0:
            H = mpo_compress(H, eps, ctx=ctx)
    return mpo_compress(H, eps, ctx=ctx)


def build_H_mpo(nx, ny, gd, g, t, m, ctx=None):
    """build_H_mpo (mps_lgt.jl:183-197): the bare Hamiltonian."""
    return assemble_mpo(node_dims(nx, ny, gd), lgt_terms(nx, ny, gd, g, t, m), ctx=ctx)


def build_penalized_H(nx, ny, gd, g, t, m, gch, lam, ctx=None):
    """build_penalized_H (mps_lgt.jl:527-535): H + Lambda sum_i (G_i - g_i)^2."""
    return assemble_mpo(node_dims(nx, ny, gd),
                        lgt_terms(nx, ny, gd, g, t, m) + gauss_penalty_terms(nx, ny, gd, gch, lam), ctx=ctx)


def product_mps(configs, dims):
    """product_mps (mps_lgt.jl:200-202): configs[p] in 1..d (one-hot)."""
    out = []
    for cfg, d in zip(configs, dims):
        a = np.zeros((1, 1, d))
        a[0, 0, cfg - 1] = 1.0
        out.append(a)
    return out


def staggered_mps(nx, ny, gd):
    """staggered_mps (mps_lgt.jl:615-624): n_f = 1 on odd ix + iy, E = 0."""
    chain, _ = snake_nodes(nx, ny)
    cfg = []
    for (ix, iy) in chain:
        _, dR, dU = lattice.site_dims(ix, iy, nx, ny, gd)
        nf = 1 if (ix + iy) % 2 == 1 else 0
        iR = 1 if dR == 1 else (gd - 1) // 2 + 1
        iU = 1 if dU == 1 else (gd - 1) // 2 + 1
        cfg.append(nf * dR * dU + (iR - 1) * dU + iU)
    return product_mps(cfg, node_dims(nx, ny, gd))


def string_breaking_mps(nx, ny, gd, iy_str=None):
    """string_breaking_mps (mps_lgt.jl:624-637): the vacuum with n_f(2, iy_str)
    = 1 and E_R(1, iy_str) = 1 (the sector of string_charges)."""
    iy_str = ny // 2 if iy_str is None else iy_str
    chain, _ = snake_nodes(nx, ny)
    cfg = []
    for (ix, iy) in chain:
        _, dR, dU = lattice.site_dims(ix, iy, nx, ny, gd)
        nf = 1 if (ix == 2 and iy == iy_str) else 0
        eR = 1 if (ix == 1 and iy == iy_str) else 0
        iR = 1 if dR == 1 else eR + (gd - 1) // 2 + 1
        iU = 1 if dU == 1 else (gd - 1) // 2 + 1
        cfg.append(nf * dR * dU + (iR - 1) * dU + iU)
    return product_mps(cfg, node_dims(nx, ny, gd))


def string_charges(nx, ny, iy_str=None):
    """string_charges (mps_lgt.jl:640-643): g(1, iy_str) = 1, g(2, iy_str) = -2."""
    iy_str = ny // 2 if iy_str is None else iy_str
    g = np.zeros((nx, ny), dtype=np.int64)
    g[0, iy_str - 1], g[1, iy_str - 1] = 1, -2
    return g


def random_mps(dims, D, seed):
    """A seeded random MPS with bonds min(D, growth) (the reference's default
    start, real entries)."""
    rng = np.random.default_rng(seed)
    out, dl = [], 1
    for p, d in enumerate(dims):
        dr = 1 if p == len(dims) - 1 else min(D, dl * d)
        out.append(rng.standard_normal((dl, dr, d)))
        dl = dr
    return out


def dmrg_ground_state(H, dims, D=40, nsweeps=8, psi0=None, seed=1, tol=1e-7, maxiter=60, krylovdim=12,
                      etol=1e-7, ctx=None):
    """dmrg_ground_state (mps_lgt.jl:418-467) on the GPU: (E, psi, sweeps)."""
    psi0 = random_mps(dims, D, seed) if psi0 is None else psi0
    return peps.dmrg(H, psi0, D, nsweeps, tol, maxiter, krylovdim, etol, ctx=ctx)


def identity_mpo(dims):
    return [np.eye(d).reshape(1, 1, d, d) for d in dims]


def mps_overlap(psi, ctx=None):
    """<psi|psi> (mps_overlap, mps_lgt.jl:204-215, of an MPS with itself)."""
    return peps.mpo_expect(identity_mpo([a.shape[2] for a in psi]), psi, ctx=ctx)


def mps_local_expect(psi, p, O, ctx=None):
    """mps_local_expect (mps_lgt.jl:317-330)."""
    dims = [a.shape[2] for a in psi]
    W = identity_mpo(dims)
    W[p] = np.asarray(O, dtype=np.float64).reshape(1, 1, dims[p], dims[p])
    return peps.mpo_expect(W, psi, ctx=ctx) / mps_overlap(psi, ctx=ctx)


def measure_mps(psi, nx, ny, gd, t_now=0.0, ctx=None):
    """measure_mps (mps_lgt.jl:600-612): site-resolved <n_f>, <E^2> (node_nf_op,
    node_E2_op, :589-598), all in one pass of overlap environments; psi real
    or complex."""
    _, pos = snake_nodes(nx, ny)
    E, _, _, nf = _ops(gd)
    sites, ops, where = [], [], []
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            p = pos[(ix, iy)]
            O = np.zeros((psi[p].shape[2],) * 2)
            if ix < nx:
                O += _embed(ix, iy, nx, ny, gd, R=E @ E)
            if iy < ny:
                O += _embed(ix, iy, nx, ny, gd, Up=E @ E)
            sites += [p, p]
            ops += [_embed(ix, iy, nx, ny, gd, f=nf), O]
            where.append((ix - 1, iy - 1))
    vals = peps.mps_local_expect_many(psi, sites, ops, ctx=ctx)
    nfg, e2g = np.zeros((nx, ny)), np.zeros((nx, ny))
    for i, (a, b) in enumerate(where):
        nfg[a, b], e2g[a, b] = vals[2 * i], vals[2 * i + 1]
    par = (np.add.outer(np.arange(1, nx + 1), np.arange(1, ny + 1)) % 2 == 0)
    return dict(t=t_now, nf_mean=float(nfg.mean()), nf_even=float(nfg[par].mean()), nf_odd=float(nfg[~par].mean()),
                E2_mean=float(e2g.mean()), nf_grid=nfg, E2_grid=e2g)


def gauss_violation(psi, nx, ny, gd, gch, ctx=None):
    """gauss_violation (mps_lgt.jl:663-670): <sum_i (G_i - g_i)^2>."""
    dims = node_dims(nx, ny, gd)
    Hg = assemble_mpo(dims, gauss_penalty_terms(nx, ny, gd, gch, 1.0), ctx=ctx)
    return peps.mpo_expect(Hg, psi, ctx=ctx) / mps_overlap(psi, ctx=ctx) + float(np.sum(np.asarray(gch) ** 2))
