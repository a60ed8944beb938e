"""CPU oracle for the exact diagonalisation of the finite U(1) lattice gauge
theory (finite_ed.jl).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` alone; the library's ED
(lgt_amd/csrc/ed.hip) never calls it.

Restated from the reference, function by function:
  * ``encode``            finite_ed.jl:79-95   (the key layout; links take
                          ``link_bits(dg)`` bits, 2 for dg = 1 as in the
                          reference, whose 2-bit field only holds dg = 1)
  * ``build_basis``       finite_ed.jl:125-233 (Gauss-law enumeration, in the
                          reference's order: fermion masks ascending, then the
                          free E_U values lexicographically with E_U(1,1) the
                          slowest, each running -dg..dg)
  * ``build_hamiltonian`` finite_ed.jl:252-326 (diagonal electric + staggered
                          mass, hopping -t with the link raised on a forward
                          hop; no plaquette term)
  * ``find_ground_state`` finite_ed.jl:333-341 (lowest eigenpair; the
                          reference uses KrylovKit's Lanczos, here ARPACK's
                          through scipy, or the dense eigensolver when small)
  * ``measure_ed``        finite_ed.jl:369-395
  * ``evolve_ed``         finite_ed.jl:404-430 (exp(-i dt H) psi per step,
                          renormalised; scipy's expm_multiply)

Pins (tests/test_oracle_ed.py): the basis against a brute-force Gauss-law
filter of the full product space, the Hamiltonian against one assembled from
explicit local operators on that space, the ground state against the dense
eigensolver and the t = 0 closed form, the observables against the full-space
state, and the particle-number invariant.
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla


def link_bits(dg):
    return max(2, int(np.ceil(np.log2(2 * dg + 1))))


def key_bits(nx, ny, dg):
    return nx * ny + link_bits(dg) * ((nx - 1) * ny + nx * (ny - 1))


def encode(n_f, e_R, e_U, dg):
    """finite_ed.jl:79-95: n_f bits (iy outer, ix inner), then e_R + dg and
    e_U + dg in link_bits(dg)-bit fields in the same order.  Arrays carry a
    leading state axis; returns uint64 keys."""
    n_f = np.asarray(n_f)
    nx, ny = n_f.shape[1], n_f.shape[2]
    lb = link_bits(dg)
    key = np.zeros(n_f.shape[0], np.uint64)
    off = 0
    for iy in range(ny):
        for ix in range(nx):
            key |= n_f[:, ix, iy].astype(np.uint64) << np.uint64(off)
            off += 1
    for iy in range(ny):
        for ix in range(nx - 1):
            key |= (e_R[:, ix, iy].astype(np.int64) + dg).astype(np.uint64) << np.uint64(off)
            off += lb
    for iy in range(ny - 1):
        for ix in range(nx):
            key |= (e_U[:, ix, iy].astype(np.int64) + dg).astype(np.uint64) << np.uint64(off)
            off += lb
    return key


class Basis:
    """The gauge-sector basis: per-state n_f (nx, ny), e_R (nx-1, ny),
    e_U (nx, ny-1) and the encoded keys, in the reference's order."""

    def __init__(self, nx, ny, dg, n_f, e_R, e_U):
        self.nx, self.ny, self.dg = nx, ny, dg
        self.n_f, self.e_R, self.e_U = n_f, e_R, e_U
        self.keys = encode(n_f, e_R, e_U, dg)
        self.index = {int(k): i for i, k in enumerate(self.keys)}

    @property
    def dim(self):
        return len(self.keys)


def build_basis(nx, ny, dg, g_charges):
    """finite_ed.jl:125-233.  g_charges[ix-1][iy-1] is the background charge."""
    g = np.asarray(g_charges, dtype=np.int64).reshape(nx, ny)
    nsite = nx * ny
    nf_target = -int(g.sum())
    assert 0 <= nf_target <= nsite, f"Inconsistent sector: N_f = {nf_target}"
    F = (nx - 1) * (ny - 1)
    vals = np.arange(-dg, dg + 1)
    # every free E_U assignment, E_U(1,1) the most significant digit
    combos = np.array(list(itertools.product(vals, repeat=F)), dtype=np.int64).reshape(-1, F)
    out_nf, out_eR, out_eU = [], [], []
    for mask in range(2 ** nsite):
        if bin(mask).count("1") != nf_target:
            continue
        n_f = np.array([[(mask >> (iy * nx + ix)) & 1 for iy in range(ny)] for ix in range(nx)], np.int64)
        q = n_f + g
        S = np.cumsum(q.sum(axis=0))
        if S[ny - 1] != 0:
            continue
        nc = combos.shape[0]
        eU = np.zeros((nc, nx, max(ny - 1, 0)), np.int64)
        ok = np.ones(nc, bool)
        for iy in range(ny - 1):
            free = combos[:, iy * (nx - 1):(iy + 1) * (nx - 1)]
            eU[:, :nx - 1, iy] = free
            last = S[iy] - free.sum(axis=1)
            ok &= (last >= -dg) & (last <= dg)
            eU[:, nx - 1, iy] = last
        eR = np.zeros((nc, max(nx - 1, 0), ny), np.int64)
        for iy in range(ny):
            acc = np.zeros(nc, np.int64)
            for ix in range(nx):
                now = eU[:, ix, iy] if iy < ny - 1 else 0
                prev = eU[:, ix, iy - 1] if iy > 0 else 0
                acc = acc + q[ix, iy] - now + prev
                if ix < nx - 1:
                    ok &= (acc >= -dg) & (acc <= dg)
                    eR[:, ix, iy] = acc
                else:
                    ok &= acc == 0
        k = int(ok.sum())
        if k:
            out_nf.append(np.broadcast_to(n_f, (k, nx, ny)).copy())
            out_eR.append(eR[ok])
            out_eU.append(eU[ok])
    if not out_nf:
        z = np.zeros
        return Basis(nx, ny, dg, z((0, nx, ny), np.int8), z((0, nx - 1, ny), np.int8), z((0, nx, ny - 1), np.int8))
    return Basis(nx, ny, dg, np.concatenate(out_nf).astype(np.int8), np.concatenate(out_eR).astype(np.int8),
                 np.concatenate(out_eU).astype(np.int8))


def diagonal(b, g, m):
    """The on-site part of finite_ed.jl:266-278 for every state."""
    e2 = (b.e_R.astype(float) ** 2).sum(axis=(1, 2)) + (b.e_U.astype(float) ** 2).sum(axis=(1, 2))
    ix = np.arange(1, b.nx + 1)[:, None]
    iy = np.arange(1, b.ny + 1)[None, :]
    stag = np.where((ix + iy) % 2 == 0, 1.0, -1.0)
    return (g * g / 2.0) * e2 + m * (b.n_f.astype(float) * stag[None]).sum(axis=(1, 2))


def build_hamiltonian(b, g, t, m):
    """finite_ed.jl:252-326 as a scipy CSR matrix (real: every entry is).
    The reference's loop over states runs here over all states at once per
    bond: the forward-hop condition, the hopped state's key, its index."""
    d = diagonal(b, g, m)
    I, J, V = [np.arange(b.dim)], [np.arange(b.dim)], [d]
    nx, ny, dg = b.nx, b.ny, b.dg
    order = np.argsort(b.keys)
    sorted_keys = b.keys[order]

    def link(src, jdx):
        k = encode(*src, dg)
        pos = np.searchsorted(sorted_keys, k)
        pos = np.minimum(pos, b.dim - 1)
        hit = sorted_keys[pos] == k
        j = order[pos[hit]]
        i = jdx[hit]
        I.extend([j, i]); J.extend([i, j]); V.extend([np.full(len(i), -t), np.full(len(i), -t)])

    for iy in range(ny):
        for ix in range(nx - 1):
            e = b.e_R[:, ix, iy]
            s = np.nonzero((b.n_f[:, ix, iy] == 0) & (b.n_f[:, ix + 1, iy] == 1) & (e < dg))[0]
            nf = b.n_f[s].copy(); eR = b.e_R[s].copy()
            nf[:, ix, iy], nf[:, ix + 1, iy] = 1, 0
            eR[:, ix, iy] += 1
            link((nf, eR, b.e_U[s]), s)
    for iy in range(ny - 1):
        for ix in range(nx):
            e = b.e_U[:, ix, iy]
            s = np.nonzero((b.n_f[:, ix, iy] == 0) & (b.n_f[:, ix, iy + 1] == 1) & (e < dg))[0]
            nf = b.n_f[s].copy(); eU = b.e_U[s].copy()
            nf[:, ix, iy], nf[:, ix, iy + 1] = 1, 0
            eU[:, ix, iy] += 1
            link((nf, b.e_R[s], eU), s)
    return sps.csr_matrix((np.concatenate(V), (np.concatenate(I), np.concatenate(J))), shape=(b.dim, b.dim))


def find_ground_state(H):
    """finite_ed.jl:333-341: (E0, psi0) of the lowest eigenvalue."""
    n = H.shape[0]
    if n <= 2000:
        w, v = np.linalg.eigh(H.toarray())
        return float(w[0]), v[:, 0]
    w, v = spla.eigsh(H, k=1, which="SA", tol=1e-13)
    return float(w[0]), v[:, 0]


def measure_ed(psi, b):
    """finite_ed.jl:369-395 -> dict(nf_mean, nf_even, nf_odd, E2_mean,
    nf_grid, E2_grid)."""
    p = np.abs(psi) ** 2
    nf_grid = np.einsum("s,sxy->xy", p, b.n_f.astype(float))
    e2 = np.zeros((b.dim, b.nx, b.ny))
    e2[:, :b.nx - 1, :] += b.e_R.astype(float) ** 2
    e2[:, :, :b.ny - 1] += b.e_U.astype(float) ** 2
    E2_grid = np.einsum("s,sxy->xy", p, e2)
    ix = np.arange(1, b.nx + 1)[:, None]
    iy = np.arange(1, b.ny + 1)[None, :]
    even = (ix + iy) % 2 == 0
    return dict(nf_mean=float(nf_grid.mean()), nf_even=float(nf_grid[even].mean()),
                nf_odd=float(nf_grid[~even].mean()), E2_mean=float(E2_grid.mean()),
                nf_grid=nf_grid, E2_grid=E2_grid)


def evolve_ed(psi, H, dt, n_steps):
    """finite_ed.jl:404-430: psi <- exp(-i dt H) psi, renormalised, n_steps
    times; returns the list of states after each step."""
    out = []
    psi = np.asarray(psi, dtype=complex)
    for _ in range(n_steps):
        psi = spla.expm_multiply(-1j * dt * H, psi)
        psi = psi / np.linalg.norm(psi)
        out.append(psi)
    return out
