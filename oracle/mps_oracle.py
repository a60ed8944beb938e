"""CPU oracle for the snake-MPS / DMRG engine (mps_lgt.jl).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` alone; the library's DMRG
(lgt_amd/csrc/dmrg.cpp, driven by lgt_amd/mps.py) never calls it.

Instead of restating the MPO algebra a second time, the oracle builds the
lattice Hamiltonian directly on the full product space of the snake chain
from the PEPS oracle's site and bond operators (H_onsite_site, H_hop_h_site,
H_hop_v_site, u1_lgt_hamiltonian.jl:658-801), the Gauss-law values G_i of
every product state, and ``mpo_to_dense`` (mps_lgt.jl:649-661) to bring an
MPO to the same space.  Small lattices only (2x2 at gd = 3 is 1296 states).

  * ``snake_nodes``, ``node_dims``  mps_lgt.jl:51-67
  * ``full_hamiltonian``            the sum that lgt_terms (mps_lgt.jl:83-116)
                                    lists, embedded operator by operator
  * ``gauss_values``                G_i = E_R(i) - E_R(i-x) + E_U(i) - E_U(i-y)
                                    - n_f(i) per product state (mps_lgt.jl:466-514)
  * ``mpo_to_dense``                mps_lgt.jl:649-661
  * ``krylov_evolve_dense``         krylov_evolve_step (mps_lgt.jl:546-586)
                                    on full-space vectors: the same Lanczos
                                    recurrence, full reorthogonalisation and
                                    exp(-i dt T) e_1, with exact sums in place
                                    of MPS direct sums and compressions
  * ``product_vector``              product_mps (mps_lgt.jl:200-202) as a
                                    full-space vector

Pins (tests/test_oracle_mps.py): the full-space Hamiltonian restricted to a
Gauss-law sector has exactly the spectrum of the ED oracle's Hamiltonian of
that sector (finite_ed.jl, itself pinned against the reference's fixtures),
and the Gauss values reproduce the ED basis size of every sector tested.
The Krylov step is pinned against exp(-i dt H) psi from the eigenvectors of
H (agreement to the Lanczos truncation error, which shrinks as m grows and
vanishes once m reaches the dimension of the Krylov space), and the product
vector against the Gauss sector of its charges.
"""
from __future__ import annotations

import numpy as np

from oracle import peps_oracle as P


def snake_nodes(nx, ny):
    """mps_lgt.jl:51-61 (positions from 0)."""
    chain = []
    for iy in range(1, ny + 1):
        xs = range(1, nx + 1) if iy % 2 == 1 else range(nx, 0, -1)
        chain += [(ix, iy) for ix in xs]
    return chain, {node: p for p, node in enumerate(chain)}


def node_dims(nx, ny, gd):
    """mps_lgt.jl:64-67."""
    chain, _ = snake_nodes(nx, ny)
    return [P.site_dims(ix, iy, nx, ny, gd)[0] for (ix, iy) in chain]


def _apply_local(dims, ops):
    """The full-space matrix of the product of `ops` ({chain position:
    (d, d) matrix}), identity elsewhere (first chain site slowest)."""
    N = int(np.prod(dims))
    out = np.zeros((N, N))
    eye = np.eye(N)
    # act on the identity, one factor at a time
    T = eye.reshape(list(dims) + [N])
    for p, O in ops.items():
        T = np.moveaxis(np.tensordot(np.real(O), T, axes=([1], [p])), 0, p)
    out[:] = T.reshape(N, N)
    return out


def full_hamiltonian(nx, ny, gd, g, t, m):
    """The lattice Hamiltonian on the snake chain's full product space."""
    _, pos = snake_nodes(nx, ny)
    dims = node_dims(nx, ny, gd)
    H = np.zeros((int(np.prod(dims)),) * 2)
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            H += _apply_local(dims, {pos[(ix, iy)]: P.H_onsite_site(ix, iy, nx, ny, gd, g, m)})
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            h2 = np.real(P.H_hop_h_site(ix, iy, nx, ny, gd, t))
            H += _apply_pair(dims, pos[(ix, iy)], pos[(ix + 1, iy)], h2)
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            h2 = np.real(P.H_hop_v_site(ix, iy, nx, ny, gd, t))
            H += _apply_pair(dims, pos[(ix, iy)], pos[(ix, iy + 1)], h2)
    return H


def _apply_pair(dims, pa, pb, h2):
    """Full-space matrix of a two-site operator h2 on (site pa) x (site pb)
    (pa's index slowest in h2)."""
    da, db = dims[pa], dims[pb]
    # split h2 into a sum of products by its operator-Schmidt decomposition
    M = h2.reshape(da, db, da, db).transpose(0, 2, 1, 3).reshape(da * da, db * db)
    U, S, Vt = np.linalg.svd(M)
    out = 0.0
    for k in range(len(S)):
        if S[k] <= 1e-15 * S[0]:
            break
        A = (U[:, k] * S[k]).reshape(da, da)
        B = Vt[k].reshape(db, db)
        out = out + _apply_local(dims, {pa: A, pb: B})
    return out


def gauss_values(nx, ny, gd):
    """G_i of every product state of the chain, shape (N, nx, ny)."""
    chain, pos = snake_nodes(nx, ny)
    dims = node_dims(nx, ny, gd)
    N = int(np.prod(dims))
    idx = np.array(np.unravel_index(np.arange(N), dims)).T      # (N, sites) local indices, 0-based
    nf = np.zeros((N, nx, ny))
    eR = np.zeros((N, nx, ny))
    eU = np.zeros((N, nx, ny))
    for (ix, iy), p in pos.items():
        _, dR, dU = P.site_dims(ix, iy, nx, ny, gd)
        loc = idx[:, p] + 1
        for s in range(N):
            f, lR, lU = P.decode_phys(int(loc[s]), dR, dU, gd)
            nf[s, ix - 1, iy - 1] = f
            eR[s, ix - 1, iy - 1] = lR / 2.0 if dR > 1 else 0.0
            eU[s, ix - 1, iy - 1] = lU / 2.0 if dU > 1 else 0.0
    G = eR + eU - nf
    G[:, 1:, :] -= eR[:, :-1, :]
    G[:, :, 1:] -= eU[:, :, :-1]
    return G


def mpo_to_dense(W):
    """mps_lgt.jl:649-661 (first site slowest)."""
    op = np.transpose(W[0][0], (1, 2, 0))                      # (k, b, Dr)
    for Wp in W[1:]:
        Dl, Dr, d, _ = Wp.shape
        K, B, _ = op.shape
        new = np.zeros((K * d, B * d, Dr))
        for r in range(Dr):
            for l in range(Dl):
                new[:, :, r] += np.kron(op[:, :, l], Wp[l, r])
        op = new
    return op[:, :, 0]


def product_vector(configs, dims):
    """product_mps (mps_lgt.jl:200-202) as a full-space vector: configs[p]
    in 1..dims[p], first chain site slowest."""
    idx = np.ravel_multi_index(tuple(int(c) - 1 for c in configs), dims)
    v = np.zeros(int(np.prod(dims)), dtype=complex)
    v[idx] = 1.0
    return v


def krylov_evolve_dense(H, psi, dt, m=8):
    """krylov_evolve_step (mps_lgt.jl:546-586) on a full-space vector: Lanczos
    with the three-term recurrence, full reorthogonalisation against every
    basis vector, stop on beta < 1e-12; psi_new = V exp(-i dt T) e_1,
    renormalised."""
    V = [psi / np.linalg.norm(psi)]
    al, be = [], []
    w = H @ V[0]
    al.append(float(np.real(np.vdot(V[0], w))))
    w = w - al[0] * V[0]
    for j in range(2, m + 1):
        b = float(np.linalg.norm(w))
        if b < 1e-12:
            break
        be.append(b)
        V.append(w / b)
        w = H @ V[-1]
        a = float(np.real(np.vdot(V[-1], w)))
        al.append(a)
        w = w - a * V[-1] - b * V[-2]
        for vk in V:
            w = w - np.vdot(vk, w) * vk
    k = len(al)
    T = np.diag(al) + np.diag(be[:k - 1], 1) + np.diag(be[:k - 1], -1)
    wT, QT = np.linalg.eigh(T)
    coeffs = QT @ (np.exp(-1j * dt * wT) * QT[0])
    out = sum(c * v for c, v in zip(coeffs, V))
    return out / np.linalg.norm(out)
