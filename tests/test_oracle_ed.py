"""Pins of the ED oracle (oracle/ed_oracle.py, finite_ed.jl) on the CPU.

Every check uses something the restatement does not: the Gauss law applied
by brute force to the whole product space, a Hamiltonian assembled from
explicit local operators by Kronecker products, the dense eigensolver, the
t = 0 closed form, the full-space state for the observables and the dense
matrix exponential for the real-time step."""
import itertools

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sps

from oracle import ed_oracle as ED


def staggered(nx, ny):
    return np.array([[-1 if (ix + iy) % 2 else 0 for iy in range(1, ny + 1)] for ix in range(1, nx + 1)])


def product_space(nx, ny, dg):
    """Every (n_f, e_R, e_U) of the lattice, in Kronecker order: the fermion
    of site (ix, iy) (iy outer) first, then the e_R links, then the e_U links,
    the last factor fastest."""
    ns, nh, nv = nx * ny, (nx - 1) * ny, nx * (ny - 1)
    vals = range(-dg, dg + 1)
    out = []
    for occ in itertools.product((0, 1), repeat=ns):
        for er in itertools.product(vals, repeat=nh):
            for eu in itertools.product(vals, repeat=nv):
                n_f = np.array(occ).reshape(ny, nx).T
                e_R = np.array(er, dtype=int).reshape(ny, nx - 1).T if nh else np.zeros((nx - 1, ny), int)
                e_U = np.array(eu, dtype=int).reshape(ny - 1, nx).T if nv else np.zeros((nx, ny - 1), int)
                out.append((n_f, e_R, e_U))
    return out


def gauss(n_f, e_R, e_U, nx, ny):
    G = np.zeros((nx, ny), int)
    for ix in range(nx):
        for iy in range(ny):
            r = e_R[ix, iy] if ix < nx - 1 else 0
            l_ = e_R[ix - 1, iy] if ix > 0 else 0
            u = e_U[ix, iy] if iy < ny - 1 else 0
            d = e_U[ix, iy - 1] if iy > 0 else 0
            G[ix, iy] = r - l_ + u - d - n_f[ix, iy]
    return G


SECTORS = [(2, 2, 1, "stag"), (2, 2, 2, "stag"), (2, 3, 1, "stag"), (3, 2, 1, "stag"), (2, 2, 1, "zero"),
           (3, 2, 1, "string")]


def charges(nx, ny, kind):
    if kind == "stag":
        return staggered(nx, ny)
    g = np.zeros((nx, ny), int)
    if kind == "string":  # finite_ed.jl:17-19 quench A: g(1,2) = +1, g(2,2) = -2
        g[0, 1], g[1, 1] = 1, -2
    return g


@pytest.mark.parametrize("nx,ny,dg,kind", SECTORS)
def test_basis_is_the_brute_force_gauss_sector_in_reference_order(nx, ny, dg, kind):
    g = charges(nx, ny, kind)
    b = ED.build_basis(nx, ny, dg, g)
    keep = [s for s in product_space(nx, ny, dg) if np.array_equal(gauss(*s, nx, ny), g)]

    def order(s):  # finite_ed.jl:140-206: fermion mask, then the free E_U values
        n_f, _, e_U = s
        mask = sum(int(n_f[ix, iy]) << (iy * nx + ix) for iy in range(ny) for ix in range(nx))
        free = tuple(int(e_U[ix, iy]) for iy in range(ny - 1) for ix in range(nx - 1))
        return (mask, free)

    keep.sort(key=order)
    assert b.dim == len(keep) > 0
    np.testing.assert_array_equal(b.n_f, np.array([s[0] for s in keep]))
    np.testing.assert_array_equal(b.e_R, np.array([s[1] for s in keep]).reshape(b.e_R.shape))
    np.testing.assert_array_equal(b.e_U, np.array([s[2] for s in keep]).reshape(b.e_U.shape))
    assert len(set(b.keys.tolist())) == b.dim


def test_key_layout_matches_the_reference_for_dg1():
    # finite_ed.jl:79-95 by hand on a 2x2 state: n_f bits 0..3, e_R + 1 at
    # bits 4..7 (2 bits each), e_U + 1 at bits 8..11
    n_f = np.array([[[1, 0], [0, 1]]])       # n_f[ix, iy]
    e_R = np.array([[[1, -1]]])               # e_R[1,1] = 1, e_R[1,2] = -1
    e_U = np.array([[[0], [1]]])              # e_U[1,1] = 0, e_U[2,1] = 1
    want = (1 << 0) | (1 << 3) | (2 << 4) | (0 << 6) | (1 << 8) | (2 << 10)
    assert int(ED.encode(n_f, e_R, e_U, 1)[0]) == want


def local_hamiltonian(nx, ny, dg, g, t, m):
    """H on the full product space from local operators (Kronecker order of
    product_space): g^2/2 E^2 on every link, m (-1)^(ix+iy) n_f on every
    site, and -t (c+_L c_R U+ + h.c.) on every bond, with the hard-core
    hopping of finite_ed.jl:280-320 (no fermionic sign)."""
    ns, nh, nv = nx * ny, (nx - 1) * ny, nx * (ny - 1)
    dl = 2 * dg + 1
    dims = [2] * ns + [dl] * (nh + nv)
    n = np.diag([0.0, 1.0])
    sp = np.array([[0.0, 0.0], [1.0, 0.0]])   # |1><0|
    E = np.diag(np.arange(-dg, dg + 1, dtype=float))
    Up = np.diag(np.ones(dl - 1), -1)          # e -> e + 1

    def op(factors):
        out = sps.identity(1, format="csr")
        for k, d in enumerate(dims):
            out = sps.kron(out, sps.csr_matrix(factors[k]) if k in factors else sps.identity(d), format="csr")
        return out

    site = lambda ix, iy: iy * nx + ix
    hlink = lambda ix, iy: ns + iy * (nx - 1) + ix
    vlink = lambda ix, iy: ns + nh + iy * nx + ix
    H = 0.0
    for iy in range(ny):
        for ix in range(nx):
            stag = 1.0 if (ix + iy) % 2 == 0 else -1.0
            H = H + m * stag * op({site(ix, iy): n})
            if ix < nx - 1:
                H = H + (g * g / 2) * op({hlink(ix, iy): E @ E})
                hop = op({site(ix, iy): sp, site(ix + 1, iy): sp.T, hlink(ix, iy): Up})
                H = H - t * (hop + hop.T)
            if iy < ny - 1:
                H = H + (g * g / 2) * op({vlink(ix, iy): E @ E})
                hop = op({site(ix, iy): sp, site(ix, iy + 1): sp.T, vlink(ix, iy): Up})
                H = H - t * (hop + hop.T)
    return H


def full_index(nx, ny, dg, s):
    n_f, e_R, e_U = s
    digits = [int(n_f[ix, iy]) for iy in range(ny) for ix in range(nx)]
    digits += [int(e_R[ix, iy]) + dg for iy in range(ny) for ix in range(nx - 1)]
    digits += [int(e_U[ix, iy]) + dg for iy in range(ny - 1) for ix in range(nx)]
    dims = [2] * (nx * ny) + [2 * dg + 1] * (len(digits) - nx * ny)
    k = 0
    for d, v in zip(dims, digits):
        k = k * d + v
    return k


@pytest.mark.parametrize("nx,ny,dg,kind", [(2, 2, 1, "stag"), (2, 2, 2, "stag"), (3, 2, 1, "string")])
def test_hamiltonian_is_the_local_operator_sum_restricted_to_the_sector(nx, ny, dg, kind):
    g_, t, m = 0.7, 1.3, 0.4
    b = ED.build_basis(nx, ny, dg, charges(nx, ny, kind))
    H = ED.build_hamiltonian(b, g_, t, m).toarray()
    idx = [full_index(nx, ny, dg, (b.n_f[i], b.e_R[i], b.e_U[i])) for i in range(b.dim)]
    Hf = sps.csr_matrix(local_hamiltonian(nx, ny, dg, g_, t, m))
    cols = Hf[:, idx].tocsr()
    np.testing.assert_allclose(H, cols[idx, :].toarray(), rtol=0, atol=1e-14)
    # the sector is invariant: no amplitude leaks out of it
    out = np.setdiff1d(np.arange(Hf.shape[0]), idx)
    assert cols[out, :].count_nonzero() == 0


def test_ground_state_dense_vs_lanczos_and_closed_forms():
    b = ED.build_basis(3, 3, 1, staggered(3, 3))
    H = ED.build_hamiltonian(b, 1.0, 1.0, 0.5)
    assert b.dim > 2000  # the Lanczos (ARPACK) branch
    E0, psi = ED.find_ground_state(H)
    w = np.linalg.eigvalsh(H.toarray())
    assert abs(E0 - w[0]) <= 1e-10 * abs(w[0])
    assert np.linalg.norm(H @ psi - E0 * psi) <= 1e-9 * abs(E0)
    # t = 0: the Hamiltonian is diagonal
    H0 = ED.build_hamiltonian(b, 1.0, 0.0, 0.5)
    assert ED.find_ground_state(H0)[0] == pytest.approx(ED.diagonal(b, 1.0, 0.5).min(), abs=1e-12)


def test_measure_matches_the_full_space_state_and_conserves_particles():
    nx, ny, dg = 3, 2, 1
    g = charges(nx, ny, "string")
    b = ED.build_basis(nx, ny, dg, g)
    E0, psi = ED.find_ground_state(ED.build_hamiltonian(b, 1.0, 1.0, 2.0))
    obs = ED.measure_ed(psi, b)
    assert obs["nf_grid"].sum() == pytest.approx(-g.sum(), abs=1e-12)
    # the same observables from the state embedded in the product space
    full = product_space(nx, ny, dg)
    amp = np.zeros(len(full))
    for i in range(b.dim):
        amp[full_index(nx, ny, dg, (b.n_f[i], b.e_R[i], b.e_U[i]))] = psi[i]
    p = amp ** 2
    nf = sum(p[k] * full[k][0] for k in np.nonzero(p)[0])
    np.testing.assert_allclose(obs["nf_grid"], nf, atol=1e-13)
    ix = np.arange(1, nx + 1)[:, None] + np.arange(1, ny + 1)[None, :]
    assert obs["nf_even"] == pytest.approx(nf[ix % 2 == 0].mean(), abs=1e-13)


def test_evolve_matches_the_dense_exponential():
    b = ED.build_basis(3, 2, 1, charges(3, 2, "string"))
    H = ED.build_hamiltonian(b, 1.0, 1.0, 2.0)
    rng = np.random.default_rng(0)
    psi = rng.standard_normal(b.dim) + 1j * rng.standard_normal(b.dim)
    psi /= np.linalg.norm(psi)
    out = ED.evolve_ed(psi, H, 0.05, 3)
    U = sla.expm(-1j * 0.05 * H.toarray())
    np.testing.assert_allclose(out[-1], U @ U @ U @ psi, atol=1e-12)


def test_library_state_key_matches_the_oracle_encoding():
    """lgt_amd.ed_quench.state_key (_encode, finite_ed.jl:79-96) gives the
    oracle's key for every state of two sectors (3x2 string sector at
    dg = 1, 2x2 staggered at dg = 2)."""
    from lgt_amd import ed_quench
    from oracle import peps_oracle as P
    for nx, ny, dg, g in ((3, 2, 1, ed_quench.string_charges(3, 2)),
                          (2, 2, 2, np.asarray(P.staggered_charges(2, 2)))):
        b = ED.build_basis(nx, ny, dg, g)
        assert b.dim > 0
        for s in range(0, b.dim, max(1, b.dim // 50)):
            k = ed_quench.state_key(b.n_f[s], b.e_R[s], b.e_U[s], dg)
            assert k == int(b.keys[s]), (nx, ny, dg, s)
