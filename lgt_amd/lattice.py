"""Host-side lattice set-up: site dimensions, Gauss-law labels and the seeded
random initial GaugePEPS used by the configs (small numpy arrays, uploaded
once).  Mirrors u1_lgt_hamiltonian.jl:605-618, finite_peps_gauge_invariant.jl
:58-67,109-136,180-218 with flux labels stored as integers 2E.
"""
from __future__ import annotations

import numpy as np


def boundary_label(gd):
    """Label of a missing boundary link: 0 for odd gd (reference), +1 for even gd."""
    return 0 if gd % 2 == 1 else 1


def link_labels(gd):
    return np.array([2 * i - (gd - 1) for i in range(gd)], dtype=np.int64)


def site_dims(ix, iy, nx, ny, gd):
    dR = gd if ix < nx else 1
    dU = gd if iy < ny else 1
    return 2 * dR * dU, dR, dU


def phys_labels(ix, iy, nx, ny, gd):
    """(n_f, lR, lU) of every physical index of site (ix, iy)."""
    d, dR, dU = site_dims(ix, iy, nx, ny, gd)
    p0 = np.arange(d)
    nf = p0 // (dR * dU)
    rem = p0 % (dR * dU)
    b0 = boundary_label(gd)
    lR = np.full(d, b0) if dR == 1 else 2 * (rem // dU) - (gd - 1)
    lU = np.full(d, b0) if dU == 1 else 2 * (rem % dU) - (gd - 1)
    return nf.astype(np.int64), np.asarray(lR, np.int64), np.asarray(lU, np.int64)


def staggered_charges(nx, ny):
    """g = -1 on odd ix+iy (gs_benchmark_v4.jl:46-52)."""
    ix = np.arange(1, nx + 1)[:, None]
    iy = np.arange(1, ny + 1)[None, :]
    return np.where((ix + iy) % 2 == 1, -1, 0).astype(np.int64)


def default_bond_labels(gd, D):
    labs = sorted(link_labels(gd).tolist(), key=lambda x: (abs(x), x))
    return np.array([labs[i % gd] for i in range(D)], dtype=np.int64)


def gauss_mask(ix, iy, nx, ny, gd, q, fl, fr, fu, fd):
    nf, lR, lU = phys_labels(ix, iy, nx, ny, gd)
    P, R, U = nf[:, None, None, None, None], lR[:, None, None, None, None], lU[:, None, None, None, None]
    L_ = np.asarray(fl)[None, :, None, None, None]
    Rr = np.asarray(fr)[None, None, :, None, None]
    Uu = np.asarray(fu)[None, None, None, :, None]
    Dd = np.asarray(fd)[None, None, None, None, :]
    return (Rr == R) & (Uu == U) & (R + U - L_ - Dd == 2 * (P + q))


def random_gauge_peps_arrays(nx, ny, gd, gch, D, seed, complex_entries=False):
    """Seeded random gauge-invariant PEPS at bond dimension D.

    i.i.d. N(0,1) entries from numpy.random.default_rng(seed) (PCG64) drawn
    site by site (iy outer, ix inner) over the full (p,l,r,u,d) array in
    column-major order, Gauss-forbidden entries zeroed, each site normalised;
    all weights 1, bond labels from default_bond_labels.
    """
    rng = np.random.default_rng(seed)
    b0 = boundary_label(gd)
    gch = np.asarray(gch).reshape(nx, ny)
    lh = [[np.ones(D) for _ in range(ny)] for _ in range(nx - 1)]
    lv = [[np.ones(D) for _ in range(ny - 1)] for _ in range(nx)]
    qh = [[default_bond_labels(gd, D) for _ in range(ny)] for _ in range(nx - 1)]
    qv = [[default_bond_labels(gd, D) for _ in range(ny - 1)] for _ in range(nx)]
    tens = [[None] * ny for _ in range(nx)]
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d = site_dims(ix, iy, nx, ny, gd)[0]
            shp = (d, D if ix > 1 else 1, D if ix < nx else 1, D if iy < ny else 1, D if iy > 1 else 1)
            n = int(np.prod(shp))
            v = rng.standard_normal(n)
            if complex_entries:
                v = v + 1j * rng.standard_normal(n)
            T = np.reshape(v.astype(np.complex128 if complex_entries else np.float64), shp, order="F").copy()
            fl = qh[ix - 2][iy - 1] if ix > 1 else [b0]
            fr = qh[ix - 1][iy - 1] if ix < nx else [b0]
            fu = qv[ix - 1][iy - 1] if iy < ny else [b0]
            fd = qv[ix - 1][iy - 2] if iy > 1 else [b0]
            T[~gauss_mask(ix, iy, nx, ny, gd, gch[ix - 1, iy - 1], fl, fr, fu, fd)] = 0
            nrm = np.sqrt(np.sum(np.abs(T) ** 2))
            if nrm > 1e-15:
                T /= nrm
            tens[ix - 1][iy - 1] = T
    return dict(tensors=tens, lh=lh, lv=lv, qh=qh, qv=qv, gch=gch, gd=gd)


def init_gauge_peps_arrays(nx, ny, gd, gch, noise=0.0, seed=None):
    """init_gauge_peps (finite_peps_gauge_invariant.jl:180-218): the bond
    dimension 1 product state in which every link carries the same flux
    label b0 (0 for odd gd, the reference's zero-flux state; +1, i.e.
    E = +1/2, for even gd, which has no zero label) and n_f = -g(ix, iy) on
    every site, which satisfies the Gauss law.  noise > 0 adds noise * z to
    the site tensor (z complex Gaussian of unit variance, the reference's
    randn(ComplexF64), here from numpy.random.default_rng(seed)), re-applies
    the Gauss mask and renormalises each site."""
    gch = np.asarray(gch).reshape(nx, ny)
    b0 = boundary_label(gd)
    lh = [[np.ones(1) for _ in range(ny)] for _ in range(nx - 1)]
    lv = [[np.ones(1) for _ in range(ny - 1)] for _ in range(nx)]
    qh = [[np.array([b0], np.int64) for _ in range(ny)] for _ in range(nx - 1)]
    qv = [[np.array([b0], np.int64) for _ in range(ny - 1)] for _ in range(nx)]
    rng = np.random.default_rng(seed)
    cplx = noise > 0
    tens = [[None] * ny for _ in range(nx)]
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d, dR, dU = site_dims(ix, iy, nx, ny, gd)
            nf0 = -int(gch[ix - 1, iy - 1])
            if not 0 <= nf0 <= 1:
                raise ValueError(f"init_gauge_peps: n_f = {nf0} out of range at ({ix},{iy})")
            T = np.zeros((d, 1, 1, 1, 1), np.complex128 if cplx else np.float64)
            nf, lR, lU = phys_labels(ix, iy, nx, ny, gd)
            T[np.where((nf == nf0) & (lR == b0) & (lU == b0))[0][0], 0, 0, 0, 0] = 1.0
            if cplx:
                T = T + noise * (rng.standard_normal(T.shape) + 1j * rng.standard_normal(T.shape)) / np.sqrt(2.0)
                T[~gauss_mask(ix, iy, nx, ny, gd, gch[ix - 1, iy - 1], [b0], [b0], [b0], [b0])] = 0
                nrm = np.sqrt(np.sum(np.abs(T) ** 2))
                if nrm > 1e-15:
                    T /= nrm
            tens[ix - 1][iy - 1] = T
    return dict(tensors=tens, lh=lh, lv=lv, qh=qh, qv=qv, gch=gch, gd=gd)


def init_finite_peps_v2_arrays(nx, ny, gd, noise=0.0, seed=None):
    """init_finite_peps_v2 (finite_peps_full_update.jl:29-62): the label-free
    bond dimension 1 product state with n_f = 0 on even and 1 on odd ix+iy
    and links at the flux label b0 (as init_gauge_peps_arrays), plus
    noise * z on every entry (no mask), each site normalised."""
    b0 = boundary_label(gd)
    lh = [[np.ones(1) for _ in range(ny)] for _ in range(nx - 1)]
    lv = [[np.ones(1) for _ in range(ny - 1)] for _ in range(nx)]
    qh = [[np.array([b0], np.int64) for _ in range(ny)] for _ in range(nx - 1)]
    qv = [[np.array([b0], np.int64) for _ in range(ny - 1)] for _ in range(nx)]
    rng = np.random.default_rng(seed)
    cplx = noise > 0
    tens = [[None] * ny for _ in range(nx)]
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d, dR, dU = site_dims(ix, iy, nx, ny, gd)
            nf0 = 0 if (ix + iy) % 2 == 0 else 1
            T = np.zeros((d, 1, 1, 1, 1), np.complex128 if cplx else np.float64)
            nf, lR, lU = phys_labels(ix, iy, nx, ny, gd)
            T[np.where((nf == nf0) & (lR == b0) & (lU == b0))[0][0], 0, 0, 0, 0] = 1.0
            if cplx:
                T = T + noise * (rng.standard_normal(T.shape) + 1j * rng.standard_normal(T.shape)) / np.sqrt(2.0)
            nrm = np.sqrt(np.sum(np.abs(T) ** 2))
            if nrm > 1e-15:
                T /= nrm
            tens[ix - 1][iy - 1] = T
    return dict(tensors=tens, lh=lh, lv=lv, qh=qh, qv=qv, gch=staggered_charges(nx, ny), gd=gd)


def init_finite_peps_arrays(nx, ny, gd, D, noise=0.0, seed=None, nf_of=None):
    """init_finite_peps (finite_peps_ground_state.jl:122-160): the label-free
    FinitePEPS at bond dimension D (1 on the open edges) with all weights 1
    and the product state |n_f, E_R = 0, E_U = 0> in the [., 1, 1, 1, 1]
    corner of every site tensor, n_f = 0 on even and 1 on odd ix + iy (or
    nf_of(ix, iy)); noise > 0 adds noise * z to every entry (z complex
    Gaussian of unit variance, the reference's randn(ComplexF64), here from
    numpy.random.default_rng(seed)) without renormalising.  The zero field
    E = 0 needs odd gd (the reference's dg = (gd-1)/2).  Entries are complex,
    as the reference's ComplexF64 tensors; the flux labels are placeholders
    (a FinitePEPS carries none)."""
    if gd % 2 == 0:
        raise ValueError("init_finite_peps: the E = 0 product state needs an odd link dimension")
    rng = np.random.default_rng(seed)
    lh = [[np.ones(D) for _ in range(ny)] for _ in range(nx - 1)]
    lv = [[np.ones(D) for _ in range(ny - 1)] for _ in range(nx)]
    qh = [[np.zeros(D, np.int64) for _ in range(ny)] for _ in range(nx - 1)]
    qv = [[np.zeros(D, np.int64) for _ in range(ny - 1)] for _ in range(nx)]
    tens = [[None] * ny for _ in range(nx)]
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d, dR, dU = site_dims(ix, iy, nx, ny, gd)
            shp = (d, D if ix > 1 else 1, D if ix < nx else 1, D if iy < ny else 1, D if iy > 1 else 1)
            T = np.zeros(shp, np.complex128)
            nf0 = (0 if (ix + iy) % 2 == 0 else 1) if nf_of is None else int(nf_of(ix, iy))
            nf, lR, lU = phys_labels(ix, iy, nx, ny, gd)
            T[np.where((nf == nf0) & (lR == 0) & (lU == 0))[0][0], 0, 0, 0, 0] = 1.0
            if noise > 0:
                T = T + noise * (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)) / np.sqrt(2.0)
            tens[ix - 1][iy - 1] = T
    return dict(tensors=tens, lh=lh, lv=lv, qh=qh, qv=qv, gch=staggered_charges(nx, ny), gd=gd)


def site_operators(ix, iy, nx, ny, gd):
    """The single-site operators the quench protocols use, embedded in the
    site's physical space (fermion slowest, then the right link, the up link
    fastest: embed_f_site / embed_R_site / embed_U_site,
    u1_lgt_hamiltonian.jl:629-643): n_f, c^dagger (op_cdag, :147), U on the
    right link and U on the up link (op_U_gauge, U|l> = |l+2> in label units;
    None for a missing link), as complex (d, d) arrays."""
    d, dR, dU = site_dims(ix, iy, nx, ny, gd)
    c = np.zeros((2, 2))
    c[0, 1] = 1.0
    nfo = np.diag([0.0, 1.0])
    U = np.zeros((gd, gd))
    for k in range(gd - 1):
        U[k + 1, k] = 1.0
    I2, IR, IU = np.eye(2), np.eye(dR), np.eye(dU)
    out = dict(nf=np.kron(np.kron(nfo, IR), IU), cdag=np.kron(np.kron(c.T, IR), IU),
               U_R=np.kron(np.kron(I2, U), IU) if dR > 1 else None,
               U_U=np.kron(np.kron(I2, IR), U) if dU > 1 else None)
    return {k: (None if v is None else v.astype(np.complex128)) for k, v in out.items()}
