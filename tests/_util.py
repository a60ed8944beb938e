"""Shared helpers for the GPU parity tests: move states between the oracle
and the device library."""
import numpy as np

from oracle import peps_oracle as O


def to_device(gp_o, ctx):
    from lgt_amd import peps
    cp = any(np.iscomplexobj(t) and np.any(t.imag != 0) for col in gp_o.tensors for t in col)
    return peps.GaugePEPS.from_arrays(gp_o.tensors, gp_o.lh, gp_o.lv, gp_o.qh, gp_o.qv, gp_o.gch,
                                      gp_o.gd, is_complex=cp, ctx=ctx)


def to_oracle(gp_d):
    a = gp_d.to_arrays()
    g = O.GaugePEPS(gp_d.nx, gp_d.ny, gp_d.gd, gp_d.gch)
    g.tensors = [[np.asarray(t, dtype=np.complex128) for t in col] for col in a["tensors"]]
    g.lh, g.lv, g.qh, g.qv = a["lh"], a["lv"], a["qh"], a["qv"]
    return g


def su_state(nx, ny, gd, D, steps, seed, cplx=False, tau=0.05, g=1.0, t=1.0, m=0.5):
    gp = O.init_gauge_peps_random(nx, ny, gd, O.staggered_charges(nx, ny), D, seed, complex_entries=cplx)
    for _ in range(steps):
        O.trotter_step_gauge(gp, g, t, m, tau, D)
    return gp


def random_peps(nx, ny, gd, D, seed, cplx=False):
    """Unmasked random FinitePEPS (bond labels all 0, unused) with random
    positive decreasing weights."""
    rng = np.random.default_rng(seed)
    gp = O.GaugePEPS(nx, ny, gd, O.staggered_charges(nx, ny))
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d = O.site_dims(ix, iy, nx, ny, gd)[0]
            dims = (d, D if ix > 1 else 1, D if ix < nx else 1, D if iy < ny else 1, D if iy > 1 else 1)
            T = rng.standard_normal(dims)
            if cplx:
                T = T + 1j * rng.standard_normal(dims)
            gp.setT(ix, iy, T.astype(np.complex128))
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            gp.lh[ix - 1][iy - 1] = np.sort(rng.uniform(0.2, 1.0, D))[::-1]
            gp.qh[ix - 1][iy - 1] = np.zeros(D, np.int64)
    for iy in range(1, ny):
        for ix in range(1, nx + 1):
            gp.lv[ix - 1][iy - 1] = np.sort(rng.uniform(0.2, 1.0, D))[::-1]
            gp.qv[ix - 1][iy - 1] = np.zeros(D, np.int64)
    return gp


def mps_dense(mps):
    v = mps[0].reshape(mps[0].shape[0] * mps[0].shape[1], mps[0].shape[2], order="F")
    for A in mps[1:]:
        v = (v @ A.reshape(A.shape[0], -1, order="F")).reshape(-1, A.shape[2], order="F")
    return v[:, 0]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def fixture_arrays(z, prefix, nx, ny):
    """Tensors, weights and labels stored by tests/golden/make_config_fixtures.py."""
    tens = [[z[f"{prefix}T_{ix}_{iy}"] for iy in range(1, ny + 1)] for ix in range(1, nx + 1)]
    lh = [[z[f"{prefix}lh_{ix}_{iy}"] for iy in range(1, ny + 1)] for ix in range(1, nx)]
    qh = [[z[f"{prefix}qh_{ix}_{iy}"] for iy in range(1, ny + 1)] for ix in range(1, nx)]
    lv = [[z[f"{prefix}lv_{ix}_{iy}"] for iy in range(1, ny)] for ix in range(1, nx + 1)]
    qv = [[z[f"{prefix}qv_{ix}_{iy}"] for iy in range(1, ny)] for ix in range(1, nx + 1)]
    return dict(tensors=tens, lh=lh, lv=lv, qh=qh, qv=qv)


def fixture_state(z, prefix, nx, ny, gd, ctx):
    from lgt_amd import lattice, peps
    a = fixture_arrays(z, prefix, nx, ny)
    return peps.GaugePEPS.from_arrays(a["tensors"], a["lh"], a["lv"], a["qh"], a["qv"],
                                      lattice.staggered_charges(nx, ny), gd, ctx=ctx)
