"""FP64 parity of the dense simple update with plaquettes
(finite_peps_ground_state.jl) on the GPU with the CPU oracle.  Singular
vectors carry gauge freedom (signs, phases), so states are compared through
the contracted dense state and the bond weights; plaquette gate splits
through their reconstruction.
"""
import numpy as np
import pytest

from oracle import peps_oracle as O
from tests._util import random_peps, rel, to_device, to_oracle

pytestmark = pytest.mark.gpu

G, T_HOP, M = 1.0, 1.0, 0.5


def _bond_dims(gp):
    return ([len(gp.LH(ix, iy)) for iy in range(1, gp.ny + 1) for ix in range(1, gp.nx)],
            [len(gp.LV(ix, iy)) for iy in range(1, gp.ny) for ix in range(1, gp.nx + 1)])


def _compare(gd_dev, go, tol_psi, tol_lam):
    gg = to_oracle(gd_dev)
    assert _bond_dims(gg) == _bond_dims(go)
    for iy in range(1, go.ny + 1):
        for ix in range(1, go.nx):
            assert np.abs(np.asarray(gg.LH(ix, iy)) - go.LH(ix, iy)).max() < tol_lam
    for iy in range(1, go.ny):
        for ix in range(1, go.nx + 1):
            assert np.abs(np.asarray(gg.LV(ix, iy)) - go.LV(ix, iy)).max() < tol_lam
    assert rel(O.dense_psi(gg), O.dense_psi(go)) < tol_psi


@pytest.mark.parametrize("nx,ny,gd,D,Dtr,cplx", [(3, 3, 2, 2, 3, False), (3, 3, 2, 3, 2, False),
                                                (2, 3, 3, 2, 4, True)])
def test_dense_bond_updates(ctx, nx, ny, gd, D, Dtr, cplx):
    from lgt_amd import peps
    print(f"dense bond updates {nx}x{ny} gd={gd} D={D} Dtr={Dtr} cplx={cplx}", flush=True)
    go = random_peps(nx, ny, gd, D, 17, cplx)
    gdv = to_device(go, ctx)
    rng = np.random.default_rng(5)
    bonds = [("h", ix, iy) for iy in range(1, ny + 1) for ix in range(1, nx)] + \
            [("v", ix, iy) for iy in range(1, ny) for ix in range(1, nx + 1)]
    for o, ix, iy in bonds:
        if o == "h":
            Gt = O.gate(O.H_merged_h_site(ix, iy, nx, ny, gd, G, T_HOP, M), 0.05)
        else:
            Gt = O.gate(O.H_merged_v_site(ix, iy, nx, ny, gd, G, T_HOP, M), 0.05)
        if cplx:  # a generic complex gate: a conjugated or transposed gate fails
            Gt = Gt + 0.1j * rng.standard_normal(Gt.shape)
        (O.update_bond_h if o == "h" else O.update_bond_v)(go, ix, iy, Gt, Dtr)
        (peps.update_bond_h if o == "h" else peps.update_bond_v)(gdv, ix, iy, Gt, Dtr)
    _compare(gdv, go, 1e-10, 1e-12)


@pytest.mark.parametrize("gd,tau", [(2, 0.05), (3, 0.1)])
def test_plaquette_gate_split(ctx, gd, tau):
    from lgt_amd import peps
    Gp = O.gate(O.H_plaquette_gauge(gd, G), tau)
    GA, GB, GC = peps.plaquette_gate(gd, G, tau, ctx=ctx)
    oA, oB, oC = O.decompose_plaquette_LShape(Gp, gd, cutoff=O.SVD_CUT)
    assert GB.shape == oB.shape and GC.shape == oC.shape and GA.shape == oA.shape
    assert np.abs(O.reconstruct_plaquette_LShape(GA, GB, GC) - Gp).max() < 1e-12


@pytest.mark.parametrize("nx,ny,gd,D,plaq,cplx", [(3, 3, 2, 2, [(1, 1), (2, 2)], False),
                                                 (3, 2, 2, 2, [(2, 1)], True),
                                                 (2, 2, 3, 2, [(1, 1)], False)])
def test_update_plaquette(ctx, nx, ny, gd, D, plaq, cplx):
    from lgt_amd import peps
    print(f"update_plaquette {nx}x{ny} gd={gd} D={D} at {plaq} cplx={cplx}", flush=True)
    go = random_peps(nx, ny, gd, D, 23, cplx)
    gdv = to_device(go, ctx)
    GA, GB, GC = O.decompose_plaquette_LShape(O.gate(O.H_plaquette_gauge(gd, G), 0.1), gd,
                                              cutoff=O.SVD_CUT)
    for ix, iy in plaq:
        O.update_plaquette(go, ix, iy, GA, GB, GC, 3)
        peps.update_plaquette(gdv, ix, iy, GA, GB, GC, 3)
    _compare(gdv, go, 1e-10, 1e-12)


def test_dense_trotter_step(ctx):
    from lgt_amd import peps
    go = random_peps(3, 3, 2, 2, 31)
    gdv = to_device(go, ctx)
    for _ in range(2):
        O.trotter_step_dense(go, G, T_HOP, M, 0.05, 3)
        peps.trotter_step(gdv, G, T_HOP, M, 0.05, 3)
    _compare(gdv, go, 1e-9, 1e-10)
    E_o = O.total_energy(go, G, T_HOP, M)
    E_g = peps.total_energy(gdv, G, T_HOP, M)
    print(f"total_energy gpu {E_g:.15f} oracle {E_o:.15f}", flush=True)
    assert abs(E_g - E_o) < 1e-10 * max(1.0, abs(E_o))


@pytest.mark.parametrize("nx,ny,gd,D,cplx", [(3, 3, 2, 3, False), (2, 3, 3, 2, True)])
def test_su_observables(ctx, nx, ny, gd, D, cplx):
    from lgt_amd import peps
    go = random_peps(nx, ny, gd, D, 41, cplx)
    gdv = to_device(go, ctx)
    rng = np.random.default_rng(8)
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            dp = O.site_dims(ix, iy, nx, ny, gd)[0]
            Op = rng.standard_normal((dp, dp)) + (1j * rng.standard_normal((dp, dp)) if cplx else 0)
            e_o = O.expect_site(go, ix, iy, Op)
            assert abs(peps.expect_site(gdv, ix, iy, Op) - e_o) < 1e-12 * max(1.0, abs(e_o))
    for iy in range(1, ny):
        for ix in range(1, nx):
            e_o = O.plaquette_energy(go, ix, iy, 1.3)
            assert abs(peps.plaquette_energy(gdv, ix, iy, 1.3) - e_o) < 1e-12 * max(1.0, abs(e_o))
    E_o = O.total_energy(go, G, T_HOP, M)
    assert abs(peps.total_energy(gdv, G, T_HOP, M) - E_o) < 1e-11 * max(1.0, abs(E_o))


def test_dense_update_drops_flux_labels(ctx):
    from lgt_amd import peps
    go = random_peps(2, 2, 2, 2, 3)
    gdv = to_device(go, ctx)
    peps.forbidden_norm(gdv)  # labelled state: allowed
    Gt = O.gate(O.H_merged_h_site(1, 1, 2, 2, 2, G, T_HOP, M), 0.05)
    peps.update_bond_h(gdv, 1, 1, Gt, 2)
    with pytest.raises(peps.PepsError) as e:
        peps.forbidden_norm(gdv)
    assert e.value.code == 1
    with pytest.raises(peps.PepsError):
        peps.trotter_step_gauge(gdv, G, T_HOP, M, 0.05, 2)


@pytest.mark.parametrize("cplx", [False, True])
def test_full_update_v2_bonds_and_step(ctx, cplx):
    """full_update_bond_h!/v! (finite_peps_full_update.jl): bond by bond, then
    two trotter_step_full! steps."""
    from lgt_amd import peps
    nx, ny, gd = (3, 3, 2) if not cplx else (2, 3, 3)
    print(f"v2 full update {nx}x{ny} gd={gd} cplx={cplx}", flush=True)
    go = random_peps(nx, ny, gd, 2, 19, cplx)
    go.lh[0][0] = np.array([1.0, 3e-11])  # below the 1e-10 cut after the update
    gdv = to_device(go, ctx)
    rng = np.random.default_rng(4)
    for o, ix, iy in [("h", 1, 1), ("v", 1, 1), ("h", nx - 1, ny), ("v", nx, ny - 1)]:
        Hb = (O.H_merged_h_site if o == "h" else O.H_merged_v_site)(ix, iy, nx, ny, gd, G, T_HOP, M)
        Gt = O.gate(Hb, 0.05)
        if cplx:
            Gt = Gt + 0.1j * rng.standard_normal(Gt.shape)
        (O.full_update_bond_h_v2 if o == "h" else O.full_update_bond_v_v2)(go, ix, iy, Gt, 3)
        (peps.full_update_bond_h if o == "h" else peps.full_update_bond_v)(gdv, ix, iy, Gt, 3)
    _compare(gdv, go, 1e-10, 1e-12)
    for _ in range(2):
        O.trotter_step_full(go, G, T_HOP, M, 0.05, 3)
        peps.trotter_step_full(gdv, G, T_HOP, M, 0.05, 3)
    _compare(gdv, go, 1e-9, 1e-10)
    with pytest.raises(peps.PepsError) as e:
        peps.full_update_bond_h(gdv, 1, 1, np.eye(O.site_dims(1, 1, nx, ny, gd)[0] * O.site_dims(2, 1, nx, ny, gd)[0]),
                                3, use_gauss=True)
    assert e.value.code == 7


@pytest.mark.parametrize("cplx", [False, True])
def test_measure_all_finite(ctx, cplx):
    from lgt_amd import peps
    nx, ny, gd = (3, 3, 2) if not cplx else (2, 3, 3)
    go = random_peps(nx, ny, gd, 3, 51, cplx)
    gdv = to_device(go, ctx)
    mo, mg = O.measure_all_finite(go, 0.25), peps.measure_all_finite(gdv, 0.25)
    for k in ("nf_mean", "nf_even", "nf_odd", "E2_mean", "S_h_mean", "S_v_mean", "S_mean"):
        assert abs(mo[k] - mg[k]) < 1e-12 * max(1.0, abs(mo[k])), k
    np.testing.assert_allclose(mg["nf_grid"], mo["nf_grid"], atol=1e-12)
    np.testing.assert_allclose(mg["E2_grid"], mo["E2_grid"], atol=1e-12)
    lam = go.LH(1, 1)
    assert abs(peps.bond_entropy(lam) - O.bond_entropy(lam)) < 1e-14


@pytest.mark.parametrize("nx,ny,gd,D,chi,cplx", [(3, 3, 2, 2, 64, False), (3, 3, 2, 2, 3, False),
                                                (3, 4, 2, 3, 12, False), (2, 3, 3, 2, 64, True)])
def test_plaquette_expect(ctx, nx, ny, gd, D, chi, cplx):
    """Boundary-MPS <W> on every plaquette against the oracle (exact and
    truncated chi)."""
    from lgt_amd import peps
    go = random_peps(nx, ny, gd, D, 61, cplx)
    gdv = to_device(go, ctx)
    for iy in range(1, ny):
        for ix in range(1, nx):
            wo = O.plaquette_expect(go, ix, iy, chi)
            wg = peps.plaquette_expect(gdv, ix, iy, chi)
            assert abs(wg - wo) < 1e-10 * max(1.0, abs(wo)), (ix, iy, wg, wo)


def test_batch_trotter_step_matches_sequential(ctx):
    """Independent instances stepped concurrently (one context and host
    thread each) give the same states as stepping them one by one; states
    sharing a context are refused."""
    import time
    from lgt_amd import peps
    from tests._util import su_state
    specs = [(3, 3, 2, 2, 11), (3, 4, 2, 3, 3), (4, 3, 2, 2, 7), (3, 3, 2, 3, 5)]
    olds = [su_state(nx, ny, gd, D, 3, seed) for nx, ny, gd, D, seed in specs]
    ctxs = [peps.Context(0) for _ in specs]
    batch = [to_device(go, c) for go, c in zip(olds, ctxs)]
    seq = [to_device(go, ctx) for go in olds]
    t0 = time.perf_counter()
    st = peps.trotter_step_batch(batch, peps.PEPS_STEP_V4, G, T_HOP, M, 0.05, 3, 12, n_als=2)
    t_batch = time.perf_counter() - t0
    assert list(st) == [0] * len(specs)
    t0 = time.perf_counter()
    for gp in seq:
        peps.trotter_step_v4(gp, G, T_HOP, M, 0.05, 3, 12, n_als=2)
    t_seq = time.perf_counter() - t0
    print(f"batch of {len(specs)}: {t_batch:.3f} s, one by one: {t_seq:.3f} s", flush=True)
    for gb, gs in zip(batch, seq):
        a, b = to_oracle(gb), to_oracle(gs)
        assert _bond_dims(a) == _bond_dims(b)
        for iy in range(1, a.ny + 1):
            for ix in range(1, a.nx + 1):
                assert rel(a.T(ix, iy), b.T(ix, iy)) < 1e-12
    with pytest.raises(peps.PepsError):
        peps.trotter_step_batch([seq[0], seq[1]], peps.PEPS_STEP_V4, G, T_HOP, M, 0.05, 3, 12, n_als=2)
