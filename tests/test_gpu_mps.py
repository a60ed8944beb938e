"""The snake-MPS / DMRG engine on the GPU (lgt_amd/mps.py over peps_dmrg and
peps_mpo_expect; mps_lgt.jl) against the oracle's full-space Hamiltonian and
the ED oracle (finite_ed.jl), as the reference's own self-test does
(selftest_mps_lgt, mps_lgt.jl:680-712: g = t = 1, m = 0.5, Lambda = 20,
D = 60, 8 sweeps; the dense check at 2x2, the Gauss diagnostic at 2x3)."""
import numpy as np
import pytest

from oracle import ed_oracle as ED
from oracle import mps_oracle as M
from oracle import peps_oracle as P

pytestmark = pytest.mark.gpu

G_, T_, M_, LAM = 1.0, 1.0, 0.5, 20.0


def _ed(nx, ny, gch):
    b = ED.build_basis(nx, ny, 1, gch)
    E0, psi = ED.find_ground_state(ED.build_hamiltonian(b, G_, T_, M_))
    return E0, ED.measure_ed(psi, b)


def test_mpo_assembly_matches_full_hamiltonian(ctx):
    """build_H_mpo and build_penalized_H (term MPOs summed and compressed
    with the library's SVD) densify to the oracle's full-space Hamiltonian,
    and to it plus Lambda sum_i ((G_i - g_i)^2 - g_i^2), at 2x2, gd = 3."""
    from lgt_amd import mps
    nx, ny, gd = 2, 2, 3
    print("mpo assembly 2x2 gd=3", flush=True)
    ref = M.full_hamiltonian(nx, ny, gd, G_, T_, M_)
    H = mps.build_H_mpo(nx, ny, gd, G_, T_, M_, ctx=ctx)
    assert np.abs(M.mpo_to_dense(H) - ref).max() < 1e-12 * np.abs(ref).max()
    gch = np.asarray(P.staggered_charges(nx, ny))
    Hp = mps.build_penalized_H(nx, ny, gd, G_, T_, M_, gch, LAM, ctx=ctx)
    G = M.gauss_values(nx, ny, gd)
    want = ref + np.diag(LAM * (((G - gch[None]) ** 2).sum(axis=(1, 2)) - (gch ** 2).sum()))
    assert np.abs(M.mpo_to_dense(Hp) - want).max() < 1e-11 * np.abs(want).max()
    # compressed: the bare MPO's bonds stay far below the 12 summed terms
    assert max(w.shape[1] for w in H) <= 8


def test_dmrg_2x2_dense_and_ed(ctx):
    """The reference's dense gate: DMRG on the penalized MPO reaches its
    lowest dense eigenvalue; the state's bare energy is the sector's ED
    ground energy, it violates Gauss's law by < 1e-8, and its n_f and E^2
    grids are ED's."""
    from lgt_amd import mps
    nx, ny, gd = 2, 2, 3
    print("dmrg 2x2 gd=3", flush=True)
    gch = np.asarray(P.staggered_charges(nx, ny))
    dims = mps.node_dims(nx, ny, gd)
    Hp = mps.build_penalized_H(nx, ny, gd, G_, T_, M_, gch, LAM, ctx=ctx)
    E, psi, sweeps = mps.dmrg_ground_state(Hp, dims, D=60, nsweeps=8, seed=1, ctx=ctx)
    E_dense = np.linalg.eigvalsh(M.mpo_to_dense(Hp))[0]
    assert abs(E - E_dense) < 1e-8, (E, E_dense, sweeps)
    E0, obs = _ed(nx, ny, gch)
    Hb = mps.build_H_mpo(nx, ny, gd, G_, T_, M_, ctx=ctx)
    Eb = mps.peps.mpo_expect(Hb, psi, ctx=ctx) / mps.mps_overlap(psi, ctx=ctx)
    assert abs(Eb - E0) < 1e-8, (Eb, E0)
    assert mps.gauss_violation(psi, nx, ny, gd, gch, ctx=ctx) < 1e-8
    got = mps.measure_mps(psi, nx, ny, gd, ctx=ctx)
    assert np.abs(got["nf_grid"] - obs["nf_grid"]).max() < 1e-7
    assert np.abs(got["E2_grid"] - obs["E2_grid"]).max() < 1e-7
    assert abs(got["nf_even"] - obs["nf_even"]) < 1e-7 and abs(got["nf_odd"] - obs["nf_odd"]) < 1e-7


def test_dmrg_2x3_ed(ctx):
    """2x3 (1.4e5 product states, too many for the dense check): from a
    seeded random MPS (the reference's default start), DMRG's bare energy is
    the ED ground energy of the staggered sector to 1e-7, the Gauss violation
    is < 1e-7 and the observables are ED's to 1e-6.  (From the staggered
    product state two-site updates cannot move charge along the vertical
    bonds, which are three chain sites apart, and stall above the ground
    state.)"""
    from lgt_amd import mps
    nx, ny, gd = 2, 3, 3
    print("dmrg 2x3 gd=3", flush=True)
    gch = np.asarray(P.staggered_charges(nx, ny))
    dims = mps.node_dims(nx, ny, gd)
    Hp = mps.build_penalized_H(nx, ny, gd, G_, T_, M_, gch, LAM, ctx=ctx)
    E, psi, sweeps = mps.dmrg_ground_state(Hp, dims, D=60, nsweeps=8, seed=1, ctx=ctx)
    E0, obs = _ed(nx, ny, gch)
    Hb = mps.build_H_mpo(nx, ny, gd, G_, T_, M_, ctx=ctx)
    Eb = mps.peps.mpo_expect(Hb, psi, ctx=ctx) / mps.mps_overlap(psi, ctx=ctx)
    assert abs(Eb - E0) < 1e-7, (Eb, E0, E, sweeps)
    assert abs(E - (E0 - LAM * float((gch ** 2).sum()))) < 1e-7
    assert mps.gauss_violation(psi, nx, ny, gd, gch, ctx=ctx) < 1e-7
    got = mps.measure_mps(psi, nx, ny, gd, ctx=ctx)
    assert np.abs(got["nf_grid"] - obs["nf_grid"]).max() < 1e-6
    assert np.abs(got["E2_grid"] - obs["E2_grid"]).max() < 1e-6


def test_dmrg_refuses_bad_input(ctx):
    """Open boundaries, matching bonds and real tensors are required."""
    from lgt_amd import mps, peps
    W = mps.identity_mpo([2, 2])
    psi = [np.ones((1, 1, 2)), np.ones((1, 1, 2))]
    with pytest.raises(peps.PepsError):
        peps.dmrg(W, [np.ones((1, 2, 2)), np.ones((1, 1, 2))], 4, ctx=ctx)   # bond mismatch
    with pytest.raises(peps.PepsError):
        peps.dmrg(W[:1], psi[:1], 4, ctx=ctx)                                  # one site
    assert abs(peps.mpo_expect(W, psi, ctx=ctx) - 4.0) < 1e-15


def test_gs_benchmark_mps_task_row(tmp_path):
    """run_one_mps (gs_benchmark_mps.jl:41-96) at 2x3 from a random start:
    the CSV row carries the reference's columns, the DMRG energy is ED's to
    1e-7 and the differences are consistent."""
    from lgt_amd import gs_benchmark as gb, peps
    print("gs_benchmark_mps task 7 at 2x3", flush=True)
    ctx = peps.Context(0)
    try:
        row = gb.run_one_mps(7, str(tmp_path), ctx=ctx, nx=2, ny=3, D=60, start="random")
    finally:
        ctx.close()
    lines = (tmp_path / "gs_mps_task7.csv").read_text().splitlines()
    assert lines[0] == ",".join(gb.MPS_COLUMNS) and len(lines) == 2
    cells = dict(zip(gb.MPS_COLUMNS, lines[1].split(",")))
    assert int(cells["task"]) == 7 and float(cells["m"]) == 0.5 and float(cells["g"]) == 0.5
    assert row["dE_rel"] < 1e-7 and row["gauss_viol"] < 1e-6 and 1 <= row["bond"] <= 60
    assert row["dE_rel"] == pytest.approx(abs(row["E_mps"] - row["E_ed"]) / abs(row["E_ed"]), rel=1e-12)
    assert abs(row["nf_even_mps"] - row["nf_even_ed"]) < 1e-6


def _dense(psi):
    """The full-space vector of an MPS (first chain site slowest)."""
    T = np.asarray(psi[0])[0].T                       # (d, Dr)
    for A in psi[1:]:
        T = np.einsum("ir,rsk->iks", T, np.asarray(A)).reshape(-1, A.shape[1])
    return T[:, 0]


def test_mps_local_expect_and_complex_mpo_expect(ctx):
    """mps_local_expect (mps_lgt.jl:317-330, one pass of environments for
    many operators) and mpo_expect of a complex MPS against the dense
    vector of a random complex MPS."""
    from lgt_amd import mps, peps
    rng = np.random.default_rng(11)
    dims = [3, 4, 2, 3]
    psi = [a + 1j * rng.standard_normal(a.shape) for a in mps.random_mps(dims, 5, 4)]
    v = _dense(psi)
    ops = [rng.standard_normal((d, d)) for d in dims]
    got = peps.mps_local_expect_many(psi, list(range(4)), ops, ctx=ctx)
    for p, O in enumerate(ops):
        want = np.real(np.vdot(v, M._apply_local(dims, {p: O}) @ v)) / np.vdot(v, v).real
        assert abs(got[p] - want) < 1e-12 * max(1.0, abs(want))
    W = mps.term_mpo(dims, (0.7, {1: ops[1], 3: ops[3]}))
    want = np.real(np.vdot(v, M.mpo_to_dense(W) @ v))
    assert abs(peps.mpo_expect(W, psi, ctx=ctx) - want) < 1e-11 * max(1.0, abs(want))


def test_krylov_evolve_matches_dense_krylov_2x2(ctx):
    """Quench A at 2x2 (string state, g = 1, m = 2, dt = 0.05): ten
    krylov_evolve steps on the GPU (m = 8, no bond cap reached) give the
    dense Krylov recurrence's state to 1e-9 and the exact exp(-i t H) psi
    to 1e-8; one call of ten steps equals ten calls of one; the energy is
    conserved."""
    from lgt_amd import mps, peps
    nx, ny, gd = 2, 2, 3
    print("krylov evolve 2x2 string state", flush=True)
    H = mps.build_H_mpo(nx, ny, gd, 1.0, 1.0, 2.0, ctx=ctx)
    psi0 = mps.string_breaking_mps(nx, ny, gd)
    dims = mps.node_dims(nx, ny, gd)
    cfg = [int(np.argmax(a.reshape(-1))) + 1 for a in psi0]
    Hd = M.full_hamiltonian(nx, ny, gd, 1.0, 1.0, 2.0)
    w, Q = np.linalg.eigh(Hd)
    x = M.product_vector(cfg, dims)
    y = x.copy()
    psi = psi0
    E0 = peps.mpo_expect(H, psi0, ctx=ctx)
    for step in range(10):
        psi = peps.krylov_evolve(H, psi, 0.05, nsteps=1, m=8, Dmax=200, ctx=ctx)
        x = M.krylov_evolve_dense(Hd, x, 0.05, 8)
        y = Q @ (np.exp(-1j * 0.05 * w) * (Q.conj().T @ y))
        v = _dense(psi)
        assert np.abs(v - x).max() < 1e-9, step
        assert np.abs(v - y).max() < 1e-8, step
        assert abs(peps.mpo_expect(H, psi, ctx=ctx) - E0) < 1e-8
    psi10 = peps.krylov_evolve(H, psi0, 0.05, nsteps=10, m=8, Dmax=200, ctx=ctx)
    assert np.abs(_dense(psi10) - _dense(psi)).max() < 1e-12
    assert all(np.iscomplexobj(a) for a in psi10)


def test_quench_B_2x3_against_ed():
    """Quench B at 2x3 (mps_quench.jl:89-103 at a smaller lattice): the
    penalised-DMRG ground state at m = 5 evolved by evolve_mps with m = 0.1
    (dt = 0.05, eight steps) reproduces the ED evolution of the ED ground
    state (finite_ed.jl:404-430): n_f and E^2 on every site to 1e-6 at every
    step."""
    from lgt_amd import mps, mps_quench, peps
    nx, ny, gd = 2, 3, 3
    print("quench B 2x3 against ED", flush=True)
    ctx = peps.Context(0)
    try:
        gch = np.asarray(P.staggered_charges(nx, ny))
        dims = mps.node_dims(nx, ny, gd)
        Hp = mps.build_penalized_H(nx, ny, gd, 1.0, 1.0, 5.0, gch, 5.0, ctx=ctx)
        _, psi0, _ = mps.dmrg_ground_state(Hp, dims, D=60, nsweeps=8, seed=1, ctx=ctx)
        H = mps.build_H_mpo(nx, ny, gd, 1.0, 1.0, 0.1, ctx=ctx)
        data, psi = mps_quench.evolve_mps(psi0, H, nx, ny, gd, 0.05, 8, 120, ctx=ctx)
    finally:
        ctx.close()
    b = ED.build_basis(nx, ny, 1, gch)
    _, g0 = ED.find_ground_state(ED.build_hamiltonian(b, 1.0, 1.0, 5.0))
    states = [g0] + ED.evolve_ed(g0, ED.build_hamiltonian(b, 1.0, 1.0, 0.1), 0.05, 8)
    assert len(data) == 9
    for step, (d, s) in enumerate(zip(data, states)):
        obs = ED.measure_ed(s, b)
        assert np.abs(d["nf_grid"] - obs["nf_grid"]).max() < 1e-6, step
        assert np.abs(d["E2_grid"] - obs["E2_grid"]).max() < 1e-6, step
        assert d["t"] == pytest.approx(0.05 * step)
    # the evolution moved the state (the total n_f is conserved, the site
    # occupations and fields are not): far more than the tolerance
    moved = max(np.abs(data[-1]["nf_grid"] - data[0]["nf_grid"]).max(),
                np.abs(data[-1]["E2_grid"] - data[0]["E2_grid"]).max())
    assert moved > 1e-3, moved


def test_krylov_evolve_refuses_bad_input(ctx):
    from lgt_amd import mps, peps
    W = mps.identity_mpo([2, 2])
    psi = [np.ones((1, 1, 2)), np.ones((1, 1, 2))]
    with pytest.raises(peps.PepsError):
        peps.krylov_evolve(W, [np.ones((1, 2, 2)), np.ones((1, 1, 2))], 0.1, ctx=ctx)  # bond mismatch
    with pytest.raises(peps.PepsError):
        peps.krylov_evolve(W, psi, 0.1, Dmax=0, ctx=ctx)
    out = peps.krylov_evolve(W, psi, 0.1, nsteps=0, ctx=ctx)
    assert np.allclose(_dense(out), _dense(psi))
