"""The GPU exact diagonalisation (finite_ed.jl) against the ED oracle.

Basis keys bit-exact and in the same order; H x to 1e-14; the ground-state
energy to 1e-10 and the state to an overlap of 1 - 1e-9; the observables to
1e-12 on a given state.  At 4x4 (beyond the oracle's reach in a test) the
size-independent properties: H symmetric, the particle number, the ground
state's residual and its energy below every diagonal entry."""
import numpy as np
import pytest

from oracle import ed_oracle as ED
from oracle import peps_oracle as P

pytestmark = pytest.mark.gpu


def staggered(nx, ny):
    return np.array([[-1 if (ix + iy) % 2 else 0 for iy in range(1, ny + 1)] for ix in range(1, nx + 1)])


def string(nx, ny):  # finite_ed.jl:17-19, quench A
    g = np.zeros((nx, ny), int)
    g[0, 1], g[1, 1] = 1, -2
    return g


SECTORS = [(3, 4, 1, "stag"), (3, 4, 1, "string"), (2, 2, 2, "stag"), (2, 3, 1, "stag"), (4, 2, 3, "stag")]


def charges(nx, ny, kind):
    return staggered(nx, ny) if kind == "stag" else string(nx, ny)


@pytest.mark.parametrize("nx,ny,dg,kind", SECTORS)
def test_basis_keys_and_order(ctx, nx, ny, dg, kind):
    from lgt_amd import peps
    g = charges(nx, ny, kind)
    ed = peps.build_basis(nx, ny, dg, g, ctx=ctx)
    ref = ED.build_basis(nx, ny, dg, g)
    print(f"ED basis {nx}x{ny} dg={dg} {kind}: dim {ed.dim}", flush=True)
    assert ed.dim == ref.dim
    np.testing.assert_array_equal(ed.keys(), ref.keys)


@pytest.mark.parametrize("nx,ny,dg,kind", [(3, 4, 1, "stag"), (2, 2, 2, "stag"), (3, 4, 1, "string")])
def test_hamiltonian_action(ctx, nx, ny, dg, kind):
    from lgt_amd import peps
    g = charges(nx, ny, kind)
    ed = peps.build_basis(nx, ny, dg, g, ctx=ctx).hamiltonian(0.8, 1.1, 0.6)
    H = ED.build_hamiltonian(ED.build_basis(nx, ny, dg, g), 0.8, 1.1, 0.6)
    x = np.random.default_rng(1).standard_normal(ed.dim)
    y, yr = ed.apply(x), H @ x
    assert np.linalg.norm(y - yr) <= 1e-14 * np.linalg.norm(yr)


@pytest.mark.parametrize("g,t,m", [(1.0, 1.0, 0.5), (0.5, 1.0, 0.25)])
def test_ground_state_and_observables_3x4(ctx, g, t, m):
    from lgt_amd import peps
    gch = staggered(3, 4)
    ed = peps.build_basis(3, 4, 1, gch, ctx=ctx).hamiltonian(g, t, m)
    E0, psi = ed.ground_state(1e-10)
    b = ED.build_basis(3, 4, 1, gch)
    H = ED.build_hamiltonian(b, g, t, m)
    E0r, psir = ED.find_ground_state(H)
    assert abs(E0 - E0r) <= 1e-10 * abs(E0r)
    assert abs(abs(psi @ psir) - 1.0) <= 1e-9
    assert np.linalg.norm(H @ psi - E0 * psi) <= 1e-9
    obs, ref = ed.measure(), ED.measure_ed(psi, b)
    for k in ("nf_mean", "nf_even", "nf_odd", "E2_mean"):
        assert abs(obs[k] - ref[k]) <= 1e-12
    np.testing.assert_allclose(obs["nf_grid"], ref["nf_grid"], atol=1e-12)
    np.testing.assert_allclose(obs["E2_grid"], ref["E2_grid"], atol=1e-12)
    assert obs["nf_mean"] == pytest.approx(0.5, abs=1e-12)  # N_f = 6 of 12 sites


def test_4x4_properties(ctx):
    from lgt_amd import peps
    gch = staggered(4, 4)
    ed = peps.build_basis(4, 4, 1, gch, ctx=ctx).hamiltonian(1.0, 1.0, 0.5)
    print(f"ED 4x4 staggered: dim {ed.dim}", flush=True)
    keys = ed.keys()
    assert len(np.unique(keys)) == ed.dim
    # every key holds N_f = 8 fermions in its low 16 bits
    assert np.all(np.array([bin(int(k) & 0xFFFF).count("1") for k in keys[:: max(1, ed.dim // 4096)]]) == 8)
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(ed.dim), rng.standard_normal(ed.dim)
    Hy, Hx = ed.apply(y), ed.apply(x)
    assert abs(x @ Hy - y @ Hx) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(Hy)
    E0, psi = ed.ground_state(1e-9)
    assert np.linalg.norm(ed.apply(psi) - E0 * psi) <= 1e-9
    diag = np.array([ed.apply(np.eye(1, ed.dim, i)[0])[i] for i in range(0, ed.dim, max(1, ed.dim // 64))])
    assert E0 <= diag.min()
    obs = ed.measure()
    assert obs["nf_grid"].sum() == pytest.approx(8.0, abs=1e-9)


def test_evolve_matches_expm_multiply(ctx):
    """evolve_ED! (finite_ed.jl:405-430) on the GPU: ten steps of dt = 0.05
    from a seeded complex state of the 3x2 staggered sector equal the
    oracle's expm_multiply evolution to 1e-9 (state and observables), with
    the state renormalised every step."""
    from lgt_amd import peps
    nx, ny, dg = 3, 2, 1
    print("ED evolve 3x2", flush=True)
    gch = np.asarray(P.staggered_charges(nx, ny))
    b = ED.build_basis(nx, ny, dg, gch)
    H = ED.build_hamiltonian(b, 1.0, 1.0, 0.5)
    rng = np.random.default_rng(2)
    psi0 = rng.standard_normal(b.dim) + 1j * rng.standard_normal(b.dim)
    psi0 /= np.linalg.norm(psi0)
    ed = peps.ExactDiag(nx, ny, dg, gch, ctx=ctx).hamiltonian(1.0, 1.0, 0.5)
    data, psi, proj = ed.evolve(psi0, 0.05, 10)
    want = ED.evolve_ed(psi0, H, 0.05, 10)
    assert np.abs(psi - want[-1]).max() < 1e-9
    assert len(data) == 11 and proj >= 10
    for s, d in enumerate(data):
        obs = ED.measure_ed(psi0 if s == 0 else want[s - 1], b)
        assert np.abs(d["nf_grid"] - obs["nf_grid"]).max() < 1e-9, s
        assert np.abs(d["E2_grid"] - obs["E2_grid"]).max() < 1e-9, s
        assert d["t"] == pytest.approx(0.05 * s)


def test_quench_A_ED_matches_oracle_and_mps():
    """run_quench_A_ED (finite_ed.jl:491-551) at 2x2: the string state found
    through its key, evolved 8 steps of dt = 0.05; the observables equal the
    oracle ED evolution's (1e-9) and the MPS Krylov quench's on the same
    lattice (mps_quench.jl's protocol, 1e-7), and the CSV has the reference's
    columns."""
    import tempfile
    from lgt_amd import ed_quench, mps, mps_quench, peps
    nx, ny, dg = 2, 2, 1
    print("ED quench A 2x2", flush=True)
    ctx = peps.Context(0)
    try:
        with tempfile.TemporaryDirectory() as d:
            data, path = ed_quench.run_quench_A_ED(nx=nx, ny=ny, dt=0.05, n_steps=8, results_dir=d, ctx=ctx,
                                                   log=None)
            lines = open(path).read().splitlines()
        assert lines[0] == ",".join(ed_quench.ED_CSV_COLUMNS) and len(lines) == 10
        H = mps.build_H_mpo(nx, ny, 2 * dg + 1, 1.0, 1.0, 2.0, ctx=ctx)
        mdata, _ = mps_quench.evolve_mps(mps.string_breaking_mps(nx, ny, 2 * dg + 1), H, nx, ny, 2 * dg + 1,
                                         0.05, 8, 200, ctx=ctx)
    finally:
        ctx.close()
    g = ed_quench.string_charges(nx, ny)
    b = ED.build_basis(nx, ny, dg, g)
    n_f = np.zeros((1, nx, ny), np.int8)
    n_f[0, 1, 0] = 1
    e_R = np.zeros((1, nx - 1, ny), np.int8)
    e_R[0, 0, 0] = 1
    k0 = ED.encode(n_f, e_R, np.zeros((1, nx, ny - 1), np.int8), dg)[0]
    psi0 = np.zeros(b.dim, complex)
    psi0[int(np.nonzero(b.keys == k0)[0][0])] = 1.0
    states = [psi0] + ED.evolve_ed(psi0, ED.build_hamiltonian(b, 1.0, 1.0, 2.0), 0.05, 8)
    for s in range(9):
        obs = ED.measure_ed(states[s], b)
        for key in ("nf_grid", "E2_grid"):
            assert np.abs(data[s][key] - obs[key]).max() < 1e-9, (s, key)
            assert np.abs(mdata[s][key] - obs[key]).max() < 1e-7, (s, key)
    assert np.abs(data[-1]["nf_grid"] - data[0]["nf_grid"]).max() > 1e-3
