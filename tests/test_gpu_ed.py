"""The GPU exact diagonalisation (finite_ed.jl) against the ED oracle.

Basis keys bit-exact and in the same order; H x to 1e-14; the ground-state
energy to 1e-10 and the state to an overlap of 1 - 1e-9; the observables to
1e-12 on a given state.  At 4x4 (beyond the oracle's reach in a test) the
size-independent properties: H symmetric, the particle number, the ground
state's residual and its energy below every diagonal entry."""
import numpy as np
import pytest

from oracle import ed_oracle as ED

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
