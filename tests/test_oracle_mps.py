"""The MPS/DMRG oracle (oracle/mps_oracle.py) against the ED oracle, and the
library's host-side term lists (lgt_amd/mps.py, mps_lgt.jl:83-116,474-514)
against the oracle's full-space Hamiltonian.  CPU only."""
import numpy as np
import pytest

from oracle import ed_oracle as ED
from oracle import mps_oracle as M
from oracle import peps_oracle as P


def _sector(G, gch):
    return np.all(np.abs(G - np.asarray(gch)[None]) < 1e-12, axis=(1, 2))


@pytest.mark.parametrize("charges", ["staggered", "string"])
def test_full_hamiltonian_sector_spectrum_matches_ed(charges):
    """Restricted to a Gauss-law sector, the full-space Hamiltonian of the
    snake chain has the ED Hamiltonian's spectrum (2x2, gd = 3, i.e. the
    reference's dg = 1), and the sector has the ED basis' size."""
    nx, ny, gd = 2, 2, 3
    if charges == "staggered":
        gch = np.asarray(P.staggered_charges(nx, ny))
    else:  # string_charges (mps_lgt.jl:640-643), iy_str = 1
        gch = np.zeros((nx, ny), np.int64)
        gch[0, 0], gch[1, 0] = 1, -2
    H = M.full_hamiltonian(nx, ny, gd, 1.0, 1.0, 0.5)
    assert np.abs(H - H.T).max() < 1e-14
    sel = _sector(M.gauss_values(nx, ny, gd), gch)
    b = ED.build_basis(nx, ny, 1, gch)
    assert sel.sum() == b.dim > 0
    w_full = np.linalg.eigvalsh(H[np.ix_(sel, sel)])
    w_ed = np.linalg.eigvalsh(ED.build_hamiltonian(b, 1.0, 1.0, 0.5).toarray())
    assert np.abs(w_full - w_ed).max() < 1e-12


def test_library_terms_sum_to_the_full_hamiltonian():
    """lgt_terms: the sum of coef x (product of the listed site operators)
    is the oracle's full-space Hamiltonian; gauss_penalty_terms sum to
    Lambda sum_i ((G_i - g_i)^2 - g_i^2), diagonal in the product basis."""
    from lgt_amd import mps
    nx, ny, gd, lam = 2, 2, 3, 7.0
    dims = M.node_dims(nx, ny, gd)
    assert mps.node_dims(nx, ny, gd) == dims
    assert mps.snake_nodes(3, 2)[0] == [(1, 1), (2, 1), (3, 1), (3, 2), (2, 2), (1, 2)]

    def total(terms):
        return sum(c * M._apply_local(dims, ops) for c, ops in terms)

    H = total(mps.lgt_terms(nx, ny, gd, 1.3, 0.7, 0.4))
    assert np.abs(H - M.full_hamiltonian(nx, ny, gd, 1.3, 0.7, 0.4)).max() < 1e-13
    gch = np.asarray(P.staggered_charges(nx, ny))
    Hg = total(mps.gauss_penalty_terms(nx, ny, gd, gch, lam))
    G = M.gauss_values(nx, ny, gd)
    want = lam * (((G - gch[None]) ** 2).sum(axis=(1, 2)) - (gch ** 2).sum())
    assert np.abs(Hg - np.diag(want)).max() < 1e-12


def test_mpo_to_dense_of_product_terms():
    """mpo_to_dense of a bond-1 term MPO is the Kronecker product of its site
    operators (first site slowest), and of a direct sum the sum."""
    from lgt_amd import mps
    rng = np.random.default_rng(3)
    dims = [2, 3, 2]
    A, B = rng.standard_normal((2, 2)), rng.standard_normal((2, 2))
    W1 = mps.term_mpo(dims, (0.5, {0: A, 2: B}))
    W2 = mps.term_mpo(dims, (-1.5, {1: rng.standard_normal((3, 3))}))
    D1, D2 = M.mpo_to_dense(W1), M.mpo_to_dense(W2)
    assert np.abs(D1 - 0.5 * np.kron(np.kron(A, np.eye(3)), B)).max() < 1e-15
    assert np.abs(M.mpo_to_dense(mps.mpo_axpby(1.0, W1, 2.0, W2)) - (D1 + 2.0 * D2)).max() < 1e-14


def test_krylov_step_matches_the_exact_exponential():
    """krylov_evolve_dense (krylov_evolve_step, mps_lgt.jl:546-586, on
    vectors) against exp(-i dt H) psi from the eigenvectors of H, 2x2, gd = 3,
    a generic complex start: the Lanczos error falls with m (m = 8 at
    dt = 0.05: ~1e-11 after five steps), and the opposite sign of dt is far
    off."""
    H = M.full_hamiltonian(2, 2, 3, 1.0, 1.0, 2.0)
    w, Q = np.linalg.eigh(H)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(H.shape[0]) + 1j * rng.standard_normal(H.shape[0])
    v /= np.linalg.norm(v)

    def exact(sign):
        y = v.copy()
        for _ in range(5):
            y = Q @ (np.exp(-1j * sign * 0.05 * w) * (Q.conj().T @ y))
        return y

    errs = {}
    for m in (4, 8):
        x = v.copy()
        for _ in range(5):
            x = M.krylov_evolve_dense(H, x, 0.05, m)
        errs[m] = np.abs(x - exact(1.0)).max()
    assert errs[8] < 1e-9 and errs[4] > 1e-7
    assert np.abs(x - exact(-1.0)).max() > 1e-2


def test_string_product_vector_sector():
    """string_breaking_mps (mps_lgt.jl:624-637) as a full-space vector lies
    in the Gauss sector of string_charges (:640-643), 2x2 and 2x3."""
    from lgt_amd import mps
    for nx, ny in ((2, 2), (2, 3)):
        gd = 3
        st = mps.string_breaking_mps(nx, ny, gd)
        cfg = [int(np.argmax(a.reshape(-1))) + 1 for a in st]
        dims = M.node_dims(nx, ny, gd)
        if int(np.prod(dims)) > 20000:
            # the Gauss values of the single product state
            idx = np.ravel_multi_index(tuple(c - 1 for c in cfg), dims)
            chain, pos = M.snake_nodes(nx, ny)
            G = np.zeros((nx, ny))
            for (ix, iy), p in pos.items():
                _, dR, dU = P.site_dims(ix, iy, nx, ny, gd)
                f, lR, lU = P.decode_phys(cfg[p], dR, dU, gd)
                eR = lR / 2.0 if dR > 1 else 0.0
                eU = lU / 2.0 if dU > 1 else 0.0
                G[ix - 1, iy - 1] += eR + eU - f
                if ix < nx:
                    G[ix, iy - 1] -= eR
                if iy < ny:
                    G[ix - 1, iy] -= eU
            assert idx >= 0
        else:
            v = M.product_vector(cfg, dims)
            G = M.gauss_values(nx, ny, gd)[int(np.argmax(np.abs(v)))]
        assert np.array_equal(G, mps.string_charges(nx, ny))


def test_quench_csv_writers(tmp_path):
    """write_mps_csv (mps_quench.jl:64-75) and data_to_df_ED
    (finite_ed.jl:437-445) write the reference's five columns, one row per
    measurement, with Julia's float printing; the drivers' CLIs refuse an
    unknown protocol."""
    from lgt_amd import ed_quench, mps_quench
    rows = [dict(t=0.0, nf_mean=1 / 12, nf_even=1 / 6, nf_odd=0.0, E2_mean=1 / 12, extra=1),
            dict(t=0.05, nf_mean=0.0833, nf_even=0.1654, nf_odd=0.0012, E2_mean=0.084, extra=2)]
    p1 = mps_quench.write_mps_csv(rows, "mps_quench_A", str(tmp_path))
    p2 = ed_quench.data_to_csv(rows, str(tmp_path / "finite_ed_quench_A_data.csv"))
    for p in (p1, p2):
        lines = open(p).read().splitlines()
        assert lines[0] == "t,nf_mean,nf_even,nf_odd,E2_mean" and len(lines) == 3
        assert lines[1].split(",")[0] == "0.0" and lines[2].split(",")[3] == "0.0012"
    assert p1.endswith("mps_quench_A_data.csv")
    with pytest.raises(SystemExit):
        mps_quench.main(["4", "--results", str(tmp_path)])
    with pytest.raises(SystemExit):
        ed_quench.main(["0", "--results", str(tmp_path)])
