"""FP64 parity of the real-time quench path (finite_peps_quench.jl) and the
ground-state benchmark driver (gs_benchmark_v4.jl) on the GPU with the CPU
oracle.  States are compared through gauge-invariant quantities: the bond
weights, the contracted dense state (small lattices) and the observables of
measure_all_finite."""
import numpy as np
import pytest

from oracle import peps_oracle as O
from tests._util import random_peps, rel, to_oracle

pytestmark = pytest.mark.gpu


def _weights(gp):
    return ([np.asarray(gp.LH(ix, iy)) for iy in range(1, gp.ny + 1) for ix in range(1, gp.nx)]
            + [np.asarray(gp.LV(ix, iy)) for iy in range(1, gp.ny) for ix in range(1, gp.nx + 1)])


def _compare_weights(gg, go, tol):
    for a, b in zip(_weights(gg), _weights(go)):
        assert a.shape == b.shape
        assert np.abs(a - b).max() < tol


def _compare_obs(a, b, tol):
    for k in ("nf_mean", "nf_even", "nf_odd", "E2_mean", "S_h_mean", "S_v_mean", "S_mean"):
        assert abs(a[k] - b[k]) < tol, (k, a[k], b[k])
    assert np.abs(np.asarray(a["nf_grid"]) - b["nf_grid"]).max() < tol
    assert np.abs(np.asarray(a["E2_grid"]) - b["E2_grid"]).max() < tol


def _device(go, ctx):
    from lgt_amd import peps
    return peps.GaugePEPS.from_arrays(go.tensors, go.lh, go.lv, go.qh, go.qv, go.gch, go.gd, is_complex=True,
                                      ctx=ctx)


@pytest.mark.parametrize("Dtr", [2, 4])
def test_realtime_step_matches_oracle(ctx, Dtr):
    """Two realtime_trotter_step calls on a complex 3x3 (gd=2, D=2) state:
    weights and the dense state to 1e-10 (Dtr=2 truncates every bond)."""
    from lgt_amd import peps
    print(f"realtime step 3x3 gd=2 D=2 Dtr={Dtr}", flush=True)
    go = random_peps(3, 3, 2, 2, 23, cplx=True)
    gd_ = _device(go, ctx)
    for _ in range(2):
        peps.realtime_trotter_step(gd_, 1.0, 1.0, 0.5, 0.05, Dtr)
        O.realtime_trotter_step(go, 1.0, 1.0, 0.5, 0.05, Dtr)
    gg = to_oracle(gd_)
    _compare_weights(gg, go, 1e-10)
    assert rel(O.dense_psi(gg), O.dense_psi(go)) < 1e-10


def test_realtime_step_refuses_real_states_and_mu(ctx):
    from lgt_amd import peps
    go = random_peps(2, 2, 2, 2, 3, cplx=False)
    gp = peps.GaugePEPS.from_arrays(go.tensors, go.lh, go.lv, go.qh, go.qv, go.gch, go.gd, is_complex=False,
                                    ctx=ctx)
    with pytest.raises(peps.PepsError):
        peps.realtime_trotter_step(gp, 1.0, 1.0, 0.5, 0.05, 2)
    p = peps._params(1.0, 1.0, 0.5, 0.05, 2, 8, 0, mu=0.5)
    import ctypes
    assert peps.lib().peps_trotter_step(gp.h, peps.PEPS_STEP_GAUGE, ctypes.byref(p)) == 1  # PEPS_EARG


def test_quench_A_string_breaking(ctx):
    """run_quench_A's set-up and three real-time steps on its 3x4 lattice
    (dg=1, m=2, dt=0.02) at D_bond=1, D_max=3 (the oracle's dense SVDs at the
    reference's D_max=12 would take hours): the observables after every step
    agree with the oracle's to 1e-10."""
    from lgt_amd import lattice, peps, quench
    nx, ny, gd, D = 3, 4, 3, 1
    print("quench A 3x4 gd=3", flush=True)
    gp = quench.string_state(nx, ny, gd, D, ctx=ctx)
    a = lattice.init_finite_peps_arrays(nx, ny, gd, D, nf_of=lambda ix, iy: 0)
    go = O.GaugePEPS(nx, ny, gd, a["gch"])
    go.tensors = [[np.asarray(t, np.complex128) for t in col] for col in a["tensors"]]
    go.lh, go.lv, go.qh, go.qv = a["lh"], a["lv"], a["qh"], a["qv"]
    O.apply_site_op(go, 1, ny // 2, lattice.site_operators(1, ny // 2, nx, ny, gd)["U_R"])
    O.apply_site_op(go, 2, ny // 2, lattice.site_operators(2, ny // 2, nx, ny, gd)["cdag"])
    _compare_obs(peps.measure_all_finite(gp), O.measure_all_finite(go), 1e-14)
    data = quench.evolve(gp, 1.0, 1.0, 2.0, 0.02, 3, 3)
    for k in range(3):
        O.realtime_trotter_step(go, 1.0, 1.0, 2.0, 0.02, 3)
        _compare_obs(data[k + 1], O.measure_all_finite(go, 0.02 * (k + 1)), 1e-10)
        assert data[k + 1]["t"] == pytest.approx(0.02 * (k + 1))
    _compare_weights(to_oracle(gp), go, 1e-10)
    # the string spreads: the fermion density leaves its site
    assert data[-1]["nf_grid"][1, ny // 2 - 1] < 1.0


def test_quench_B_preparation_with_mu_and_csv(ctx, tmp_path):
    """run_quench_B at reduced sizes (2x3, dg=1, D_bond=1, D_max=2: 3 dense
    ITE steps with plaquettes and mu = 1 at m = 5, then 2 real-time steps at
    m = 0.1): observables to 1e-10 against the oracle; the CSV and snapshot
    files are written."""
    from lgt_amd import lattice, quench
    nx, ny, gd = 2, 3, 3
    print("quench B 2x3 gd=3", flush=True)
    data = quench.run_quench_B(nx=nx, ny=ny, gd=gd, D_bond=1, D_max=2, n_ite=3, n_steps=2,
                               results_dir=str(tmp_path), ctx=ctx)
    a = lattice.init_finite_peps_arrays(nx, ny, gd, 1)
    go = O.GaugePEPS(nx, ny, gd, a["gch"])
    go.tensors = [[np.asarray(t, np.complex128) for t in col] for col in a["tensors"]]
    go.lh, go.lv, go.qh, go.qv = a["lh"], a["lv"], a["qh"], a["qv"]
    for _ in range(3):
        O.trotter_step_dense(go, 1.0, 1.0, 5.0, 0.05, 2, mu=1.0)
    _compare_obs(data[0], O.measure_all_finite(go), 1e-10)
    for k in range(2):
        O.realtime_trotter_step(go, 1.0, 1.0, 0.1, 0.02, 2)
        _compare_obs(data[k + 1], O.measure_all_finite(go), 1e-10)
    lines = (tmp_path / "finite_peps_quench_B_data.csv").read_text().splitlines()
    assert lines[0] == ",".join(quench.CSV_COLUMNS) and len(lines) == 4
    assert float(lines[2].split(",")[0]) == pytest.approx(0.02)
    snap = (tmp_path / "finite_peps_quench_B_final.txt").read_text()
    assert "⟨n_f⟩" in snap and "⟨E²⟩" in snap and "iy=3:" in snap


def test_gs_benchmark_task_row(ctx, tmp_path):
    """One gs_benchmark_v4 task at a short schedule (8 ITE steps): the CSV row
    carries the reference's columns and equals the library's direct answers
    (ED E0 of the sector, differences consistent)."""
    from lgt_amd import gs_benchmark as gb, lattice, peps
    print("gs_benchmark task 3 (8 ITE steps)", flush=True)
    row = gb.run_one_v4(3, str(tmp_path), ctx=ctx, n_ite=8, seed=4)
    lines = (tmp_path / "gs_benchmark_003.csv").read_text().splitlines()
    assert lines[0] == ",".join(gb.V4_COLUMNS) and len(lines) == 2
    cells = dict(zip(gb.V4_COLUMNS, lines[1].split(",")))
    assert int(cells["task_id"]) == 3 and float(cells["m"]) == 0.25 and float(cells["g"]) == 1.0
    for k in ("peps_energy", "ed_E0", "d_energy", "gauge_viol"):
        assert float(cells[k]) == row[k]
    assert row["gauge_viol"] < 1e-9
    assert row["d_energy"] == pytest.approx(row["peps_energy"] - row["ed_E0"], abs=1e-14)
    assert row["d_cdw"] == pytest.approx((row["peps_nf_even"] - row["peps_nf_odd"])
                                         - (row["ed_nf_even"] - row["ed_nf_odd"]), abs=1e-14)
    ed = peps.build_basis(3, 4, 1, lattice.staggered_charges(3, 4), ctx=ctx)
    peps.build_hamiltonian(ed, 1.0, 1.0, 0.25)
    E0, _ = peps.find_ground_state(ed)
    assert row["ed_E0"] == pytest.approx(E0, rel=1e-10)
    assert row["ed_basis_dim"] == ed.dim
    # compute_energy_v4 normalises per bond, so it is not a variational
    # bound on E0; the row only has to carry a finite energy and bonds <= D_max
    assert np.isfinite(row["peps_energy"]) and 1 <= row["D_final"] <= 4
