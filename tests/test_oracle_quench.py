"""Pins for the oracle's real-time quench path (finite_peps_quench.jl) and
for the host-side output formats.

* The real-time gate exp(-i dt H) against its spectral form, unitarity and
  the group law (a sign error or a missing factor i fails).
* realtime_trotter_step on a product state with t_hop = 0: the bond gates
  then factor into on-site phases, so each site must end as
  exp(-i dt H_onsite) phi exactly (the on-site weights of the bonds touching
  a site sum to one; h bonds take two half steps, v bonds one full step).
  A wrong dt factor on either sweep or a wrong sign fails.
* apply_site_op through run_quench_A's set-up: c^dagger on the vacuum puts
  one fermion on its site and U on the right link one unit of flux there,
  which measure_all_finite must show exactly.
* The chemical potential mu of the merged bond Hamiltonians.
* Julia's float printing and the CSV schemas of data_to_df
  (finite_peps_quench.jl:239-250) and gs_benchmark_v4.jl:97-113."""
import numpy as np
import pytest
import scipy.linalg as sla

from lgt_amd import lattice
from lgt_amd.gs_benchmark import V4_COLUMNS, V4_PARAM_GRID
from lgt_amd.quench import CSV_COLUMNS, julia_float, write_csv
from oracle import peps_oracle as O


def _oracle_state(a):
    gp = O.GaugePEPS(len(a["tensors"]), len(a["tensors"][0]), a["gd"], a["gch"])
    gp.tensors = [[np.asarray(t, np.complex128) for t in col] for col in a["tensors"]]
    gp.lh, gp.lv, gp.qh, gp.qv = a["lh"], a["lv"], a["qh"], a["qv"]
    return gp


@pytest.mark.parametrize("dt", [0.02, 0.7])
def test_realtime_gate_is_the_unitary_exponential(dt):
    H = O.H_merged_h_site(1, 2, 3, 3, 3, 1.0, 1.0, 0.5)
    G = O.realtime_gate(H, dt)
    assert np.allclose(G, sla.expm(-1j * dt * H), atol=1e-12)  # Pade, as Julia's exp
    assert np.allclose(G @ G.conj().T, np.eye(len(H)), atol=1e-12)
    assert np.allclose(O.realtime_gate(H, dt / 2) @ O.realtime_gate(H, dt / 2), G, atol=1e-12)
    # a two-level check of the sign: exp(-i dt diag(0, 1)) = diag(1, e^{-i dt})
    assert np.allclose(O.realtime_gate(np.diag([0.0, 1.0]), dt), np.diag([1.0, np.exp(-1j * dt)]))


@pytest.mark.parametrize("gd", [2, 3])
def test_realtime_step_without_hopping_is_the_onsite_evolution(gd):
    nx, ny, dt, g, m = 3, 3, 0.3, 1.3, 0.7
    rng = np.random.default_rng(5)
    gp = O.GaugePEPS(nx, ny, gd, O.staggered_charges(nx, ny))
    phi = {}
    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            d = O.site_dims(ix, iy, nx, ny, gd)[0]
            v = rng.standard_normal(d) + 1j * rng.standard_normal(d)
            phi[ix, iy] = v / np.linalg.norm(v)
            gp.setT(ix, iy, phi[ix, iy].reshape(d, 1, 1, 1, 1).copy())
    gp.lh = [[np.ones(1) for _ in range(ny)] for _ in range(nx - 1)]
    gp.lv = [[np.ones(1) for _ in range(ny - 1)] for _ in range(nx)]
    gp.qh = [[np.zeros(1, np.int64) for _ in range(ny)] for _ in range(nx - 1)]
    gp.qv = [[np.zeros(1, np.int64) for _ in range(ny - 1)] for _ in range(nx)]
    O.realtime_trotter_step(gp, g, 0.0, m, dt, 4)
    for (ix, iy), v in phi.items():
        Hs = O.H_onsite_site(ix, iy, nx, ny, gd, g, m)
        w, V = np.linalg.eigh(Hs)
        want = (V * np.exp(-1j * dt * w)) @ V.T @ v
        got = gp.T(ix, iy).reshape(-1)
        assert got.size == want.size  # the product state stays at bond dimension 1
        # equal up to a complex factor (the SVD's phase and the weight split)
        cos = abs(np.vdot(got, want)) / (np.linalg.norm(got) * np.linalg.norm(want))
        assert cos > 1 - 1e-12, ((ix, iy), cos)
        # the opposite sign of i is a different state
        wrong = (V * np.exp(1j * dt * w)) @ V.T @ v
        assert abs(np.vdot(got, wrong)) / (np.linalg.norm(got) * np.linalg.norm(wrong)) < 1 - 1e-6


def test_string_state_puts_one_fermion_and_one_flux_unit():
    nx, ny, gd, D = 3, 4, 3, 4
    gp = _oracle_state(lattice.init_finite_peps_arrays(nx, ny, gd, D, nf_of=lambda ix, iy: 0))
    vac = O.measure_all_finite(gp)
    assert np.all(vac["nf_grid"] == 0) and np.all(vac["E2_grid"] == 0)
    iy = ny // 2
    O.apply_site_op(gp, 1, iy, lattice.site_operators(1, iy, nx, ny, gd)["U_R"])
    O.apply_site_op(gp, 2, iy, lattice.site_operators(2, iy, nx, ny, gd)["cdag"])
    obs = O.measure_all_finite(gp)
    want_nf = np.zeros((nx, ny))
    want_nf[1, iy - 1] = 1.0
    want_e2 = np.zeros((nx, ny))
    want_e2[0, iy - 1] = 1.0  # E_R = +1 on the right link of (1, iy)
    assert np.array_equal(obs["nf_grid"], want_nf)
    assert np.array_equal(obs["E2_grid"], want_e2)
    assert obs["nf_mean"] == pytest.approx(1.0 / (nx * ny), abs=1e-15)


def test_site_operators_match_the_oracle_embeddings():
    for gd in (2, 3):
        ops = lattice.site_operators(2, 2, 3, 3, gd)
        _, dR, dU = O.site_dims(2, 2, 3, 3, gd)
        assert np.array_equal(ops["cdag"], np.kron(np.kron(O.op_c().conj().T, np.eye(dR)), np.eye(dU)))
        assert np.array_equal(ops["U_R"], np.kron(np.kron(np.eye(2), O.op_U_gauge(gd)), np.eye(dU)))
        assert np.array_equal(ops["nf"], np.kron(np.kron(O.op_nf(), np.eye(dR)), np.eye(dU)))
    assert lattice.site_operators(3, 3, 3, 3, 3)["U_R"] is None


def test_chemical_potential_enters_the_merged_hamiltonians():
    nx, ny, gd, mu = 3, 3, 3, 0.37
    for ix, iy in [(1, 1), (2, 2)]:
        dL = O.site_dims(ix, iy, nx, ny, gd)[0]
        dR = O.site_dims(ix + 1, iy, nx, ny, gd)[0]
        nL = np.kron(np.kron(O.op_nf(), np.eye(dL // 2 // O.site_dims(ix, iy, nx, ny, gd)[2])),
                     np.eye(O.site_dims(ix, iy, nx, ny, gd)[2]))
        nR = np.kron(np.kron(O.op_nf(), np.eye(dR // 2 // O.site_dims(ix + 1, iy, nx, ny, gd)[2])),
                     np.eye(O.site_dims(ix + 1, iy, nx, ny, gd)[2]))
        diff = O.H_merged_h_site(ix, iy, nx, ny, gd, 1.0, 1.0, 0.5, mu) - O.H_merged_h_site(ix, iy, nx, ny, gd,
                                                                                            1.0, 1.0, 0.5)
        want = mu * (O.onsite_weight(ix, iy, nx, ny) * np.kron(nL, np.eye(dR))
                     + O.onsite_weight(ix + 1, iy, nx, ny) * np.kron(np.eye(dL), nR))
        assert np.allclose(diff, want, atol=1e-15)


def test_julia_float_printing():
    cases = {1e-5: "1.0e-5", 0.0001: "0.0001", 1e15: "1.0e15", 123456789.0: "1.23456789e8",
             100000.0: "100000.0", 1e6: "1.0e6", 2.0: "2.0", -0.5: "-0.5", 0.1: "0.1", 1.5e-7: "1.5e-7",
             123.456: "123.456", 0.0: "0.0", -3.75e-12: "-3.75e-12", 0.3333333333333333: "0.3333333333333333"}
    for x, s in cases.items():
        assert julia_float(x) == s, (x, julia_float(x), s)
        assert float(s) == x


def test_csv_schemas(tmp_path):
    # data_to_df (finite_peps_quench.jl:239-250)
    assert CSV_COLUMNS == ("t", "nf_mean", "nf_even", "nf_odd", "E2_mean", "S_h_mean", "S_v_mean", "S_mean")
    # gs_benchmark_v4.jl:97-113
    header = ("task_id,m,g,t_hop,nx,ny,dg,D_max,n_ite,noise,chi,n_als,D_final,gauge_viol,peps_energy,"
              "peps_nf_mean,peps_nf_even,peps_nf_odd,peps_E2_mean,peps_S_h,peps_S_v,peps_S_mean,peps_time_s,"
              "ed_E0,ed_nf_mean,ed_nf_even,ed_nf_odd,ed_E2_mean,ed_basis_dim,ed_time_s,d_energy,d_nf_mean,"
              "d_nf_even,d_nf_odd,d_E2_mean,d_cdw")
    assert ",".join(V4_COLUMNS) == header
    assert len(V4_PARAM_GRID) == 9 and V4_PARAM_GRID[0] == (0.25, 0.25) and V4_PARAM_GRID[-1] == (0.5, 4.0)
    rows = [dict(t=0.0, nf_mean=0.5, nf_even=0.25, nf_odd=0.75, E2_mean=1e-5, S_h_mean=0.0, S_v_mean=2.0,
                 S_mean=1.0), dict(t=0.02, nf_mean=0.5, nf_even=0.125, nf_odd=0.875, E2_mean=3.0, S_h_mean=1.0,
                                   S_v_mean=1.0, S_mean=1.0)]
    p = tmp_path / "q.csv"
    write_csv(rows, CSV_COLUMNS, str(p))
    lines = p.read_text().splitlines()
    assert lines[0] == ",".join(CSV_COLUMNS)
    assert lines[1] == "0.0,0.5,0.25,0.75,1.0e-5,0.0,2.0,1.0"
    assert lines[2] == "0.02,0.5,0.125,0.875,3.0,1.0,1.0,1.0"
