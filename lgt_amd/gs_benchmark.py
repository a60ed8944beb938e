"""Ground-state benchmark of the boundary-MPS full update against exact
diagonalisation, with the reference's per-task CSV schema
(gs_benchmark_v4.jl:54-120), so that gs_benchmark_collect.jl and the plot
scripts read this driver's files unchanged.

One task = one (m, g) point of V4_PARAM_GRID (task_id 1-based, wrapping as
the SLURM array index does): ite_ground_state_v4 from init_gauge_peps(noise),
measure_all_finite, forbidden_norm and compute_energy_v4 on the GPU, then ED
of the same Gauss sector (build_basis / build_hamiltonian /
find_ground_state / measure_ED, also on the GPU), written as one CSV row to
<outdir>/gs_benchmark_<task_id:03d>.csv.

Usage: python -m lgt_amd.gs_benchmark [task_id] [--outdir DIR] [--n-ite N]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

from . import lattice, peps
from .quench import write_csv

# gs_benchmark_v4.jl:30-44
V4_PARAM_GRID = [
    (0.25, 0.25), (0.25, 0.50), (0.25, 1.00),
    (0.25, 2.00), (0.25, 4.00),
    (0.50, 0.25), (0.50, 0.50), (0.50, 2.00),
    (0.50, 4.00),
]
V4_NX, V4_NY, V4_DG, V4_D_MAX, V4_N_ITE = 3, 4, 1, 4, 200
V4_T_HOP, V4_NOISE, V4_CHI, V4_NALS = 1.0, 0.1, 48, 30

# the DataFrame of gs_benchmark_v4.jl:97-113, column order kept
V4_COLUMNS = (
    "task_id", "m", "g", "t_hop", "nx", "ny", "dg",
    "D_max", "n_ite", "noise", "chi", "n_als",
    "D_final", "gauge_viol",
    "peps_energy",
    "peps_nf_mean", "peps_nf_even", "peps_nf_odd",
    "peps_E2_mean", "peps_S_h", "peps_S_v",
    "peps_S_mean", "peps_time_s",
    "ed_E0", "ed_nf_mean", "ed_nf_even",
    "ed_nf_odd", "ed_E2_mean",
    "ed_basis_dim", "ed_time_s",
    "d_energy",
    "d_nf_mean",
    "d_nf_even",
    "d_nf_odd",
    "d_E2_mean",
    "d_cdw",
)


def run_one_v4(task_id=1, outdir=None, ctx=None, n_ite=V4_N_ITE, nx=V4_NX, ny=V4_NY, dg=V4_DG,
               D_max=V4_D_MAX, chi=V4_CHI, n_als=V4_NALS, noise=V4_NOISE, seed=None, log=None):
    """run_one_v4 (gs_benchmark_v4.jl:54-120).  Returns the row (a dict keyed
    by V4_COLUMNS); writes it when outdir is given.  seed fixes the noise of
    the initial state (the reference draws it from Julia's global RNG)."""
    m, g = V4_PARAM_GRID[(task_id - 1) % len(V4_PARAM_GRID)]
    gd = 2 * dg + 1
    gch = lattice.staggered_charges(nx, ny)
    ctx = ctx or peps.default_context()
    t0 = time.time()
    gp = peps.ite_ground_state_v4(nx, ny, gd, D_max, g, V4_T_HOP, m, gch, n_ite=n_ite, noise=noise, chi=chi,
                                  n_als=n_als, seed=seed, ctx=ctx)
    obs = peps.measure_all_finite(gp, 0.0)
    gviol = peps.forbidden_norm(gp)
    E_peps = peps.compute_energy_v4(gp, g, V4_T_HOP, m, chi)
    ctx.sync()
    t_peps = time.time() - t0
    if log:
        log(f"[PEPS]  E={E_peps:.6f}  <n_f>={obs['nf_mean']:.6f}  nf_e={obs['nf_even']:.6f}  "
            f"nf_o={obs['nf_odd']:.6f}  <E2>={obs['E2_mean']:.6f}  viol={gviol:.1e}  ({t_peps:.1f}s)")
    t1 = time.time()
    ed = peps.build_basis(nx, ny, dg, gch, ctx=ctx)
    peps.build_hamiltonian(ed, g, V4_T_HOP, m)
    E0, psi0 = peps.find_ground_state(ed)
    obs_ed = peps.measure_ED(ed, psi0)
    t_ed = time.time() - t1
    if log:
        log(f"[ED ]  E0={E0:.6f}  nf_e={obs_ed['nf_even']:.6f}  <E2>={obs_ed['E2_mean']:.6f}  ({t_ed:.1f}s)")
    D_final = max([len(gp.bond(peps.PEPS_H, ix, iy)[1]) for iy in range(1, ny + 1) for ix in range(1, nx)],
                  default=1)
    row = dict(
        task_id=int(task_id), m=m, g=g, t_hop=V4_T_HOP, nx=nx, ny=ny, dg=dg,
        D_max=D_max, n_ite=int(n_ite), noise=noise, chi=chi, n_als=n_als,
        D_final=int(D_final), gauge_viol=gviol,
        peps_energy=E_peps,
        peps_nf_mean=obs["nf_mean"], peps_nf_even=obs["nf_even"], peps_nf_odd=obs["nf_odd"],
        peps_E2_mean=obs["E2_mean"], peps_S_h=obs["S_h_mean"], peps_S_v=obs["S_v_mean"],
        peps_S_mean=obs["S_mean"], peps_time_s=t_peps,
        ed_E0=E0, ed_nf_mean=obs_ed["nf_mean"], ed_nf_even=obs_ed["nf_even"],
        ed_nf_odd=obs_ed["nf_odd"], ed_E2_mean=obs_ed["E2_mean"],
        ed_basis_dim=int(ed.dim), ed_time_s=t_ed,
        d_energy=E_peps - E0,
        d_nf_mean=obs["nf_mean"] - obs_ed["nf_mean"],
        d_nf_even=obs["nf_even"] - obs_ed["nf_even"],
        d_nf_odd=obs["nf_odd"] - obs_ed["nf_odd"],
        d_E2_mean=obs["E2_mean"] - obs_ed["E2_mean"],
        d_cdw=(obs["nf_even"] - obs["nf_odd"]) - (obs_ed["nf_even"] - obs_ed["nf_odd"]),
    )
    row = {k: (float(v) if not isinstance(v, int) else v) for k, v in row.items()}
    if outdir:
        path = os.path.join(outdir, f"gs_benchmark_{int(task_id):03d}.csv")
        write_csv([row], V4_COLUMNS, path)
        if log:
            log(f"  Saved: {path}")
    return row


# gs_benchmark_mps.jl:26-39: the same (m, g) grid as gs_benchmark_v3.jl, the
# 3x4 lattice at dg = 1, DMRG bond 40, at most 8 sweeps, Lambda = 5
MPS_PARAM_GRID = [(0.25, 0.25), (0.25, 0.50), (0.25, 1.00), (0.25, 2.00), (0.25, 4.00),
                  (0.50, 0.25), (0.50, 0.50), (0.50, 2.00), (0.50, 4.00)]
MPS_NX, MPS_NY, MPS_DG, MPS_T_HOP, MPS_D, MPS_NSW, MPS_LAM = 3, 4, 1, 1.0, 40, 8, 5.0
MPS_COLUMNS = ("task", "m", "g", "E_ed", "E_mps", "dE_rel", "nf_mean_ed", "nf_mean_mps", "nf_even_ed",
               "nf_even_mps", "nf_odd_ed", "nf_odd_mps", "E2_mean_ed", "E2_mean_mps", "gauss_viol", "bond")


def run_one_mps(task_id=1, outdir=None, ctx=None, nx=MPS_NX, ny=MPS_NY, dg=MPS_DG, D=MPS_D, nsweeps=MPS_NSW,
                lam=MPS_LAM, start="staggered", seed=1, log=None):
    """run_one_mps (gs_benchmark_mps.jl:41-96): the sector's ED ground state
    and the Gauss-penalised DMRG one (GPU), compared in one CSV row keyed by
    MPS_COLUMNS (gs_mps_task<id>.csv).  start "staggered" is the reference's
    product start in the target sector; "random" a seeded random MPS."""
    from . import mps
    m, g = MPS_PARAM_GRID[(task_id - 1) % len(MPS_PARAM_GRID)]
    gd = 2 * dg + 1
    gch = lattice.staggered_charges(nx, ny)
    ctx = ctx or peps.default_context()
    t0 = time.time()
    ed = peps.build_basis(nx, ny, dg, gch, ctx=ctx)
    peps.build_hamiltonian(ed, g, MPS_T_HOP, m)
    E_ed, psi_ed = peps.find_ground_state(ed)
    obs_ed = peps.measure_ED(ed, psi_ed)
    if log:
        log(f"  [ED] E0 = {E_ed:.8f}  ({ed.dim} states, {time.time() - t0:.1f}s)")
    t1 = time.time()
    dims = mps.node_dims(nx, ny, gd)
    Hpen = mps.build_penalized_H(nx, ny, gd, g, MPS_T_HOP, m, gch, lam, ctx=ctx)
    psi0 = mps.staggered_mps(nx, ny, gd) if start == "staggered" else None
    _, psi, sweeps = mps.dmrg_ground_state(Hpen, dims, D=D, nsweeps=nsweeps, psi0=psi0, seed=seed, ctx=ctx)
    Hbare = mps.build_H_mpo(nx, ny, gd, g, MPS_T_HOP, m, ctx=ctx)
    E_mps = peps.mpo_expect(Hbare, psi, ctx=ctx) / mps.mps_overlap(psi, ctx=ctx)
    obs = mps.measure_mps(psi, nx, ny, gd, ctx=ctx)
    viol = mps.gauss_violation(psi, nx, ny, gd, gch, ctx=ctx)
    bond = max(t.shape[1] for t in psi)
    t_mps = time.time() - t1
    if log:
        log(f"  [MPS] E = {E_mps:.8f}  bond={bond}  Gauss-viol={viol:.2e}  ({t_mps:.1f}s, {sweeps} sweeps)")
        log(f"  dE/|E| = {abs(E_mps - E_ed) / abs(E_ed):.2e}")
    row = dict(task=int(task_id), m=m, g=g, E_ed=E_ed, E_mps=E_mps,
               dE_rel=abs(E_mps - E_ed) / max(abs(E_ed), 1e-12),
               nf_mean_ed=obs_ed["nf_mean"], nf_mean_mps=obs["nf_mean"], nf_even_ed=obs_ed["nf_even"],
               nf_even_mps=obs["nf_even"], nf_odd_ed=obs_ed["nf_odd"], nf_odd_mps=obs["nf_odd"],
               E2_mean_ed=obs_ed["E2_mean"], E2_mean_mps=obs["E2_mean"], gauss_viol=viol, bond=int(bond))
    row = {k: (float(v) if not isinstance(v, int) else v) for k, v in row.items()}
    row["time_s"] = t_mps
    if outdir:
        path = os.path.join(outdir, f"gs_mps_task{int(task_id)}.csv")
        write_csv([row], MPS_COLUMNS, path)
        if log:
            log(f"  Saved: {path}")
    return row


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("task_id", nargs="?", type=int, default=int(os.environ.get("SLURM_ARRAY_TASK_ID", "1")))
    ap.add_argument("--outdir", default=None)
    ap.add_argument("--n-ite", type=int, default=V4_N_ITE)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--mps", action="store_true", help="the MPS/DMRG benchmark task (gs_benchmark_mps.jl)")
    ap.add_argument("--start", default="staggered", choices=["staggered", "random"])
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    if a.mps:
        run_one_mps(a.task_id, a.outdir or "results", start=a.start, seed=a.seed or 1, log=log)
    else:
        run_one_v4(a.task_id, a.outdir or os.path.join("results", "gs_bench_v4"), n_ite=a.n_ite, seed=a.seed,
                   log=log)


if __name__ == "__main__":
    sys.exit(main())
