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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("task_id", nargs="?", type=int, default=int(os.environ.get("SLURM_ARRAY_TASK_ID", "1")))
    ap.add_argument("--outdir", default=os.path.join("results", "gs_bench_v4"))
    ap.add_argument("--n-ite", type=int, default=V4_N_ITE)
    ap.add_argument("--seed", type=int, default=None)
    a = ap.parse_args(argv)
    run_one_v4(a.task_id, a.outdir, n_ite=a.n_ite, seed=a.seed, log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    sys.exit(main())
