"""Real-time quenches of the snake MPS (mps_quench.jl) on the GPU.

The evolution runs in the library: `peps.krylov_evolve` is krylov_evolve_step
(mps_lgt.jl:546-586) -- Lanczos on the MPS manifold, every H v by the zip-up
MPO application, sums by direct sums, compression by the truncated sweep --
on device tensors, and `mps.measure_mps` reads the observables from one pass
of overlap environments.  This module is the host-side driver around them,
mirroring the reference's functions of the same names: the evolution loop
(evolve_mps, :42-62), the CSV of its time series (write_mps_csv, :64-75) and
the three protocols (run_quench_A/B/C_mps, :78-123: string breaking, mass
quench and coupling quench on the 3x4 lattice, ground states of B and C from
Gauss-penalised DMRG).  The exact-diagonalisation runs that main() of the
reference adds beside each protocol are not part of this module.

Usage:  python -m lgt_amd.mps_quench 1|2|3 [--steps N] [--results DIR]
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import lattice, mps, peps
from .quench import write_csv

# mps_quench.jl:30-37 (gd = 2 dg + 1)
Q_NX, Q_NY, Q_DG = 3, 4, 1
Q_GD = 2 * Q_DG + 1
Q_THOP = 1.0
Q_DT = 0.05
Q_NSTEPS = 40
Q_DMAX = 120
Q_KRYLOV = 8
Q_DGS = 40
Q_NSW = 8
Q_LAM = 5.0

MPS_CSV_COLUMNS = ("t", "nf_mean", "nf_even", "nf_odd", "E2_mean")


def evolve_mps(psi0, H, nx, ny, gd, dt, n_steps, Dmax, m=Q_KRYLOV, eps=1e-10, ctx=None, log=None, log_every=5):
    """evolve_mps (mps_quench.jl:42-62): measure at t = 0, then after every
    krylov_evolve_step; returns (the list of measure_mps dicts, the final
    MPS).  The state stays a host list of arrays between steps; each row
    also carries the largest bond and the wall time since the first step."""
    norm = np.sqrt(mps.mps_overlap(psi0, ctx=ctx))
    psi = [np.array(a, dtype=np.complex128) for a in psi0]
    psi[0] = psi[0] / norm
    data = [mps.measure_mps(psi, nx, ny, gd, t_now=0.0, ctx=ctx)]
    if log:
        log("  step  t       <n_f>    <n_f>_e  <n_f>_o   <E^2>    bond")
        log(f"  {0:4d}  {0.0:5.2f}   {data[0]['nf_mean']:.4f}   {data[0]['nf_even']:.4f}   "
            f"{data[0]['nf_odd']:.4f}   {data[0]['E2_mean']:.4f}   {max(a.shape[1] for a in psi)}")
    t0 = time.time()
    for step in range(1, n_steps + 1):
        psi = peps.krylov_evolve(H, psi, dt, nsteps=1, m=m, Dmax=Dmax, eps=eps, ctx=ctx)
        obs = mps.measure_mps(psi, nx, ny, gd, t_now=step * dt, ctx=ctx)
        obs["bond"] = max(a.shape[1] for a in psi)
        obs["wall_s"] = time.time() - t0
        data.append(obs)
        if log and (step % log_every == 0 or step == n_steps):
            log(f"  {step:4d}  {step * dt:5.2f}   {obs['nf_mean']:.4f}   {obs['nf_even']:.4f}   "
                f"{obs['nf_odd']:.4f}   {obs['E2_mean']:.4f}   {obs['bond']}   ({obs['wall_s']:.1f} s)")
    return data, psi


def write_mps_csv(data, label, results_dir):
    """write_mps_csv (mps_quench.jl:64-75): <label>_data.csv."""
    path = os.path.join(results_dir, f"{label}_data.csv")
    write_csv([{k: d[k] for k in MPS_CSV_COLUMNS} for d in data], MPS_CSV_COLUMNS, path)
    return path


def _ground_state(g, m, ctx, log):
    gch = np.asarray(lattice.staggered_charges(Q_NX, Q_NY))
    dims = mps.node_dims(Q_NX, Q_NY, Q_GD)
    Hpen = mps.build_penalized_H(Q_NX, Q_NY, Q_GD, g, Q_THOP, m, gch, Q_LAM, ctx=ctx)
    E, psi, sweeps = mps.dmrg_ground_state(Hpen, dims, D=Q_DGS, nsweeps=Q_NSW, ctx=ctx,
                                           psi0=mps.staggered_mps(Q_NX, Q_NY, Q_GD))
    if log:
        log(f"  penalised DMRG: E = {E:.10f} after {sweeps} sweeps")
    return psi


def run_quench_A_mps(results_dir, n_steps=Q_NSTEPS, ctx=None, log=print, log_every=5):
    """run_quench_A_mps (mps_quench.jl:78-86): string breaking at g = 1, m = 2."""
    g, m = 1.0, 2.0
    psi0 = mps.string_breaking_mps(Q_NX, Q_NY, Q_GD)
    H = mps.build_H_mpo(Q_NX, Q_NY, Q_GD, g, Q_THOP, m, ctx=ctx)
    data, _ = evolve_mps(psi0, H, Q_NX, Q_NY, Q_GD, Q_DT, n_steps, Q_DMAX, ctx=ctx, log=log, log_every=log_every)
    return write_mps_csv(data, "mps_quench_A", results_dir)


def run_quench_B_mps(results_dir, m_init=5.0, m_final=0.1, n_steps=Q_NSTEPS, ctx=None, log=print, log_every=5):
    """run_quench_B_mps (mps_quench.jl:89-103): ground state at m_init, evolved
    with m_final (g = 1)."""
    g = 1.0
    psi0 = _ground_state(g, m_init, ctx, log)
    H = mps.build_H_mpo(Q_NX, Q_NY, Q_GD, g, Q_THOP, m_final, ctx=ctx)
    data, _ = evolve_mps(psi0, H, Q_NX, Q_NY, Q_GD, Q_DT, n_steps, Q_DMAX, ctx=ctx, log=log, log_every=log_every)
    return write_mps_csv(data, "mps_quench_B", results_dir)


def run_quench_C_mps(results_dir, g_init=2.0, g_final=0.5, n_steps=Q_NSTEPS, ctx=None, log=print, log_every=5):
    """run_quench_C_mps (mps_quench.jl:106-120): ground state at g_init,
    evolved with g_final (m = 2)."""
    m = 2.0
    psi0 = _ground_state(g_init, m, ctx, log)
    H = mps.build_H_mpo(Q_NX, Q_NY, Q_GD, g_final, Q_THOP, m, ctx=ctx)
    data, _ = evolve_mps(psi0, H, Q_NX, Q_NY, Q_GD, Q_DT, n_steps, Q_DMAX, ctx=ctx, log=log, log_every=log_every)
    return write_mps_csv(data, "mps_quench_C", results_dir)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("quench_id", type=int, nargs="?", default=int(os.environ.get("SLURM_ARRAY_TASK_ID", "1")))
    ap.add_argument("--steps", type=int, default=Q_NSTEPS)
    ap.add_argument("--results", default="results")
    ap.add_argument("--log-every", type=int, default=5)
    a = ap.parse_args(argv)
    os.makedirs(a.results, exist_ok=True)
    t0 = time.time()
    run = {1: run_quench_A_mps, 2: run_quench_B_mps, 3: run_quench_C_mps}.get(a.quench_id)
    if run is None:
        raise SystemExit("quench_id must be 1 (A), 2 (B), or 3 (C)")
    path = run(a.results, n_steps=a.steps, log=lambda s: print(s, flush=True), log_every=a.log_every)
    print(f"  Saved: {path}  ({time.time() - t0:.1f} s)", flush=True)


if __name__ == "__main__":
    main()
