"""Real-time quenches by exact diagonalisation (finite_ed.jl) on the GPU.

The evolution runs in the library: `ExactDiag.evolve` is evolve_ED!
(finite_ed.jl:405-430) -- each step KrylovKit's exponentiate as the
reference calls it (Lanczos projections of at most 30 vectors, error
estimate <= 1e-10), the state renormalised, the observables measured -- on
the device basis and Hamiltonian.  This module is the host-side driver
around it, mirroring the reference's functions of the same names: the
initial states, the three protocols (run_quench_A/B/C_ED, :491-640) and the
CSV of their time series (data_to_df_ED, :437-445).  mps_quench.jl runs the
same protocols with its own dt and step count beside the MPS ones.

Usage:  python -m lgt_amd.ed_quench 1|2|3 [--dt DT] [--steps N] [--results DIR]
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np

from . import lattice, peps
from .quench import write_csv

# finite_ed.jl:60-65
ED_NX, ED_NY, ED_DG = 3, 4, 1
ED_G_COUP, ED_T_HOP, ED_M_MASS = 1.0, 1.0, 2.0

ED_CSV_COLUMNS = ("t", "nf_mean", "nf_even", "nf_odd", "E2_mean")


def state_key(n_f, e_R, e_U, dg):
    """_encode (finite_ed.jl:79-96): n_f bits (iy outer, ix inner), then
    E_R + dg and E_U + dg in link-width fields in the same order (2 bits for
    dg = 1, the field width of the library's basis keys)."""
    n_f, e_R, e_U = (np.asarray(a, dtype=np.int64) for a in (n_f, e_R, e_U))
    nx, ny = n_f.shape
    lb = 2
    while (1 << lb) < 2 * dg + 1:
        lb += 1
    key, off = 0, 0
    for iy in range(ny):
        for ix in range(nx):
            key |= int(n_f[ix, iy]) << off
            off += 1
    for iy in range(ny):
        for ix in range(nx - 1):
            key |= int(e_R[ix, iy] + dg) << off
            off += lb
    for iy in range(ny - 1):
        for ix in range(nx):
            key |= int(e_U[ix, iy] + dg) << off
            off += lb
    return key


def basis_state(ed: peps.ExactDiag, n_f, e_R, e_U):
    """The basis vector of one configuration (complex), or ValueError when
    it is not in the sector."""
    k0 = np.uint64(state_key(n_f, e_R, e_U, ed.dg))
    hit = np.nonzero(ed.keys() == k0)[0]
    if len(hit) != 1:
        raise ValueError("initial state not found in basis")
    psi = np.zeros(ed.dim, np.complex128)
    psi[hit[0]] = 1.0
    return psi


def string_charges(nx, ny):
    """The sector of quench A (finite_ed.jl:504-511): g(1, iy_str) = +1,
    g(2, iy_str) = -2, iy_str = ny / 2."""
    g = np.zeros((nx, ny), dtype=np.int64)
    g[0, ny // 2 - 1], g[1, ny // 2 - 1] = 1, -2
    return g


def data_to_csv(data, path):
    """data_to_df_ED (finite_ed.jl:437-445) written with CSV.write."""
    write_csv([{k: d[k] for k in ED_CSV_COLUMNS} for d in data], ED_CSV_COLUMNS, path)
    return path


def _log_rows(data, log, every):
    if not log:
        return
    log("  step  |    t        <n_f>    <n_f>_e  <n_f>_o   <E^2>")
    n = len(data) - 1
    for s, d in enumerate(data):
        if s == 0 or s % every == 0 or s == n:
            log(f"  {s:4d}  | {d['t']:6.3f}    {d['nf_mean']:.4f}   {d['nf_even']:.4f}   {d['nf_odd']:.4f}   "
                f"{d['E2_mean']:.4f}")


def run_quench_A_ED(nx=ED_NX, ny=ED_NY, dg=ED_DG, g=ED_G_COUP, t=ED_T_HOP, m=ED_M_MASS, dt=0.02, n_steps=2,
                    results_dir="results", label="finite_ed_quench_A", ctx=None, log=print, measure_every=10):
    """run_quench_A_ED (finite_ed.jl:491-551): the string state (n_f(2,
    iy_str) = 1, E_R(1, iy_str) = 1) evolved in its sector."""
    ed = peps.ExactDiag(nx, ny, dg, string_charges(nx, ny), ctx=ctx).hamiltonian(g, t, m)
    iy = ny // 2 - 1
    n_f = np.zeros((nx, ny), np.int64)
    n_f[1, iy] = 1
    e_R = np.zeros((nx - 1, ny), np.int64)
    e_R[0, iy] = 1
    psi0 = basis_state(ed, n_f, e_R, np.zeros((nx, ny - 1), np.int64))
    data, _, _ = ed.evolve(psi0, dt, n_steps)
    _log_rows(data, log, measure_every)
    return data, data_to_csv(data, os.path.join(results_dir, f"{label}_data.csv"))


def _quench_from_ground_state(nx, ny, dg, g0, m0, g1, m1, t, dt, n_steps, results_dir, label, ctx, log,
                              measure_every):
    gch = np.asarray(lattice.staggered_charges(nx, ny))
    ed = peps.ExactDiag(nx, ny, dg, gch, ctx=ctx).hamiltonian(g0, t, m0)
    E0, psi0 = ed.ground_state()        # ite_ground_state_ED (finite_ed.jl:349-359)
    if log:
        log(f"    E0 = {E0:.8f}")
    ed.hamiltonian(g1, t, m1)
    data, _, _ = ed.evolve(psi0.astype(np.complex128), dt, n_steps)
    _log_rows(data, log, measure_every)
    return data, data_to_csv(data, os.path.join(results_dir, f"{label}_data.csv"))


def run_quench_B_ED(nx=ED_NX, ny=ED_NY, dg=ED_DG, g=ED_G_COUP, t=ED_T_HOP, m_init=5.0, m_final=0.1, dt=0.02,
                    n_steps=2, results_dir="results", label="finite_ed_quench_B", ctx=None, log=print,
                    measure_every=10):
    """run_quench_B_ED (finite_ed.jl:553-599): the staggered sector's ground
    state at m_init evolved with m_final."""
    return _quench_from_ground_state(nx, ny, dg, g, m_init, g, m_final, t, dt, n_steps, results_dir, label, ctx,
                                     log, measure_every)


def run_quench_C_ED(nx=ED_NX, ny=ED_NY, dg=ED_DG, t=ED_T_HOP, m=ED_M_MASS, g_init=2.0, g_final=0.5, dt=0.02,
                    n_steps=2, results_dir="results", label="finite_ed_quench_C", ctx=None, log=print,
                    measure_every=10):
    """run_quench_C_ED (finite_ed.jl:605-640): the ground state at g_init
    evolved with g_final."""
    return _quench_from_ground_state(nx, ny, dg, g_init, m, g_final, m, t, dt, n_steps, results_dir, label, ctx,
                                     log, measure_every)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("quench_id", type=int, nargs="?", default=1)
    ap.add_argument("--dt", type=float, default=0.02)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--results", default="results")
    a = ap.parse_args(argv)
    os.makedirs(a.results, exist_ok=True)
    run = {1: run_quench_A_ED, 2: run_quench_B_ED, 3: run_quench_C_ED}.get(a.quench_id)
    if run is None:
        raise SystemExit("quench_id must be 1 (A), 2 (B), or 3 (C)")
    t0 = time.time()
    _, path = run(dt=a.dt, n_steps=a.steps, results_dir=a.results, log=lambda s: print(s, flush=True))
    print(f"  Saved: {path}  ({time.time() - t0:.1f} s)", flush=True)


if __name__ == "__main__":
    main()
