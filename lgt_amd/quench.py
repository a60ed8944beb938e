"""Real-time quenches of the finite PEPS (finite_peps_quench.jl) on the GPU.

The evolution itself runs in the library: realtime_trotter_step
(PEPS_STEP_REALTIME, finite_peps_quench.jl:81-109) applies exp(-i dt/2 H_h),
exp(-i dt H_v), exp(-i dt/2 H_h) through the dense bond update, and
measure_all_finite reads the observables from the device state.  This module
is the host-side driver around them, mirroring the reference's functions of
the same names: the initial states, the imaginary-time preparation
(ite_ground_state, :200-221), the evolution loop (evolve!, :322-353), the
three quench protocols (run_quench_A/B/C, :366-530) and their output files
(the time-series CSV of data_to_df, :239-250, and the site-resolved snapshot
of save_final_snapshot, :272-283).  Plots are not produced.

Lattice sizes follow the reference: gd = 2 dg + 1 is the link dimension
(the reference's dg = 1 is gd = 3).
"""
from __future__ import annotations

import math
import os
from decimal import Decimal

import numpy as np

from . import lattice, peps

# finite_peps_quench.jl:56-64
QNX, QNY, QDG, QD_BOND, QD_MAX = 3, 4, 1, 4, 12
QG_COUP, QT_HOP, QM_MASS = 1.0, 1.0, 2.0

# data_to_df (finite_peps_quench.jl:239-250): the scalar columns, in order
CSV_COLUMNS = ("t", "nf_mean", "nf_even", "nf_odd", "E2_mean", "S_h_mean", "S_v_mean", "S_mean")


def julia_float(x: float) -> str:
    """A Float64 as Julia prints it (the form CSV.jl writes): the shortest
    round-trip digits, plain decimal for 1e-4 <= |x| < 1e6 (always with a
    fractional part, '2.0'), otherwise d.ddde<exp> ('1.0e-5', '1.5e7')."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign, dig, exp = Decimal(repr(x)).as_tuple()
    digits = "".join(map(str, dig)).lstrip("0")
    e10 = len(digits) - 1 + exp  # value = 0.digits... scaled: d.ddd x 10^e10
    digits = digits.rstrip("0") or "0"
    sg = "-" if sign else ""
    if 1e-4 <= abs(x) < 1e6:
        if e10 >= 0:
            ip = digits[:e10 + 1].ljust(e10 + 1, "0")
            fp = digits[e10 + 1:] or "0"
        else:
            ip, fp = "0", "0" * (-e10 - 1) + digits
        return f"{sg}{ip}.{fp}"
    return f"{sg}{digits[0]}.{digits[1:] or '0'}e{e10}"


def _cell(v) -> str:
    if isinstance(v, (bool, np.bool_)):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return julia_float(float(v))
    s = str(v)
    return f'"{s}"' if any(ch in s for ch in ',"\n') else s


def write_csv(rows, columns, path):
    """CSV.write of a DataFrame: a header line, then one line per row."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w") as f:
        f.write(",".join(columns) + "\n")
        for r in rows:
            f.write(",".join(_cell(r[c]) for c in columns) + "\n")


def data_to_csv(data, path):
    """data_to_df + CSV.write (finite_peps_quench.jl:239-250,397): the scalar
    time series of evolve()."""
    write_csv(data, CSV_COLUMNS, path)


def init_finite_peps(nx, ny, gd, D, noise=0.0, seed=None, ctx=None, nf_of=None) -> peps.GaugePEPS:
    """init_finite_peps (finite_peps_ground_state.jl:122-160) on the device
    (complex, label-free)."""
    a = lattice.init_finite_peps_arrays(nx, ny, gd, D, noise, seed, nf_of)
    return peps.GaugePEPS.from_arrays(a["tensors"], a["lh"], a["lv"], a["qh"], a["qv"], a["gch"], gd,
                                      is_complex=True, ctx=ctx)


def vacuum_state(nx, ny, gd, D, ctx=None) -> peps.GaugePEPS:
    """The vacuum of run_quench_A (finite_peps_quench.jl:378-387): n_f = 0
    and E = 0 on every site, at the bond dimensions of init_finite_peps."""
    return init_finite_peps(nx, ny, gd, D, ctx=ctx, nf_of=lambda ix, iy: 0)


def string_state(nx, ny, gd, D, ctx=None) -> peps.GaugePEPS:
    """run_quench_A's initial state (finite_peps_quench.jl:378-399): the
    vacuum, U on the right link of (1, ny/2) and c^dagger on (2, ny/2)."""
    gp = vacuum_state(nx, ny, gd, D, ctx)
    iy = ny // 2
    peps.apply_site_op(gp, 1, iy, lattice.site_operators(1, iy, nx, ny, gd)["U_R"])
    peps.apply_site_op(gp, 2, iy, lattice.site_operators(2, iy, nx, ny, gd)["cdag"])
    return gp


def ite_ground_state(nx, ny, gd, D_bond, D_max, g, t_hop, m, tau_ite, n_ite, mu=0.0, ctx=None,
                     log=None) -> peps.GaugePEPS:
    """ite_ground_state (finite_peps_quench.jl:200-221): n_ite dense
    trotter_step! steps at fixed tau from init_finite_peps(D_bond)."""
    gp = init_finite_peps(nx, ny, gd, D_bond, ctx=ctx)
    for step in range(1, n_ite + 1):
        peps.trotter_step(gp, g, t_hop, m, tau_ite, D_max, mu=mu)
        if log and (step % 100 == 0 or step == n_ite):
            D_now = max([len(gp.bond(peps.PEPS_H, ix, iy)[0]) for iy in range(1, ny + 1) for ix in range(1, nx)],
                        default=1)
            log(f"    ITE step {step:4d} / {n_ite}:  <n_f> = {peps.mean_nf(gp):.5f}   D = {D_now}")
    return gp


def evolve(gp, g, t_hop, m, dt, n_steps, D_trunc, measure_every=10, log=None, progress=None):
    """evolve! (finite_peps_quench.jl:322-353): measure at t = 0, then after
    every realtime_trotter_step; returns the list of measure_all_finite
    dicts (t, scalar observables, nf_grid, E2_grid).  `log` gets the
    reference's table rows every measure_every steps; `progress` (optional)
    is called after every step with (step, t, observables)."""
    data = [peps.measure_all_finite(gp, 0.0)]
    for step in range(1, n_steps + 1):
        peps.realtime_trotter_step(gp, g, t_hop, m, dt, D_trunc)
        obs = peps.measure_all_finite(gp, step * dt)
        data.append(obs)
        if progress:
            progress(step, step * dt, obs)
        if log and (step % measure_every == 0 or step == n_steps):
            D_now = max([len(gp.bond(peps.PEPS_H, ix, iy)[0]) for iy in range(1, gp.ny + 1)
                         for ix in range(1, gp.nx)], default=1)
            log(f"  {step:4d}  | {obs['t']:6.3f}    {obs['nf_mean']:.4f}   {obs['nf_even']:.4f}   "
                f"{obs['nf_odd']:.4f}   {obs['E2_mean']:.4f}   {obs['S_mean']:.4f}   {D_now}")
    return data


def format_grid(grid, label):
    """print_grid (finite_peps_quench.jl:255-266): rows iy from the top."""
    nx, ny = grid.shape
    out = [f"  {label}  (rows = iy top→bottom, cols = ix left→right):"]
    for iy in range(ny, 0, -1):
        out.append(f"    iy={iy}:  " + "".join(f"{grid[ix - 1, iy - 1]:8.4f} " for ix in range(1, nx + 1)))
    return "\n".join(out) + "\n"


def save_final_snapshot(gp, gd, t_final, filename):
    """save_final_snapshot (finite_peps_quench.jl:272-283): the n_f and E^2
    grids as text."""
    obs = peps.measure_all_finite(gp, t_final)
    dg = (gd - 1) // 2
    with open(filename, "w") as f:
        f.write(f"Finite PEPS quench — final snapshot at t = {julia_float(t_final)}\n")
        f.write(f"Lattice: {gp.nx}×{gp.ny}  dg={dg}\n\n")
        f.write(format_grid(obs["nf_grid"], "⟨n_f⟩"))
        f.write("\n")
        f.write(format_grid(obs["E2_grid"], "⟨E²⟩"))


def _finish(gp, gd, data, results_dir, label):
    os.makedirs(results_dir, exist_ok=True)
    data_to_csv(data, os.path.join(results_dir, f"{label}_data.csv"))
    save_final_snapshot(gp, gd, data[-1]["t"], os.path.join(results_dir, f"{label}_final.txt"))
    return data


def run_quench_A(nx=QNX, ny=QNY, gd=2 * QDG + 1, D_bond=QD_BOND, D_max=QD_MAX, g=QG_COUP, t_hop=QT_HOP,
                 m=QM_MASS, results_dir="results", dt=0.02, n_steps=2, label="finite_peps_quench_A", ctx=None,
                 log=None):
    """run_quench_A (finite_peps_quench.jl:366-412): string breaking."""
    gp = string_state(nx, ny, gd, D_bond, ctx)
    data = evolve(gp, g, t_hop, m, dt, n_steps, D_max, log=log)
    return _finish(gp, gd, data, results_dir, label)


def run_quench_B(nx=QNX, ny=QNY, gd=2 * QDG + 1, D_bond=QD_BOND, D_max=QD_MAX, g=QG_COUP, t_hop=QT_HOP,
                 m_init=5.0, m_final=0.1, mu_ite=1.0, tau_ite=0.05, n_ite=300, dt=0.02, n_steps=2,
                 label="finite_peps_quench_B", results_dir="results", ctx=None, log=None):
    """run_quench_B (finite_peps_quench.jl:422-466): ground state at m_init
    (with mu_ite), evolved with m_final."""
    gp = ite_ground_state(nx, ny, gd, D_bond, D_max, g, t_hop, m_init, tau_ite, n_ite, mu_ite, ctx, log)
    data = evolve(gp, g, t_hop, m_final, dt, n_steps, D_max, log=log)
    return _finish(gp, gd, data, results_dir, label)


def run_quench_C(nx=QNX, ny=QNY, gd=2 * QDG + 1, D_bond=QD_BOND, D_max=QD_MAX, t_hop=QT_HOP, m=QM_MASS,
                 g_init=2.0, g_final=0.5, mu_ite=0.0, tau_ite=0.05, n_ite=300, dt=0.02, n_steps=2,
                 label="finite_peps_quench_C", results_dir="results", ctx=None, log=None):
    """run_quench_C (finite_peps_quench.jl:477-520): ground state at g_init,
    evolved with g_final."""
    gp = ite_ground_state(nx, ny, gd, D_bond, D_max, g_init, t_hop, m, tau_ite, n_ite, mu_ite, ctx, log)
    data = evolve(gp, g_final, t_hop, m, dt, n_steps, D_max, log=log)
    return _finish(gp, gd, data, results_dir, label)
