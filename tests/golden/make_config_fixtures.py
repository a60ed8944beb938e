"""Writes the golden fixtures of the two CPU-runnable configurations with the
CPU oracle (oracle/peps_oracle.py), here in the build container; the GPU
tests compare the library with them.

  c1: 4x4 U(1) LGT, spin-1/2 links (gd=2), staggered charges, seeded random
      gauge-invariant PEPS at D=2 (lgt_amd.lattice, seed 0), 200 simple-update
      steps (trotter_step_gauge, tau=0.05, t=1, m=0.5, g=1), then the
      boundary-MPS energy (compute_energy_v4) at chi=16.
  c2: 6x6, D=4, the same start (seed 0) and 300 simple-update steps; the
      state is then rescaled site-uniformly to <psi|psi> = 1 (the physical
      state is unchanged; the reference's absolute compression cut needs an
      O(1) norm) and the norm, the energy and every plaquette <W> are
      evaluated at chi=64.
  v4: the full-update self-test of finite_peps_fullupdate_v4.jl:479-504 at
      its default settings: 3x4, gd=3 (dg=1), staggered charges,
      init_gauge_peps(noise=0.1, seed 1), eight trotter_step_v4 steps
      (g=1, t=1, m=0.25, tau=0.05, D_max=4, chi=64, 30 ALS sweeps); the
      energy (compute_energy_v4, chi=64), <n_f> and the Gauss-law violation
      after every step.
  v4spread: the same run from the same start with every site tensor
      multiplied entrywise by (1 + 1e-14 r), r ~ N(0,1) from seeds 1 and 2
      (v4_spread_<seed>.npz): the oracle's own sensitivity to rounding-level
      input changes, which bounds how far two correct FP64 runs of the
      chaotic ALS trajectory can drift apart.

Usage: python tests/golden/make_config_fixtures.py c1|c2|v4|v4spread [seed]
(c2 takes about an hour of CPU time, v4 about half an hour per run)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from lgt_amd import lattice  # noqa: E402  (host-side initial state, numpy only)
from oracle import peps_oracle as O  # noqa: E402

G, T, M, TAU = 1.0, 1.0, 0.5, 0.05


def start(nx, D, seed=0):
    gch = lattice.staggered_charges(nx, nx)
    a = lattice.random_gauge_peps_arrays(nx, nx, 2, gch, D, seed)
    gp = O.GaugePEPS(nx, nx, 2, gch)
    gp.tensors = [[np.asarray(t, np.complex128) for t in col] for col in a["tensors"]]
    gp.lh, gp.lv, gp.qh, gp.qv = a["lh"], a["lv"], a["qh"], a["qv"]
    return gp


def pack(prefix, gp, out):
    for iy in range(1, gp.ny + 1):
        for ix in range(1, gp.nx + 1):
            out[f"{prefix}T_{ix}_{iy}"] = np.real(gp.T(ix, iy)).copy()
    for iy in range(1, gp.ny + 1):
        for ix in range(1, gp.nx):
            out[f"{prefix}lh_{ix}_{iy}"] = np.asarray(gp.LH(ix, iy))
            out[f"{prefix}qh_{ix}_{iy}"] = np.asarray(gp.qh[ix - 1][iy - 1], np.int64)
    for iy in range(1, gp.ny):
        for ix in range(1, gp.nx + 1):
            out[f"{prefix}lv_{ix}_{iy}"] = np.asarray(gp.LV(ix, iy))
            out[f"{prefix}qv_{ix}_{iy}"] = np.asarray(gp.qv[ix - 1][iy - 1], np.int64)


def rescale(gp, chi):
    f_tot = 1.0
    for _ in range(4):
        n2 = O.peps_norm2(gp, chi).real
        f = n2 ** (-0.5 / (gp.nx * gp.ny))
        f_tot *= f
        gp.tensors = [[t * f for t in col] for col in gp.tensors]
        if abs(np.log10(n2)) < 1e-6:
            break
    return f_tot


def v4_selftest(perturb_seed=None):
    gp = O.init_gauge_peps(3, 4, 3, O.staggered_charges(3, 4), noise=0.1, seed=1)
    if perturb_seed is not None:
        rng = np.random.default_rng(perturb_seed)
        gp.tensors = [[t * (1.0 + 1e-14 * rng.standard_normal(t.shape)) for t in col] for col in gp.tensors]
    E, nf, forbidden = [], [], []
    for s in range(8):
        O.trotter_step_v4(gp, 1.0, 1.0, 0.25, 0.05, 4, 64, 30)
        E.append(O.compute_energy_v4(gp, 1.0, 1.0, 0.25, 64))
        nf.append(O.mean_nf(gp))
        forbidden.append(O.forbidden_norm(gp))
        print(f"v4: step {s + 1}/8 E = {E[-1]:.15f} <n_f> = {nf[-1]:.12f}", flush=True)
    name = "v4_selftest.npz" if perturb_seed is None else f"v4_spread_{perturb_seed}.npz"
    np.savez_compressed(os.path.join(HERE, name), energy=np.array(E), mean_nf=np.array(nf),
                        forbidden_norm=np.array(forbidden), nx=3, ny=4, gd=3, noise=0.1, seed=1,
                        g=1.0, t=1.0, m=0.25, tau=0.05, D_max=4, chi=64, n_als=30,
                        perturbation=(0.0 if perturb_seed is None else 1e-14))
    print(f"v4: wrote {name}", flush=True)


def main(which, arg=None):
    if which == "v4":
        return v4_selftest()
    if which == "v4spread":
        return v4_selftest(int(arg or 1))
    if which == "c1":
        nx, D, steps, chi = 4, 2, 200, 16
    else:
        nx, D, steps, chi = 6, 4, 300, 64
    gp = start(nx, D)
    for s in range(steps):
        O.trotter_step_gauge(gp, G, T, M, TAU, D)
        if (s + 1) % 25 == 0:
            print(f"{which}: SU step {s + 1}/{steps}", flush=True)
    out = dict(nx=nx, D=D, steps=steps, chi=chi, g=G, t=T, m=M, tau=TAU)
    pack("su_", gp, out)
    if which == "c1":
        out["energy"] = O.compute_energy_v4(gp, G, T, M, chi)
    else:
        out["scale"] = rescale(gp, chi)
        pack("n_", gp, out)
        out["norm2"] = O.peps_norm2(gp, chi).real
        out["energy"] = O.compute_energy_v4(gp, G, T, M, chi)
        print(f"{which}: energy done", flush=True)
        W = np.zeros((nx - 1, nx - 1), np.complex128)
        for iy in range(1, nx):
            for ix in range(1, nx):
                W[ix - 1, iy - 1] = O.plaquette_expect(gp, ix, iy, chi)
        out["plaquettes"] = W
    out["forbidden_norm"] = O.forbidden_norm(gp)
    np.savez_compressed(os.path.join(HERE, f"{which}.npz"), **out)
    print(f"{which}: wrote {which}.npz, E/site = {out['energy'] / nx ** 2:.15f}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "c1", sys.argv[2] if len(sys.argv) > 2 else None)
