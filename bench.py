"""Full-update throughput benchmark (one MI355X).

Workload: the full-update configuration — 8x8 open lattice, spin-1/2 links
(gd=2, site dim d=8 in the bulk), PEPS bond D=6, boundary-MPS bond chi=144,
t=1, m=0.5, g^2=1, tau=0.02, ALS at most 20 sweeps with tolerance 1e-10.
The initial state is the seeded random gauge-invariant PEPS (seed 0) brought
to D=6 by 3 simple-update steps (tau=0.05); set-up is not timed.  After 3
steps the bulk bonds are at D=6; after 10 most bonds have dropped to 1-3
(singular values below the 1e-12 charge-SVD cut), which is a different,
much smaller workload.  The JSON line reports the bond dimensions reached.

One step = one trotter_step_v4 (h half-sweep, v sweep, h half-sweep: 168
bond updates with their boundary-MPS environments and ALS solves), timed
between device synchronisations.  The step runs with PEPS_STEP_RENORM (every
boundary MPS is normalised after each row absorption, so the reference's
absolute truncation thresholds do not collapse the chi=144 boundaries of an
8x8 lattice whose <psi|psi> drifts by ~30 orders of magnitude per sweep) and
PEPS_STEP_POSITIVE_COST (an ALS refinement is kept only if its cost is >= 0:
the chi-truncated environments are not positive semidefinite, and accepting
negative-cost bonds zeroes the 8x8 state within a sweep).  Prints ONE JSON
line.

Precision: the default tier is FP64, the reference's precision.  It is also
the fastest tier on this workload: in the FP32 tier the contraction rounding
noise (~1e-7 relative) lies far above the boundary compression's 1e-14 cut,
so every boundary MPS stays at chi=144, and the ALS needs about twice the
sweeps; --tier fp32 --compare-tier fp64 measures both.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "full-update steps/s at 8x8 D=6 chi=144; speedup vs plain FP64 GPU at dE/site<1e-6"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
PEAK_HBM_GBS = 8000.0
# the dominant kernel of each tier: its launches are timed as one class
GEMM_CLASS = {"fp64": "gemm_f64_tiled", "fp32": "gemm_f32_tiled", "bf16": "gemm_bf16"}
# HBM bytes per launch of a kernel class measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE over the timed step of a workload (copy-calibrated factors 2.0
# and 1.0), keyed by (nx, ny, D, chi, gd, tier, class); profiles/ holds the
# summaries.
PROFILED_TRAFFIC = {
    # first full-update step (profiles/gemm_f64_tiled_fu8x8_D6_chi144_hbm_traffic.md):
    # 2.0 x 4.5306e5 KiB fetched + 1.0 x 1.4377e5 KiB written per launch
    (8, 8, 6, 144, 2, "fp64", "gemm_f64_tiled"): 1024.0 * (2.0 * 4.5306e5 + 1.4377e5),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _clocks():
    """(CLOCK_BOOTTIME, CLOCK_MONOTONIC) in ns: logged around the timed region so
    that a kernel trace of this script (rocprofv3 --kernel-trace) can be cut to
    it by timestamp."""
    return time.clock_gettime_ns(time.CLOCK_BOOTTIME), time.clock_gettime_ns(time.CLOCK_MONOTONIC)


def make_state(ctx, nx, ny, gd, D, seed, su_steps, g, t, m):
    from lgt_amd import lattice, peps
    gch = lattice.staggered_charges(nx, ny)
    a = lattice.random_gauge_peps_arrays(nx, ny, gd, gch, D, seed)
    gp = peps.GaugePEPS.from_arrays(a["tensors"], a["lh"], a["lv"], a["qh"], a["qv"], gch, gd, ctx=ctx)
    for s in range(su_steps):
        peps.trotter_step_gauge(gp, g, t, m, 0.05, D)
    ctx.sync()
    return gp


def normalise(gp, nx, ny, chi):
    """Rescale the site tensors so that <psi|psi> = 1 (the physical state is
    unchanged).  The boundary-MPS truncation keeps singular values above an
    absolute 1e-14 and the charge SVD above 1e-12, as the reference does; on
    an unnormalised 8x8 state <psi|psi> is ~1e-100 and every boundary would
    collapse to bond dimension 1.  Each site is first scaled to unit
    weighted norm, then all sites by the common factor <psi|psi>^(-1/2N),
    repeated until <psi|psi> = 1 to 0.1%."""
    from lgt_amd import peps

    def lam(orient, ix, iy):
        return np.sqrt(gp.bond(orient, ix, iy)[0])

    for iy in range(1, ny + 1):
        for ix in range(1, nx + 1):
            T = gp.site(ix, iy)
            W = T.copy()
            if ix > 1: W = W * lam(peps.PEPS_H, ix - 1, iy)[None, :, None, None, None]
            if ix < nx: W = W * lam(peps.PEPS_H, ix, iy)[None, None, :, None, None]
            if iy < ny: W = W * lam(peps.PEPS_V, ix, iy)[None, None, None, :, None]
            if iy > 1: W = W * lam(peps.PEPS_V, ix, iy - 1)[None, None, None, None, :]
            gp.set_site(ix, iy, T / max(np.linalg.norm(W), 1e-300))
    # while <psi|psi> is tiny the boundaries truncate hard and the computed
    # norm is only an estimate; rescaling by it brings the next evaluation
    # into the range where the cutoffs are harmless
    first = None
    for it in range(6):
        n2 = peps.peps_norm2(gp, chi).real
        if not (np.isfinite(n2) and n2 > 0):
            raise RuntimeError(f"normalise: <psi|psi> = {n2}")
        first = n2 if first is None else first
        if abs(np.log10(n2)) < 1e-3:
            break
        f = n2 ** (-0.5 / (nx * ny))
        for iy in range(1, ny + 1):
            for ix in range(1, nx + 1):
                gp.set_site(ix, iy, gp.site(ix, iy) * f)
    return first, peps.peps_norm2(gp, chi).real


def cpu_baseline(gp, a, per_step):
    """The oracle (numpy FP64 complex128 on OpenBLAS, all its threads) timed on
    pieces of the same workload, taken from the device state after the timed
    steps, then scaled to one full-update step.  Interior h-bond
    (nx/2, ny/2):
      * ALS (n_als sweeps, the step's tolerance and accept rules) on the
        library's environment of the bond, and the recanonicalisation (the
        reference's charge SVD of the full d D^3 square Theta): timed at full
        size (their cost does not depend on chi);
      * one spectator step plus the environment closure, and one boundary
        row absorbed and compressed (the reference's QR/SVD sweep on the
        explicit product): timed with the boundaries cut to chi' = 16 and 32
        (environment) and 8 and 12 (row), extrapolated to the step's mean
        boundary bond by the power law through the two timings (exponent
        clamped to [2, 3]).
    One step = (h + v bonds) x h-bond time + (6 ny - 14) x row time; the
    vertical bonds are charged an h-bond's time (a lower bound: their strip
    environments are larger), and the ALS part is rescaled from the sample's
    sweeps to the step's measured sweeps per bond."""
    import threadpoolctl
    from lgt_amd import peps
    from oracle import peps_oracle as O
    threads = max([p.get("num_threads", 1) for p in threadpoolctl.threadpool_info()] or [1])
    nx, ny, D, chi = a.nx, a.ny, a.D, a.chi
    ix, iy = max(1, nx // 2), max(2, ny // 2)
    arr = gp.to_arrays()
    go = O.GaugePEPS(nx, ny, gp.gd, gp.gch)
    go.tensors = [[np.asarray(t, dtype=np.complex128) for t in col] for col in arr["tensors"]]
    go.lh, go.lv, go.qh, go.qv = arr["lh"], arr["lv"], arr["qh"], arr["qv"]
    Eg = peps.bond_environment_h(gp, ix, iy, chi)
    top = [np.asarray(t, np.complex128) for t in peps.boundary_top(gp, iy, chi)]
    bot = [np.asarray(t, np.complex128) for t in peps.boundary_bottom(gp, iy, chi)]
    t_start = time.perf_counter()
    E = {k: np.asarray(v, np.complex128) for k, v in Eg.items()}
    dpL, dpR = E["RA"].shape[1], E["RB"].shape[1]
    _, dgR, dgU = O.site_dims(ix, iy, nx, ny, gp.gd)
    G = O.gate(O.H_merged_h_site(ix, iy, nx, ny, gp.gd, 1.0, 1.0, 0.5), a.tau / 2)
    th = O.apply_gate(np.einsum("apk,bqk->apqb", E["RA"], E["RB"]), G, dpL, dpR)
    eR = [O.decode_phys(p, dgR, dgU, gp.gd)[1] for p in range(1, dpL + 1)]
    t0 = time.perf_counter()
    O.als_truncate(th, E["env"], D, eR, n_als=0, positive_cost=True)
    t_als0 = time.perf_counter() - t0
    t0 = time.perf_counter()
    X, Y, bf, info = O.als_truncate(th, E["env"], D, eR, n_als=a.n_als, return_info=True, positive_cost=True,
                                    als_tol=a.als_tol)
    t_als = time.perf_counter() - t0
    t_sweep = max(t_als - t_als0, 0.0) / max(info["sweeps"], 1)
    log(f"cpu baseline: ALS init {t_als0:.2f} s, {info['sweeps']} sweeps {t_als:.2f} s")
    Afull = np.einsum("luda,apr->plrud", E["QA"], X)
    Bfull = np.einsum("rudb,bpl->plrud", E["QB"], Y)
    O.apply_mask(go, ix, iy); O.apply_mask(go, ix + 1, iy)
    g2 = O.GaugePEPS(nx, ny, gp.gd, gp.gch)
    g2.tensors = [list(col) for col in go.tensors]
    g2.lh, g2.lv, g2.qh, g2.qv = [list(c) for c in go.lh], [list(c) for c in go.lv], \
        [list(c) for c in go.qh], [list(c) for c in go.qv]
    t0 = time.perf_counter()
    O.recanonicalise_h(g2, ix, iy, Afull, Bfull, D)
    t_recan = time.perf_counter() - t0
    log(f"cpu baseline: recanonicalisation {t_recan:.2f} s")
    row = O.row_dl(go, iy)

    def cut(mps, c):
        return [t[:min(t.shape[0], c), :, :min(t.shape[2], c)].copy() for t in mps]

    def env_piece(c):
        tp, bt = cut(top, c), cut(bot, c)
        L = np.ones((1, 1, 1, 1), np.complex128)
        for j in range(1, ix - 1):
            L = O.spectator_left(L, tp[j - 1], bt[j - 1], row[j - 1])
        t0 = time.perf_counter()
        L = O.spectator_left(L, tp[ix - 2], bt[ix - 2], row[ix - 2])
        ts = time.perf_counter() - t0
        R = np.ones((1, 1, 1, 1), np.complex128)
        for j in range(nx, ix + 1, -1):
            R = O.spectator_right(R, tp[j - 1], bt[j - 1], row[j - 1])
        t0 = time.perf_counter()
        O.bond_environment_h(go, ix, iy, c, top=tp, bot=bt, Lenv=L, Renv=R)
        return ts + time.perf_counter() - t0

    def row_piece(c):
        bt = cut(bot, c)
        t0 = time.perf_counter()
        m = O.absorb_row_from_bottom(bt, O.row_dl(go, iy))
        O.compress_mps(m, c, relative=True)
        return time.perf_counter() - t0

    chi_eff = max(1.0, per_step["bmps_bond_mean"])
    CE, CR = (16, 32), (8, 12)
    e1, e2 = env_piece(CE[0]), env_piece(CE[1])
    r1, r2 = row_piece(CR[0]), row_piece(CR[1])
    pe = min(3.0, max(2.0, np.log(max(e2, 1e-9) / max(e1, 1e-9)) / np.log(CE[1] / CE[0])))
    pr = min(3.0, max(2.0, np.log(max(r2, 1e-9) / max(r1, 1e-9)) / np.log(CR[1] / CR[0])))
    t_env = e2 * (chi_eff / CE[1]) ** pe
    t_row = r2 * (chi_eff / CR[1]) ** pr
    log(f"cpu baseline: env {e1:.3f}/{e2:.3f} s at chi' {CE}, row {r1:.3f}/{r2:.3f} s at chi' {CR}")
    bonds = per_step["bonds"]
    sweeps_per_bond = per_step["als_sweeps"] / max(bonds, 1)
    t_bond = t_env + t_als0 + sweeps_per_bond * t_sweep + t_recan
    step_s = bonds * t_bond + (6 * ny - 14) * t_row
    sample_s = time.perf_counter() - t_start
    return dict(value=1.0 / step_s, unit="steps/s", cores=int(threads), kind="port",
                seconds_per_step=step_s,
                sample=(f"oracle (numpy FP64 complex128, {threads} OpenBLAS threads) on the state after the timed "
                        f"steps, {sample_s:.1f} s of CPU work.  h-bond ({ix},{iy}): ALS init {t_als0:.3f} s + "
                        f"{info['sweeps']} sweeps at {t_sweep:.3f} s, recanonicalisation {t_recan:.3f} s (full "
                        f"size); spectator step + environment closure {e1:.3f}/{e2:.3f} s at chi'={CE[0]}/{CE[1]} "
                        f"and one boundary row absorbed+compressed {r1:.3f}/{r2:.3f} s at chi'={CR[0]}/{CR[1]}, "
                        f"extrapolated as chi^{pe:.2f} / "
                        f"chi^{pr:.2f} to the step's mean boundary bond {chi_eff:.1f}: {t_env:.2f} s / "
                        f"{t_row:.2f} s.  Step = {bonds:.0f} bonds x {t_bond:.2f} s (ALS at the step's "
                        f"{sweeps_per_bond:.2f} sweeps/bond; v-bonds charged an h-bond's time, a lower bound) + "
                        f"{6 * ny - 14} rows x {t_row:.2f} s = {step_s:.1f} s"))


def summarize_rocprof(d, log_path=None):
    """Summarise rocprofv3 output under directory d: per-kernel statistics of
    the whole run (*kernel_stats.csv), of the timed region of this script cut
    out of *kernel_trace.csv by the clock window it logs, and per-kernel means
    of collected counters (*counter_collection.csv, timed window only).  gemm
    instantiations are grouped by precision (PREC 0 f64, 1 f32, 2 bf16)."""
    import collections
    import csv
    import glob
    import re

    def gname(name):
        m = re.search(r"gemm_kernel<(\w+), (\d+),", name)
        if m:
            return "gemm_kernel PREC=" + m.group(2)
        for tiled in ("gemm_f32_tiled", "gemm_f64_tiled"):
            if tiled in name:
                return tiled
        return re.sub(r"\(.*", "", name).replace("void ", "")[:80]

    win = None
    for lf in ([log_path] if log_path else []):
        for line in open(lf, errors="replace"):
            m = re.search(r"clock_boottime_ns=(\d+)\.\.(\d+) clock_monotonic_ns=(\d+)\.\.(\d+)", line)
            if m:
                win = [(int(m.group(1)), int(m.group(2))), (int(m.group(3)), int(m.group(4)))]

    def window(rows):
        for t0, t1 in (win or []):
            sel = [r for r in rows if int(r["Start_Timestamp"]) >= t0 and int(r["End_Timestamp"]) <= t1]
            if sel:
                return sel, (t1 - t0) / 1e9
        return None, None

    for f in sorted(glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)):
        rows = sorted(csv.DictReader(open(f)), key=lambda r: -float(r["TotalDurationNs"]))
        tot = sum(float(r["TotalDurationNs"]) for r in rows)
        print(f"whole run ({os.path.basename(f)}): {len(rows)} kernels, {tot / 1e9:.3f} s kernel time")
        agg = collections.defaultdict(lambda: [0, 0.0])
        for r in rows:
            agg[gname(r["Name"])][0] += int(r["Calls"])
            agg[gname(r["Name"])][1] += float(r["TotalDurationNs"])
        for name, (n, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:15]:
            print("  %-80s calls %8d total %10.3f ms avg %9.4f ms %5.1f%%" % (name, n, t / 1e6, t / 1e6 / n, 100 * t / tot))
    for f in sorted(glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)):
        sel, wall = window(list(csv.DictReader(open(f))))
        if not sel:
            continue
        agg = collections.defaultdict(lambda: [0, 0.0])
        for r in sel:
            agg[gname(r["Kernel_Name"])][0] += 1
            agg[gname(r["Kernel_Name"])][1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        tot = sum(v[1] for v in agg.values())
        print(f"timed region: {len(sel)} dispatches, wall {wall:.3f} s, kernel time {tot / 1e9:.3f} s")
        for name, (n, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:30]:
            print("  %-80s calls %8d total %10.3f ms avg %9.4f ms %5.1f%%" % (name, n, t / 1e6, t / 1e6 / n, 100 * t / tot))
    for f in sorted(glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)):
        rows = list(csv.DictReader(open(f)))
        sel, _ = window(rows)
        rows = sel or rows
        agg = collections.defaultdict(lambda: [0.0, 0])
        for r in rows:
            key = (gname(r["Kernel_Name"]), r["Counter_Name"])
            agg[key][0] += float(r["Counter_Value"])
            agg[key][1] += 1
        print(f"counters ({'timed region' if sel else 'whole run'}, {os.path.basename(f)}):")
        for (name, cn), (v, n) in sorted(agg.items(), key=lambda kv: -kv[1][0]):
            print("  %-80s %-10s dispatches %7d mean %.4e KiB total %.4e KiB" % (name, cn, n, v / n, v))


# kernel classes of the library's statistics -> (bound, peak, unit) of the
# hardware resource that bounds them (MI355X published peaks)
def class_roof(name, tier):
    if name.startswith("gemm_f32") or name == "splitk_reduce_f32":
        return "mfma", PEAK_TFLOPS["fp32"], "TFLOP/s"
    if name.startswith("gemm_bf16"):
        return "mfma", PEAK_TFLOPS["bf16"], "TFLOP/s"
    if name.startswith(("permute", "transpose", "mask", "scale", "cast", "copy", "elementwise", "reduce",
                        "splitk")):
        return "hbm", PEAK_HBM_GBS, "GB/s"
    return "mfma", PEAK_TFLOPS["fp64"], "TFLOP/s"


def roofline_of(name, v, tier):
    bound, peak, unit = class_roof(name, tier)
    secs = v["ms"] / 1e3
    if bound == "hbm":
        achieved = v["bytes"] / secs / 1e9 if secs > 0 else 0.0
    else:
        achieved = v["flops"] / secs / 1e12 if secs > 0 else 0.0
    return {"kernel_class": name, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "launches": v["launches"],
            "avg_launch_ms": v["ms"] / max(1, v["launches"]),
            "algorithmic_flop_per_launch": v["flops"] / max(1, v["launches"]),
            "algorithmic_bytes_per_launch": v["bytes"] / max(1, v["launches"])}


def run_steps(gp, ctx, a, count, label, deadline):
    """Run up to `count` full-update steps, stopping early when another step
    (at the mean step time so far) would pass `deadline` (perf_counter);
    returns (steps run, ALS totals, step times)."""
    from lgt_amd import peps
    g, t, m = 1.0, 1.0, 0.5
    als = dict(bonds=0, accepted=0, sweeps=0)
    times = []
    for s in range(count):
        if times and time.perf_counter() + float(np.mean(times)) > deadline:
            log(f"{label}: stopping after {len(times)} of {count} steps (wall-clock budget)")
            break
        ts = time.perf_counter()
        st = peps.trotter_step_v4(gp, g, t, m, a.tau, a.D, a.chi, a.n_als, a.als_tol, renorm=True,
                                  positive_cost=True)
        ctx.sync()
        times.append(time.perf_counter() - ts)
        for k in als:
            als[k] += st[k]
        log(f"{label} step {s + 1}/{count}: {times[-1]:.2f} s  als={st}")
    return len(times), als, times


def run_c4(a):
    """The batched-sweep configuration: independent PEPS instances with
    per-instance sizes and couplings (Lx, Ly uniform in the --c4-L range, D in
    the --c4-D range, chi = min(4 D^2, 144), m uniform in [0.1, 2], g^2 uniform
    in [0.5, 2], seeds base+i), each brought to its D by 3 simple-update steps
    and normalised (not timed), then --c4-steps full-update steps each at
    tau = 0.02 (ALS <= 20 sweeps, tol 1e-10).  Every instance has its own
    context and stream; --c4-workers host threads step instances
    concurrently (the library releases the GIL), each thread an instance at a
    time; the timed phase opens each instance's context, uploads its state,
    runs its steps and closes it.  Prints one JSON line: instances/s over the
    timed phase and the p50/p99 of the per-step wall latencies."""
    from concurrent.futures import ThreadPoolExecutor
    from lgt_amd import lattice, peps
    lo_l, hi_l = [int(x) for x in a.c4_L.split(",")]
    lo_d, hi_d = [int(x) for x in a.c4_D.split(",")]
    rng = np.random.default_rng(a.seed)
    inst = []
    for i in range(a.c4_n):
        lx, ly = int(rng.integers(lo_l, hi_l + 1)), int(rng.integers(lo_l, hi_l + 1))
        D = int(rng.integers(lo_d, hi_d + 1))
        inst.append(dict(i=i, nx=lx, ny=ly, D=D, chi=min(4 * D * D, 144), m=float(rng.uniform(0.1, 2.0)),
                         g=float(np.sqrt(rng.uniform(0.5, 2.0))), seed=a.seed + 1 + i))

    def setup(p):
        # the state goes back to the host and its context is closed, so that
        # at most `workers` device pools exist at a time
        ctx = peps.Context(0)
        try:
            gp = make_state(ctx, p["nx"], p["ny"], a.gd, p["D"], p["seed"], 3, p["g"], 1.0, p["m"])
            normalise(gp, p["nx"], p["ny"], p["chi"])
            return gp.to_arrays()
        finally:
            ctx.close()

    def run(p):
        ctx = peps.Context(0)
        try:
            ctx.set_tier(a.tier)
            s = p["state"]
            gp = peps.GaugePEPS.from_arrays(s["tensors"], s["lh"], s["lv"], s["qh"], s["qv"],
                                            lattice.staggered_charges(p["nx"], p["ny"]), a.gd, ctx=ctx)
            lat = []
            for _ in range(a.c4_steps):
                t0 = time.perf_counter()
                peps.trotter_step_v4(gp, p["g"], 1.0, p["m"], 0.02, p["D"], p["chi"], 20, 1e-10, renorm=True,
                                     positive_cost=True)
                ctx.sync()
                lat.append(time.perf_counter() - t0)
            return lat
        finally:
            ctx.close()

    t0 = time.perf_counter()
    with ThreadPoolExecutor(a.c4_workers) as ex:
        for p, st in zip(inst, ex.map(setup, inst)):
            p["state"] = st
    log(f"c4: {len(inst)} instances set up in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    with ThreadPoolExecutor(a.c4_workers) as ex:
        lats = list(ex.map(run, inst))
    wall = time.perf_counter() - t0
    flat = np.array([x for l in lats for x in l])
    for p, l in zip(inst, lats):
        log(f"c4: instance {p['i']:2d} {p['nx']}x{p['ny']} D={p['D']} chi={p['chi']} m={p['m']:.3f} "
            f"g={p['g']:.3f}: steps {', '.join(f'{x:.2f}' for x in l)} s")
    print(json.dumps({
        "metric": "batched full-update sweep: instances/s and per-step latency",
        "value": len(inst) / wall, "unit": "instances/s", "n_gpus": 1, "higher_is_better": True,
        "p50_step_ms": float(np.percentile(flat, 50) * 1e3), "p99_step_ms": float(np.percentile(flat, 99) * 1e3),
        "steps_per_instance": a.c4_steps, "wall_s": wall,
        "dtype": {"fp64": "f64", "fp32": "f32 (MFMA, FP32 accumulate; FP64 factorisations)",
                  "bf16": "bf16 (MFMA, FP32 accumulate; FP64 factorisations)"}[a.tier],
        "data": "synthetic (seeded random gauge-invariant PEPS per instance, simple-update initialised)",
        "config": {"workload": f"{len(inst)} independent instances, Lx,Ly in [{lo_l},{hi_l}], D in [{lo_d},{hi_d}], "
                               "chi = min(4 D^2, 144), m in [0.1,2], g^2 in [0.5,2], tau 0.02",
                   "workers": a.c4_workers, "tier": a.tier,
                   "instances": [{k: p[k] for k in ("nx", "ny", "D", "chi", "m", "g", "seed")} for p in inst]},
    }), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--max-seconds", type=float, default=420.0,
                    help="wall-clock budget of the warm-up plus timed steps: warm-up steps are cut first, then "
                         "timed steps stop once another one would pass it (the JSON line reports the steps run)")
    ap.add_argument("--tier", default="fp64", choices=["fp64", "fp32", "bf16"])
    ap.add_argument("--cut", type=float, default=None,
                    help="singular-value cut of the boundary compression of the timed run (default 1e-14, the "
                         "reference's; a reduced-precision tier can cut at its own rounding level); a "
                         "--compare-tier run keeps the default")
    ap.add_argument("--nx", type=int, default=8)
    ap.add_argument("--ny", type=int, default=8)
    ap.add_argument("--D", type=int, default=6)
    ap.add_argument("--chi", type=int, default=144)
    ap.add_argument("--gd", type=int, default=2)
    ap.add_argument("--su-steps", type=int, default=3)
    ap.add_argument("--n-als", type=int, default=20)
    ap.add_argument("--als-tol", type=float, default=1e-10)
    ap.add_argument("--tau", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--class-timing", default="gemm", choices=["gemm", "all", "off"],
                    help="HIP-event timing in the timed steps: 'gemm' (default) times the tier's MFMA GEMM class, "
                         "the dominant kernel that `roofline` reports; 'all' times every kernel class for the "
                         "per-class breakdown (`kernel_classes`), at ~4%% of the step; 'off' times the steps only "
                         "(no roofline)")
    ap.add_argument("--energy", action="store_true", help="also report E/site after the timed steps")
    ap.add_argument("--compare-tier", default=None, choices=["fp64", "fp32", "bf16"],
                    help="after the timed steps, run the same steps from the same initial state in this tier "
                         "(own context) and report its step time, the ratio and |dE/site|")
    ap.add_argument("--compare-fp64", action="store_true",
                    help="after the timed steps, run the same steps from the same initial state in the FP64 "
                         "tier (the plain FP64 GPU form) and report its step time, the speedup and |dE/site|")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 5 times (the known-size kernel that "
                         "calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    ap.add_argument("--eigh-bench", metavar="SIZES", default=None,
                    help="only time the real symmetric eigensolver (tridiagonal reduction, divide and conquer, "
                         "reflectors) on seeded random matrices of these comma-separated orders, 20 calls each, "
                         "and exit (the ALS normal matrices of the full update are 48..288)")
    ap.add_argument("--c4", action="store_true",
                    help="only run the batched-sweep configuration (independent instances with per-instance "
                         "sizes and couplings, see run_c4) and print its own JSON line")
    ap.add_argument("--c4-n", type=int, default=64)
    ap.add_argument("--c4-L", default="4,10", help="lattice side range (inclusive)")
    ap.add_argument("--c4-D", default="2,6", help="bond dimension range (inclusive)")
    ap.add_argument("--c4-steps", type=int, default=20)
    ap.add_argument("--c4-workers", type=int, default=8)
    ap.add_argument("--summarize-rocprof", metavar="DIR", default=None,
                    help="only summarise the rocprofv3 output under DIR (with --bench-log: the log of "
                         "the profiled run, whose clock window selects the timed region) and exit")
    ap.add_argument("--bench-log", default=None)
    a = ap.parse_args()
    if a.compare_fp64 and a.compare_tier is None:
        a.compare_tier = "fp64"
    if a.summarize_rocprof:
        summarize_rocprof(a.summarize_rocprof, a.bench_log)
        return
    if a.c4:
        run_c4(a)
        return
    if a.eigh_bench:
        from lgt_amd import peps
        ctx = peps.Context(0)
        rng = np.random.default_rng(1)
        for n in [int(x) for x in a.eigh_bench.split(",")]:
            B = rng.standard_normal((n, n))
            H = (B + B.T)[None]
            w, V = peps.eigh(H, ctx=ctx)
            ref = np.linalg.eigvalsh(H[0])
            err = float(np.abs(np.sort(w[0]) - ref).max() / np.abs(ref).max())
            orth = float(np.abs(V[0].T @ V[0] - np.eye(n)).max())
            t0 = time.perf_counter()
            for _ in range(20):
                peps.eigh(H, ctx=ctx)
            dt = (time.perf_counter() - t0) / 20
            print(json.dumps({"eigh_n": n, "wall_ms": dt * 1e3, "eigval_rel_err": err, "orth_err": orth}),
                  flush=True)
        return
    if a.copy_calibration:
        import torch
        n = a.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_launch": 8 * n, "launches": 5}), flush=True)
        return
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU layout "
              "is a design note, not a measured configuration", file=sys.stderr)
        sys.exit(2)

    from lgt_amd import peps
    g, t, m = 1.0, 1.0, 0.5
    ctx = peps.Context(0)
    log(f"setup: {a.nx}x{a.ny} gd={a.gd} D={a.D} chi={a.chi}, {a.su_steps} SU steps")
    t0 = time.perf_counter()
    gp = make_state(ctx, a.nx, a.ny, a.gd, a.D, a.seed, a.su_steps, g, t, m)
    n_before, n_after = normalise(gp, a.nx, a.ny, a.chi)
    log(f"setup done in {time.perf_counter() - t0:.1f} s (<psi|psi> {n_before:.3e} -> {n_after:.6f} "
        f"after site-uniform rescaling)")
    bd = [len(gp.bond(peps.PEPS_H, ix, iy)[0]) for iy in range(1, a.ny + 1) for ix in range(1, a.nx)] + \
         [len(gp.bond(peps.PEPS_V, ix, iy)[0]) for iy in range(1, a.ny) for ix in range(1, a.nx + 1)]
    bond_dims = {"mean": float(np.mean(bd)), "max": int(max(bd)), "at_D": int(sum(1 for x in bd if x == a.D)),
                 "bonds": len(bd)}
    log(f"PEPS bond dimensions at start: {bond_dims}")
    start = gp.to_arrays() if a.compare_tier else None
    ctx.set_tier(a.tier)
    if a.cut is not None:
        ctx.set_compress_cut(a.cut)
    t_begin = time.perf_counter()
    deadline = t_begin + a.max_seconds
    warm = 0
    wtimes = []
    for w in range(a.warmup):
        # warm-up takes at most a quarter of the budget
        if wtimes and time.perf_counter() + float(np.mean(wtimes)) > t_begin + 0.25 * a.max_seconds:
            log(f"warm-up: stopping after {warm} of {a.warmup} steps (wall-clock budget)")
            break
        n1, _, tw = run_steps(gp, ctx, a, 1, "warm-up", deadline)
        warm += n1
        wtimes += tw
    ctx.stats_reset()
    ctx.set_timing(a.class_timing != "off", gemm_only=a.class_timing == "gemm")
    ctx.sync()
    c0 = _clocks()
    t0 = time.perf_counter()
    steps, als, times = run_steps(gp, ctx, a, a.steps, "timed", deadline)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    c1 = _clocks()
    log(f"timed region clock_boottime_ns={c0[0]}..{c1[0]} clock_monotonic_ns={c0[1]}..{c1[1]}")
    stats = ctx.stats()
    ctx.set_timing(False)
    value = steps / elapsed
    kern = {k: v for k, v in stats.items() if v["ms"] > 0}
    for name, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"]):
        log(f"  {name:24s} launches={v['launches']:8d} ms={v['ms']:10.1f} flop={v['flops']:.3e} "
            f"bytes={v['bytes']:.3e}")
    for name, v in sorted(stats.items()):
        if v["ms"] <= 0:  # counters without kernel time (solve ranks, fallbacks, bond sizes)
            log(f"  counter {name:24s} events={v['launches']:8d} value={v['flops']:.6g} aux={v['bytes']:.6g}")
    ranked = sorted(kern.items(), key=lambda kv: -kv[1]["ms"])
    dominant = ranked[0] if ranked else ("none", dict(flops=0.0, bytes=0.0, ms=0.0, launches=0))
    gcls = GEMM_CLASS[a.tier]
    gk = stats.get(gcls, dict(flops=0.0, ms=0.0, launches=0, bytes=0.0))
    total_flops = sum(v["flops"] for k, v in stats.items() if not k.startswith(("bmps", "top_eigs", "als", "svd[s")))
    bsum = stats.get("bmps[bond]", dict(flops=0.0, bytes=0.0))
    bmax = stats.get("bmps[max bond]", dict(flops=0.0))
    bmps_mean = bsum["flops"] / max(bsum["bytes"], 1.0)
    roof = roofline_of(dominant[0], dominant[1], a.tier)
    roof["traffic"] = PROFILED_TRAFFIC.get((a.nx, a.ny, a.D, a.chi, a.gd, a.tier, dominant[0]))
    if roof["traffic"] is not None:
        roof["traffic_source"] = ("HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE over the first "
                                  "full-update step, copy-calibrated (factors 2.0 and 1.0); profiles/")
    groof = roofline_of(gcls, gk, a.tier)
    groof["traffic"] = PROFILED_TRAFFIC.get((a.nx, a.ny, a.D, a.chi, a.gd, a.tier, gcls))
    if a.class_timing == "off":
        roof = groof = None
    kernel_time = sum(v["ms"] for v in kern.values())
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": 1e3 * elapsed / max(steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 (MFMA, FP32 accumulate; FP64 factorisations)",
                  "bf16": "bf16 (MFMA, FP32 accumulate; FP64 factorisations)"}[a.tier],
        "data": "synthetic (seeded random gauge-invariant PEPS, simple-update initialised)",
        "config": {"workload": f"finite-PEPS full update, {a.nx}x{a.ny} U(1) LGT, spin-1/2 links, D={a.D}, "
                               f"boundary chi<={a.chi}",
                   "nx": a.nx, "ny": a.ny, "D": a.D, "chi_max": a.chi, "gd": a.gd, "tau": a.tau,
                   "steps_requested": a.steps, "warmup_requested": a.warmup, "max_seconds": a.max_seconds,
                   "step_seconds": [round(x, 3) for x in times],
                   "n_als_max": a.n_als, "als_tol": a.als_tol, "tier": a.tier,
                   "compress_cut": a.cut if a.cut is not None else 1e-14,
                   "step_flags": "PEPS_STEP_RENORM|PEPS_STEP_POSITIVE_COST",
                   "bond_updates_per_step": 2 * (a.nx - 1) * a.ny + a.nx * (a.ny - 1),
                   "als_sweeps_per_step": als["sweeps"] / max(1, steps),
                   "als_accept_rate": als["accepted"] / max(1, als["bonds"]),
                   "boundary_bond_mean": bmps_mean,
                   "boundary_bond_max": int(bmax["flops"]),
                   "flop_per_step": total_flops / max(steps, 1),
                   "tflops_per_s": total_flops / max(elapsed, 1e-9) / 1e12,
                   "class_timing": a.class_timing,
                   "kernel_time_share": kernel_time / max(1e-9, 1e3 * elapsed),
                   "peps_bond_dims_at_start": bond_dims,
                   "speedup_vs_plain_fp64_gpu": 1.0 if a.tier == "fp64" else None},
        "roofline": roof,
        "gemm_roofline": groof,
        "kernel_classes": [dict(roofline_of(k, v, a.tier), share_of_step=v["ms"] / max(1e-9, 1e3 * elapsed))
                           for k, v in ranked[:8]],
    }
    e = None
    if a.energy or a.compare_tier:
        e = peps.compute_energy_v4(gp, g, t, m, a.chi, renorm=True)
        out["config"]["energy_per_site"] = e / (a.nx * a.ny)
        log(f"E/site after the timed steps ({a.tier}): {e / (a.nx * a.ny):.12f}")
    if not a.no_cpu_baseline:
        per_step = dict(bonds=als["bonds"] / max(steps, 1), als_sweeps=als["sweeps"] / max(steps, 1),
                        bmps_bond_mean=bmps_mean)
        out["cpu_baseline"] = cpu_baseline(gp, a, per_step)
        log(f"cpu baseline: {out['cpu_baseline']['sample']}")
    if a.compare_tier:
        # a context of its own (an empty Ritz-basis cache), warmed like the
        # timed run's: one norm evaluation of the start state, as normalise()
        # ends with.  The timed run's state and context are released first.
        gch, gd, cplx = gp.gch, gp.gd, gp.is_complex
        gp.__del__()
        ctx.close()
        ctx2 = peps.Context(0)
        g2 = peps.GaugePEPS.from_arrays(start["tensors"], start["lh"], start["lv"], start["qh"], start["qv"],
                                        gch, gd, is_complex=cplx, ctx=ctx2)
        peps.peps_norm2(g2, a.chi)
        ctx2.set_tier(a.compare_tier)
        ctx2.sync()
        t2 = time.perf_counter()
        n2, als2, times2 = run_steps(g2, ctx2, a, warm + steps, a.compare_tier, float("inf"))
        ms2 = 1e3 * float(np.mean(times2[warm:] if len(times2) > warm else times2))
        e2 = peps.compute_energy_v4(g2, g, t, m, a.chi, renorm=True)
        log(f"E/site after the {a.compare_tier} steps: {e2 / (a.nx * a.ny):.12f} ({time.perf_counter() - t2:.1f} s)")
        out["config"][f"{a.compare_tier}_gpu"] = {"ms_per_step": ms2, "time_ratio": ms2 / out["ms_per_step"],
                                                  "energy_per_site": e2 / (a.nx * a.ny),
                                                  "abs_dE_per_site": abs(e - e2) / (a.nx * a.ny),
                                                  "als_sweeps_per_step": als2["sweeps"] / max(1, n2),
                                                  "als_accept_rate": als2["accepted"] / max(1, als2["bonds"])}
        if a.compare_tier == "fp64":
            out["config"]["speedup_vs_plain_fp64_gpu"] = ms2 / out["ms_per_step"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
