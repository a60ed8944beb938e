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

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
# the dominant kernel of each tier: its launches are timed as one class
GEMM_CLASS = {"fp64": "gemm_f64", "fp32": "gemm_f32_tiled", "bf16": "gemm_bf16"}
GEMM_KERNEL = {"fp64": "gemm_kernel (FP64 MFMA, 128x128x16 tiles)", "fp32": "gemm_f32_tiled (FP32 MFMA, 128x128x32 tiles)",
               "bf16": "gemm_kernel (BF16 MFMA)"}
# HBM bytes per launch of the dominant kernel measured with rocprofv3
# FETCH_SIZE and WRITE_SIZE over the timed step of this workload (copy-
# calibrated factors 2.0 and 1.0), keyed by (nx, ny, D, chi, gd, tier);
# profiles/ holds the summary.
PROFILED_TRAFFIC = {(8, 8, 6, 144, 2, "fp32"): (2.0 * 2.2456e6 + 1.0 * 9.2521e5) * 1024.0,
                    (8, 8, 6, 144, 2, "fp64"): (2.0 * 1.2191e5 + 1.0 * 4.6419e4) * 1024.0}


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


def cpu_baseline(flops_per_step, seconds=12.0):
    """numpy (OpenBLAS, all host threads) on the dominant contraction shape of
    the workload — the vertical-bond environment closure, an (a^2 x K) by
    (K x b^2) product with a=b=48 — timed for a bounded ~10 s sample; the
    CPU step time is the library's per-step algorithmic flop count divided
    by that measured rate."""
    from oracle import peps_oracle  # noqa: F401  (the oracle's numpy stack)
    threads = os.cpu_count() or 1
    rng = np.random.default_rng(0)
    M = N = 48 * 48
    K = 144 * 36  # one slice of the chi^2 D^4 closure cut
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    A @ B
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        A @ B
        n += 1
    dt = time.perf_counter() - t0
    rate = 2.0 * M * N * K * n / dt
    return dict(value=rate / max(flops_per_step, 1.0), unit="steps/s", cores=threads, kind="port",
                sample=(f"numpy/OpenBLAS FP64 GEMM {M}x{K}x{N} (v-bond environment closure slice) for "
                        f"{dt:.1f} s = {rate / 1e9:.1f} GFLOP/s on {threads} threads; step time = "
                        f"{flops_per_step:.3e} algorithmic flop/step / rate"))


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
        if "gemm_f32_tiled" in name:
            return "gemm_f32_tiled"
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--tier", default="fp64", choices=["fp64", "fp32", "bf16"])
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

    # one progress line per bond update on stderr (a C3 step runs for minutes)
    os.environ.setdefault("PEPS_HIP_VERBOSE", "1")
    from lgt_amd import peps
    g, t, m = 1.0, 1.0, 0.5
    ctx = peps.Context(0)
    log(f"setup: {a.nx}x{a.ny} gd={a.gd} D={a.D} chi={a.chi}, {a.su_steps} SU steps")
    t0 = time.perf_counter()
    gp = make_state(ctx, a.nx, a.ny, a.gd, a.D, a.seed, a.su_steps, g, t, m)
    n_before, n_after = normalise(gp, a.nx, a.ny, a.chi)
    log(f"setup done in {time.perf_counter() - t0:.1f} s (<psi|psi> {n_before:.3e} -> {n_after:.6f} "
        f"after site-uniform rescaling; bond dims h {[len(gp.bond(peps.PEPS_H, 1, iy)[0]) for iy in range(1, a.ny + 1)]})")
    bd = [len(gp.bond(peps.PEPS_H, ix, iy)[0]) for iy in range(1, a.ny + 1) for ix in range(1, a.nx)] + \
         [len(gp.bond(peps.PEPS_V, ix, iy)[0]) for iy in range(1, a.ny) for ix in range(1, a.nx + 1)]
    bond_dims = {"mean": float(np.mean(bd)), "max": int(max(bd)), "at_D": int(sum(1 for x in bd if x == a.D)),
                 "bonds": len(bd)}
    log(f"bond dimensions: {bond_dims}")
    start = gp.to_arrays() if a.compare_tier else None
    ctx.set_tier(a.tier)
    for w in range(a.warmup):
        tw = time.perf_counter()
        peps.trotter_step_v4(gp, g, t, m, a.tau, a.D, a.chi, a.n_als, a.als_tol, renorm=True,
                                 positive_cost=True)
        ctx.sync()
        log(f"warmup step {w + 1}/{a.warmup}: {time.perf_counter() - tw:.2f} s")
    ctx.stats_reset()
    ctx.set_timing(True)
    ctx.sync()
    c0 = _clocks()
    t0 = time.perf_counter()
    als = dict(bonds=0, accepted=0, sweeps=0)
    for s in range(a.steps):
        ts = time.perf_counter()
        st = peps.trotter_step_v4(gp, g, t, m, a.tau, a.D, a.chi, a.n_als, a.als_tol, renorm=True,
                                 positive_cost=True)
        for k in als:
            als[k] += st[k]
        ctx.sync()
        log(f"step {s + 1}/{a.steps}: {time.perf_counter() - ts:.2f} s  als={st}")
    ctx.sync()
    elapsed = time.perf_counter() - t0
    c1 = _clocks()
    log(f"timed region clock_boottime_ns={c0[0]}..{c1[0]} clock_monotonic_ns={c0[1]}..{c1[1]}")
    stats = ctx.stats()
    ctx.set_timing(False)
    value = a.steps / elapsed
    gk = stats.get(GEMM_CLASS[a.tier], dict(flops=0.0, ms=0.0, launches=0, bytes=0.0))
    achieved = gk["flops"] / (gk["ms"] / 1e3) / 1e12 if gk["ms"] > 0 else 0.0
    total_flops = sum(v["flops"] for v in stats.values())
    svd_ms = sum(v["ms"] for k, v in stats.items() if k.startswith("svd") and k != "svd[sweeps]")
    svd_n = sum(v["launches"] for k, v in stats.items() if k.startswith("svd") and k != "svd[sweeps]")
    for name, v in sorted(stats.items(), key=lambda kv: -kv[1]["ms"]):
        rate = f" {v['bytes'] / v['ms'] / 1e6:8.1f} GB/s" if v["ms"] > 0 and v["bytes"] > 0 else ""
        log(f"  {name:12s} launches={v['launches']:8d} ms={v['ms']:10.1f} flop={v['flops']:.3e}{rate}")
    out = {
        "metric": "full-update steps/s at 8x8 D=6 chi=144",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 (MFMA, FP32 accumulate; FP64 factorisations)",
                  "bf16": "bf16 (MFMA, FP32 accumulate; FP64 factorisations)"}[a.tier],
        "data": "synthetic (seeded random gauge-invariant PEPS, simple-update initialised)",
        "config": {"workload": f"finite-PEPS full update, {a.nx}x{a.ny} U(1) LGT, spin-1/2 links",
                   "nx": a.nx, "ny": a.ny, "D": a.D, "chi": a.chi, "gd": a.gd, "tau": a.tau,
                   "n_als_max": a.n_als, "als_tol": a.als_tol, "tier": a.tier,
                   "step_flags": "PEPS_STEP_RENORM|PEPS_STEP_POSITIVE_COST",
                   "bond_updates_per_step": 2 * (a.nx - 1) * a.ny + a.nx * (a.ny - 1),
                   "als_sweeps_per_step": als["sweeps"] / max(1, a.steps),
                   "als_accept_rate": als["accepted"] / max(1, als["bonds"]),
                   "flop_per_step": total_flops / a.steps,
                   "bond_dims_at_start": bond_dims},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS[a.tier],
                     "unit": "TFLOP/s", "frac": achieved / PEAK_TFLOPS[a.tier],
                     "traffic": PROFILED_TRAFFIC.get((a.nx, a.ny, a.D, a.chi, a.gd, a.tier)),
                     "kernel": GEMM_KERNEL[a.tier],
                     "launches": gk["launches"],
                     "avg_launch_ms": gk["ms"] / max(1, gk["launches"]),
                     "algorithmic_flop_per_launch": gk["flops"] / max(1, gk["launches"]),
                     "algorithmic_bytes_per_launch": gk["bytes"] / max(1, gk["launches"])},
        # the one-sided Jacobi SVD kernels (one workgroup per matrix:
        # latency bound, no HBM or MFMA roofline applies) and their share of
        # the step, beside the GEMM's
        "svd_kernel": {"kernel": "jacobi_svd_reg_batched_kernel / jacobi_svd_reg_kernel", "bound": "latency",
                       "ms_per_step": svd_ms / a.steps, "launches": svd_n,
                       "share_of_step": svd_ms / max(1e-9, 1e3 * elapsed),
                       "gemm_share_of_step": gk["ms"] / max(1e-9, 1e3 * elapsed)},
    }
    if a.energy or a.compare_tier:
        e = peps.compute_energy_v4(gp, g, t, m, a.chi, renorm=True)
        out["config"]["energy_per_site"] = e / (a.nx * a.ny)
        log(f"E/site after the timed steps ({a.tier}): {e / (a.nx * a.ny):.12f}")
    if a.compare_tier:
        # a context of its own (an empty Ritz-basis cache), warmed like the
        # timed run's: one norm evaluation of the start state, as normalise()
        # ends with.  The timed run's state and context (and the memory its
        # pool caches) are released first.
        gch, gd, cplx = gp.gch, gp.gd, gp.is_complex
        gp.__del__()
        ctx.close()
        ctx64 = peps.Context(0)
        g64 = peps.GaugePEPS.from_arrays(start["tensors"], start["lh"], start["lv"], start["qh"], start["qv"],
                                         gch, gd, is_complex=cplx, ctx=ctx64)
        peps.peps_norm2(g64, a.chi)
        ctx64.set_tier(a.compare_tier)
        ctx64.sync()
        t64 = time.perf_counter()
        als64 = dict(bonds=0, accepted=0, sweeps=0)
        for s64 in range(a.warmup + a.steps):
            st64 = peps.trotter_step_v4(g64, g, t, m, a.tau, a.D, a.chi, a.n_als, a.als_tol, renorm=True,
                                        positive_cost=True)
            for k in als64:
                als64[k] += st64[k]
            ctx64.sync()
            log(f"{a.compare_tier} step {s64 + 1}/{a.warmup + a.steps}: {time.perf_counter() - t64:.2f} s  als={st64}")
        ms64 = 1e3 * (time.perf_counter() - t64) / (a.warmup + a.steps)
        e64 = peps.compute_energy_v4(g64, g, t, m, a.chi, renorm=True)
        log(f"E/site after the {a.compare_tier} steps: {e64 / (a.nx * a.ny):.12f}")
        out["config"][f"{a.compare_tier}_gpu"] = {"ms_per_step": ms64, "time_ratio": ms64 / out["ms_per_step"],
                                     "energy_per_site": e64 / (a.nx * a.ny),
                                     "abs_dE_per_site": abs(e - e64) / (a.nx * a.ny),
                                     "als_sweeps_per_step": als64["sweeps"] / max(1, a.warmup + a.steps),
                                     "als_accept_rate": als64["accepted"] / max(1, als64["bonds"])}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(total_flops / a.steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
