"""GPU parity of the tau-annealed drivers (ite_ground_state_v4/_v3/_v2) and of
the PEPS_STEP_RENORM boundary normalisation against the CPU oracle."""
import numpy as np
import pytest

from _util import rel, to_device, to_oracle, su_state
from oracle import peps_oracle as O

pytestmark = pytest.mark.gpu
G, T, M = 1.0, 1.0, 0.5


def _scaled(gp, c):
    g2 = O.GaugePEPS(gp.nx, gp.ny, gp.gd, gp.gch)
    g2.tensors = [[t * c for t in col] for col in gp.tensors]
    g2.lh, g2.lv = [list(x) for x in gp.lh], [list(x) for x in gp.lv]
    g2.qh, g2.qv = [list(x) for x in gp.qh], [list(x) for x in gp.qv]
    return g2


def test_trotter_step_v4_renorm_where_the_cuts_differ(ctx):
    """PEPS_STEP_RENORM: on a 3x3 state scaled by 1e-2 per site (<psi|psi> ~
    1e-36) the reference's absolute 1e-14 compression cut collapses the
    boundaries while the relative cut does not.  The GPU step with the flag
    (2 ALS sweeps, so the environments matter) matches the oracle's step with
    renorm=True: labels bit-exact, the energy to 1e-8 (the ALS
    trajectory bar; weights to 1e-6 absolute, the bar of the round-trip full
    update tests).  The oracle's step without the flag gives other bond
    weights, so the two cuts really differ on this state."""
    from lgt_amd import peps
    base = su_state(3, 3, 2, 2, 2, seed=5)
    go = _scaled(base, 1e-2)
    gd = to_device(go, ctx)
    peps.trotter_step_v4(gd, G, T, M, 0.05, 2, 16, 2, renorm=True)
    O.trotter_step_v4(go, G, T, M, 0.05, 2, 16, 2, renorm=True)
    back = to_oracle(gd)
    for ix in range(1, 3):
        for iy in range(1, 4):
            assert np.array_equal(back.qh[ix - 1][iy - 1], go.qh[ix - 1][iy - 1])
            np.testing.assert_allclose(back.LH(ix, iy), go.LH(ix, iy), rtol=0, atol=1e-6)
    e_gpu = peps.compute_energy_v4(gd, G, T, M, 16, renorm=True)
    e_ora = O.compute_energy_v4(go, G, T, M, 16, renorm=True)
    assert abs(e_gpu - e_ora) <= 1e-8 * abs(e_ora)
    g_abs = _scaled(base, 1e-2)
    O.trotter_step_v4(g_abs, G, T, M, 0.05, 2, 16, 2, renorm=False)
    diff = max(np.max(np.abs(np.asarray(g_abs.LH(ix, iy)) - np.asarray(go.LH(ix, iy))))
               if len(g_abs.LH(ix, iy)) == len(go.LH(ix, iy)) else 1.0
               for ix in range(1, 3) for iy in range(1, 4))
    assert diff > 1e-6


@pytest.mark.parametrize("kind", ["v3", "v2"])
def test_ite_ground_state_simple_update_drivers(ctx, kind):
    """ite_ground_state_v3 (finite_peps_gauge_invariant.jl:548-577) and
    ite_ground_state_v2 (finite_peps_full_update.jl:582-629) with the tau
    schedule over 16 steps (2, 2, 4, 2, 2, 4 steps at tau 0.1 ... 0.001) from
    their product-state starts (no noise) on 3x3: the GPU state equals the
    oracle's (weights and tensors to 1e-10; labels bit-exact for v3)."""
    from lgt_amd import lattice, peps
    nx = ny = 3
    D, n_ite = 2, 16
    assert [s for _, s in peps.tau_stages(n_ite)] == [2, 2, 4, 2, 2, 4]
    if kind == "v3":
        gd = peps.ite_ground_state_v3(nx, ny, 3, D, G, T, M, lattice.staggered_charges(nx, ny), n_ite=n_ite,
                                      noise=0.0, ctx=ctx)
        go = O.init_gauge_peps(nx, ny, 3, O.staggered_charges(nx, ny))
        O.ite_ground_state_v3(go, G, T, M, D, n_ite)
    else:
        gd = peps.ite_ground_state_v2(nx, ny, 2, D, G, T, M, n_ite=n_ite, noise=0.0, ctx=ctx)
        a = lattice.init_finite_peps_v2_arrays(nx, ny, 2)
        go = O.GaugePEPS(nx, ny, 2, a["gch"])
        go.tensors = [[np.asarray(t, np.complex128) for t in col] for col in a["tensors"]]
        go.lh, go.lv, go.qh, go.qv = a["lh"], a["lv"], a["qh"], a["qv"]
        O.ite_ground_state_v2(go, G, T, M, D, n_ite)
    back = to_oracle(gd)
    for iy in range(1, ny + 1):
        for ix in range(1, nx):
            np.testing.assert_allclose(back.LH(ix, iy), go.LH(ix, iy), rtol=1e-10, atol=1e-12)
            if kind == "v3":
                assert np.array_equal(back.qh[ix - 1][iy - 1], go.qh[ix - 1][iy - 1])
    E_gpu = peps.total_energy(gd, G, T, M) if kind == "v2" else peps.compute_energy_v4(gd, G, T, M, 16)
    E_ora = O.total_energy(go, G, T, M) if kind == "v2" else O.compute_energy_v4(go, G, T, M, 16)
    assert abs(E_gpu - E_ora) <= 1e-10 * max(abs(E_ora), 1.0)


def test_ite_ground_state_v4_driver(ctx):
    """ite_ground_state_v4 (finite_peps_fullupdate_v4.jl:440-467): 8 steps of
    the schedule (1, 1, 2, 1, 1, 2) on 3x3, dg=1, from the zero-flux product
    state, 3 ALS sweeps per bond, chi=16.  The GPU's energy after the run
    matches the oracle's to 1e-8 relative (the ALS trajectory is sensitive to
    rounding at the 1e-9 level), and the Gauss law holds."""
    from lgt_amd import lattice, peps
    nx = ny = 3
    gd = peps.ite_ground_state_v4(nx, ny, 3, 2, G, T, M, lattice.staggered_charges(nx, ny), n_ite=8, noise=0.0,
                                  chi=16, n_als=3, ctx=ctx)
    go = O.init_gauge_peps(nx, ny, 3, O.staggered_charges(nx, ny))
    O.ite_ground_state_v4(go, G, T, M, 2, n_ite=8, chi=16, n_als=3)
    assert peps.forbidden_norm(gd) < 1e-12
    e_gpu = peps.compute_energy_v4(gd, G, T, M, 16)
    e_ora = O.compute_energy_v4(go, G, T, M, 16)
    assert abs(e_gpu - e_ora) <= 1e-8 * abs(e_ora), (e_gpu, e_ora)


def test_batched_sweep_driver(capsys):
    """bench.py --c4 at a tiny size (3 instances, 3x3..4x4, D=2, one step,
    two workers): every instance steps, the JSON line carries instances/s and
    the per-step latency percentiles, and the instance list follows the seed."""
    import argparse
    import json
    import bench
    a = argparse.Namespace(c4_n=3, c4_L="3,4", c4_D="2,2", c4_steps=1, c4_workers=2, seed=5, gd=2,
                           tier="fp64")
    print("c4 driver 3 instances", flush=True)
    bench.run_c4(a)
    line = [ln for ln in capsys.readouterr().out.splitlines() if ln.startswith("{")][-1]
    d = json.loads(line)
    assert d["unit"] == "instances/s" and d["value"] > 0
    assert 0 < d["p50_step_ms"] <= d["p99_step_ms"]
    inst = d["config"]["instances"]
    assert len(inst) == 3 and all(3 <= p["nx"] <= 4 and 3 <= p["ny"] <= 4 and p["D"] == 2 for p in inst)
    assert all(p["chi"] == 16 and 0.1 <= p["m"] <= 2.0 for p in inst)
    assert [p["seed"] for p in inst] == [6, 7, 8]
