"""Learned ALS warm start: the GPU MLP is trained on (bond features,
environment-optimal weights of the SVD rank-one terms) pairs from the plain
full update's bonds, then the
gated warm start is compared with the plain ALS on held-out bonds (sweeps to
the 1e-10 tolerance, final costs, gate pass rate).  The fallback must
reproduce the plain ALS exactly."""
import numpy as np
import pytest

from oracle import peps_oracle as O
from tests._util import su_state, to_device

pytestmark = pytest.mark.gpu

G, T_HOP, M = 1.0, 1.0, 0.5


def _bond_problems(ctx, seeds, tau):
    """(theta, env, labels) of every bond of seeded 3x3 gd=2 D=2 states."""
    from lgt_amd import peps
    out = []
    for seed in seeds:
        go = su_state(3, 3, 2, 2, 3, seed)
        gd = to_device(go, ctx)
        for orient in (0, 1):
            bonds = [(ix, iy) for iy in range(1, 4) for ix in range(1, 3)] if orient == 0 else \
                    [(ix, iy) for iy in range(1, 3) for ix in range(1, 4)]
            for ix, iy in bonds:
                E = (peps.bond_environment_h if orient == 0 else peps.bond_environment_v)(gd, ix, iy, 16)
                H = (O.H_merged_h_site if orient == 0 else O.H_merged_v_site)(ix, iy, 3, 3, 2, G, T_HOP, M)
                th = np.einsum("apk,bqk->apqb", E["RA"], E["RB"])
                th = O.apply_gate(th, O.gate(H, tau), th.shape[1], th.shape[2]).real
                labels = O.phys_labels(ix, iy, 3, 3, 2)[1 if orient == 0 else 2]
                out.append((th, E["env"].real, labels))
    return out


def test_learned_warm_start(ctx):
    from lgt_amd import peps, warmstart as W
    train = _bond_problems(ctx, [1, 2, 3, 4, 5, 6, 7, 8], 0.1)
    held = _bond_problems(ctx, [21, 22, 23], 0.1)
    Xf, Tf = [], []
    for th, env, lab in train:
        X0, Y0, _ = W.svd_init(th, env, 2, lab, ctx)
        Xf.append(W.features(th, env, X0, Y0))
        Tf.append(W.env_optimal_targets(th, env, X0, Y0))
    Xf, Tf = np.array(Xf).T, np.array(Tf).T
    mlp = peps.Mlp(W.N_IN, 512, W.N_OUT, seed=11, ctx=ctx)
    first = mlp.train_step(Xf, Tf, 1e-3)
    for _ in range(400):
        last = mlp.train_step(Xf, Tf, 1e-3)
    print(f"warm start: {Xf.shape[1]} training bonds, loss {first:.3e} -> {last:.3e}", flush=True)
    assert last < 0.5 * first
    sw_plain = sw_warm = passed = 0
    worse = 0
    for th, env, lab in held:
        Xp, Yp, bfp, ip = peps.als_truncate_init(th, env, 2, lab, n_als=30, als_tol=1e-10, ctx=ctx)
        Xw, Yw, bfw, iw = W.warm_als(mlp, th, env, 2, lab, n_als=30, als_tol=1e-10, ctx=ctx)
        np.testing.assert_array_equal(bfp, bfw)
        sw_plain += ip["sweeps"]
        sw_warm += iw["sweeps"]
        passed += iw["warm"]
        if not iw["warm"]:  # the fallback is the plain ALS, bit for bit
            np.testing.assert_array_equal(Xp, Xw)
            np.testing.assert_array_equal(Yp, Yw)
        scale = max(abs(ip["cost0"]), 1e-300)
        if W.env_cost(th, env, Xw, Yw) > W.env_cost(th, env, Xp, Yp) + 1e-9 * scale:
            worse += 1
    n = len(held)
    print(f"warm start on {n} held-out bonds: gate passed {passed}, ALS sweeps {sw_warm} vs plain {sw_plain}, "
          f"final cost worse than plain on {worse}", flush=True)
    # regression bar (the gain is not established at this size): the gated
    # warm start never ends above the plain ALS's cost, and never needs more
    # sweeps in total
    assert worse == 0
    assert sw_warm <= sw_plain
