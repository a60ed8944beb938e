"""Learned warm start of the full-update ALS bond solve (no counterpart in the
reference).

The reference starts the ALS (finite_peps_fullupdate_v4.jl:99-111) from the
charge-resolved truncated SVD of the gated two-site tensor theta:
X = U sqrt(S), Y = sqrt(S) V^H.  The warm start keeps those singular vectors
and their sector labels and rescales the kept weights: a small MLP
(``peps.Mlp``, trained and evaluated on the GPU) maps features of the bond --
the top ``K`` singular values of theta and the environment norms of the
``K`` rank-one terms U_k S_k V_k^H, each normalised -- to log-ratios r_k
between the environment-optimal weights of those rank-one terms (the
training targets, ``env_optimal_targets``) and the SVD weights.

Fidelity gate: the proposed initial bond is used only if its environment
cost ||theta - X Y||_env is not above that of the SVD initialisation; the
ALS then runs from it and its result is kept by the reference's own rule
(cost below the SVD initialisation's).  Otherwise the plain ALS runs from
the SVD initialisation, unchanged.
"""
from __future__ import annotations

import numpy as np

from . import peps

K = 8                 # singular values / rank-one terms in the features
N_IN, N_OUT = 2 * K, K


def env_inner(env, A, B):
    """<A|env|B> for two-site tensors A, B (a, pL, pR, b): sum env[a,b,c,d] A[a,p,q,b] conj(B[c,p,q,d])."""
    return np.einsum("abcd,apqb,cpqd->", env, A, B.conj(), optimize=True)


def product(X, Y):
    """X (a, pL, k) Y (b, pR, k) -> (a, pL, pR, b)."""
    return np.einsum("apk,bqk->apqb", X, Y, optimize=True)


def env_cost(theta, env, X, Y):
    """||theta - X Y||^2 in the environment metric (the ALS cost)."""
    d = theta - product(X, Y)
    return float(env_inner(env, d, d).real)


def svd_init(theta, env, D_max, labels, ctx=None):
    """The reference's initialisation (X0, Y0, bond flux) from the library."""
    X0, Y0, bf, _ = peps.als_truncate_init(theta, env, D_max, labels, n_als=0, ctx=ctx)
    return X0, Y0, bf


def weights(X, Y):
    """Weights of the rank-one terms of an SVD-shaped bond: |X_k| |Y_k|."""
    return np.array([np.linalg.norm(X[:, :, k]) * np.linalg.norm(Y[:, :, k]) for k in range(X.shape[2])])


def features(theta, env, X0, Y0):
    """2K features: the SVD weights s_k / s_0 and the environment norms of the
    rank-one terms <t_k|env|t_k> / <theta|env|theta>, zero padded."""
    s = weights(X0, Y0)
    nth = max(float(env_inner(env, theta, theta).real), 1e-300)
    e = np.array([env_inner(env, product(X0[:, :, k:k + 1], Y0[:, :, k:k + 1]),
                            product(X0[:, :, k:k + 1], Y0[:, :, k:k + 1])).real for k in range(X0.shape[2])])
    f = np.zeros(N_IN)
    n = min(K, len(s))
    f[:n] = s[:n] / max(s[0], 1e-300)
    f[K:K + n] = e[:n] / nth
    return f


def env_optimal_targets(theta, env, X0, Y0):
    """Log-ratios of the environment-optimal weights of the SVD rank-one terms
    to the SVD weights: w* minimises ||theta - sum_k w_k t_k||_env with
    t_k = X0_k (x) Y0_k / s_k, i.e. G w* = b with G_kl = <t_l|env|t_k>,
    b_k = <theta|env|t_k>.  Clipped to [-5, 5], zero padded."""
    Dp = X0.shape[2]
    s = weights(X0, Y0)
    t = [product(X0[:, :, k:k + 1], Y0[:, :, k:k + 1]) / max(s[k], 1e-300) for k in range(Dp)]
    Gm = np.array([[env_inner(env, t[l], t[k]) for l in range(Dp)] for k in range(Dp)])
    b = np.array([env_inner(env, theta, t[k]) for k in range(Dp)])
    w = np.linalg.lstsq(Gm.real, b.real, rcond=1e-12)[0]
    out = np.zeros(N_OUT)
    n = min(K, Dp)
    out[:n] = np.clip(np.log(np.maximum(np.abs(w[:n]), 1e-300) / np.maximum(s[:n], 1e-300)), -5.0, 5.0)
    return out


def rescale(X0, Y0, r):
    """The SVD initialisation with weight k multiplied by exp(r_k)."""
    X, Y = X0.copy(), Y0.copy()
    for k in range(min(X0.shape[2], len(r))):
        f = np.exp(0.5 * float(np.clip(r[k], -5.0, 5.0)))
        X[:, :, k] *= f
        Y[:, :, k] *= f
    return X, Y


def warm_als(mlp, theta, env, D_max, labels, n_als=30, als_tol=1e-10, ctx=None):
    """ALS with the learned warm start behind the fidelity gate.  Returns
    (X, Y, bflux, info) with info['warm'] telling whether the proposal passed."""
    X0, Y0, _ = svd_init(theta, env, D_max, labels, ctx)
    r = mlp.forward(features(theta, env, X0, Y0)[:, None])[:, 0]
    Xw, Yw = rescale(X0, Y0, r)
    if env_cost(theta, env, Xw, Yw) <= env_cost(theta, env, X0, Y0):
        out = peps.als_truncate_init(theta, env, D_max, labels, n_als=n_als, als_tol=als_tol, init=(Xw, Yw), ctx=ctx)
        out[3]["warm"] = True
        return out
    out = peps.als_truncate_init(theta, env, D_max, labels, n_als=n_als, als_tol=als_tol, ctx=ctx)
    out[3]["warm"] = False
    return out
