"""The GPU MLP of the learned ALS warm start against the oracle MLP: forward
pass and every parameter gradient (FP32 MFMA products, FP64 storage), and
training by the GPU Adam step reduces the loss."""
import numpy as np
import pytest

from oracle import peps_oracle as O

pytestmark = pytest.mark.gpu


def test_mlp_forward_and_gradients(ctx):
    from lgt_amd import peps
    n_in, w, n_out, B = 32, 512, 6, 64
    m = peps.Mlp(n_in, w, n_out, seed=3, ctx=ctx)
    p = m.get_params()
    rng = np.random.default_rng(1)
    X = rng.standard_normal((n_in, B))
    T = rng.standard_normal((n_out, B))
    Y = m.forward(X)
    Yo = O.mlp_forward(p, X, n_in, w, n_out)
    assert np.abs(Y - Yo).max() < 1e-5 * max(1.0, np.abs(Yo).max())
    loss, g = m.gradients(X, T)
    lo, go = O.mlp_loss_grads(p, X, T, n_in, w, n_out)
    print(f"mlp loss gpu {loss:.8f} oracle {lo:.8f}; max |dg| {np.abs(g - go).max():.2e} of {np.abs(go).max():.2e}")
    assert abs(loss - lo) < 1e-5 * max(1.0, lo)
    assert np.abs(g - go).max() < 1e-4 * np.abs(go).max()


def test_mlp_training_reduces_loss(ctx):
    from lgt_amd import peps
    n_in, w, n_out, B = 8, 512, 3, 256
    rng = np.random.default_rng(2)
    A = rng.standard_normal((n_out, n_in)) / np.sqrt(n_in)
    X = rng.standard_normal((n_in, B))
    T = np.sin(A @ X)
    m = peps.Mlp(n_in, w, n_out, seed=5, ctx=ctx)
    first = m.train_step(X, T, 1e-3)
    for _ in range(300):
        last = m.train_step(X, T, 1e-3)
    print(f"mlp training loss {first:.4e} -> {last:.4e}")
    assert last < 0.05 * first
