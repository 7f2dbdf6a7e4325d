"""Learned warm start on the GPU: the network's forward pass and gradients
against the FP64 CPU oracle (oracle/mlp.py), one Adam step against its
closed form, the PCA basis expansion, and the warm-started Magnus flow
against cold starts on held-out frequencies.

Tolerances: forward max |y - y_oracle| <= 1e-5 max |y_oracle| and gradient
max |g - g_oracle| <= 1e-4 max |g_oracle| (FP32 kernels vs FP64 oracle);
basis expansion 1e-14 relative (FP64); held-out warm flows reach E_gs within
1e-6 Ha of the cold flow and take fewer steps on average."""
import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.mlp import forward as o_forward
from oracle.mlp import loss_grad as o_loss_grad
from qd_imsrg_amd.warmstart import MLP, Basis, WarmStart, warm_flow

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    print("[case] warm start N=6 R=6", flush=True)
    c = qd.Context(6, 6)
    yield c
    c.close()


def test_forward_and_gradients_match_oracle(ctx):
    net = MLP(ctx, 2, 256, 3, 47, seed=7)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((48, 2)).astype(np.float32)
    t = rng.standard_normal((48, 47)).astype(np.float32)
    p = net.params().astype(np.float64)
    y = net.forward(x)
    y_ref = o_forward(p, net.widths, x.astype(np.float64))[-1]
    assert np.abs(y - y_ref).max() <= 1e-5 * np.abs(y_ref).max()
    loss, g = net.loss_grad(x, t)
    loss_ref, g_ref = o_loss_grad(p, net.widths, x.astype(np.float64), t.astype(np.float64))
    print(f"loss {loss:.6e} vs oracle {loss_ref:.6e}; grad max rel {np.abs(g - g_ref).max() / np.abs(g_ref).max():.2e}")
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert np.abs(g - g_ref).max() <= 1e-4 * np.abs(g_ref).max()
    net.close()


def test_first_adam_step_closed_form(ctx):
    """With zero moments, Adam's first step is p - lr * g / (|g| + 1e-8)."""
    net = MLP(ctx, 2, 16, 2, 3, seed=9)
    rng = np.random.default_rng(10)
    x = rng.standard_normal((5, 2)).astype(np.float32)
    t = rng.standard_normal((5, 3)).astype(np.float32)
    p0 = net.params()
    _, g = net.loss_grad(x, t)
    trace = net.train(x, t, 1, 1e-2)
    p1 = net.params()
    expect = p0.astype(np.float64) - 1e-2 * g / (np.abs(g) + 1e-8)
    assert np.abs(p1 - expect).max() <= 1e-6
    assert trace.shape == (1,)
    trace = net.train(x, t, 300, 1e-2)
    assert trace[-1] < 0.05 * trace[0]
    net.close()


def test_basis_expand(ctx):
    rng = np.random.default_rng(11)
    mean = rng.standard_normal(ctx.neq)
    vecs = rng.standard_normal((3, ctx.neq))
    coef = np.array([0.5, -2.0, 0.25])
    b = Basis(ctx, mean, vecs)
    out = ctx.operator(-1)
    b.expand(coef, out)
    ref = mean + coef @ vecs
    assert np.abs(out.download() - ref).max() <= 1e-14 * np.abs(ref).max()
    b.close()


def test_warm_start_saves_steps(ctx):
    """Held-out frequencies (every 4th of 16 in [0.4, 1.0]): the learned
    warm start (MLP -> PCA coefficients -> Omega, then the Pulay-accelerated
    continuation) against the cold start as shipped and against the
    Pulay-accelerated cold start, which isolates what the network adds."""
    omegas = np.geomspace(0.4, 1.0, 16)
    held = [k for k in range(16) if k % 4 == 2]
    train = [k for k in range(16) if k % 4 != 2]
    cold, h0s = {}, {}
    for k, w in enumerate(omegas):
        h0, _, _ = qd.hf_hamiltonian(6, 6, float(w))
        H0, HS, OM = ctx.operator(1, h0), ctx.operator(1), ctx.operator(-1)
        r, _ = qd.magnus_flow(ctx, H0, HS, OM)
        assert r.converged
        cold[k] = (r.e0, r.iterations, OM.download())
        h0s[k] = h0
    ws = WarmStart(ctx, rank=64, hidden=64, layers=3, seed=1)
    info = ws.fit(omegas[train], [cold[k][2] for k in train], steps=3000, lr=1e-3)
    print("fit:", info)
    assert 1 <= info["rank"] <= len(train) - 1
    n_cold, n_cold_diis, n_warm, dE, e_cold_diis = [], [], [], [], []
    for k in held:
        H0 = ctx.operator(1, h0s[k])
        HS, OM = ctx.operator(1), ctx.operator(-1)
        rd, _ = qd.magnus_flow(ctx, H0, HS, OM, diis=6)
        HS, OM = ctx.operator(1), ctx.operator(-1)
        ws.predict(omegas[k], OM)
        r, _ = qd.magnus_flow(ctx, H0, HS, OM, warm_start=True, diis=6)
        assert r.converged and rd.converged
        n_cold.append(cold[k][1])
        n_cold_diis.append(rd.iterations)
        e_cold_diis.append(rd.e0)
        n_warm.append(r.iterations)
        dE.append(abs(r.e0 - cold[k][0]))
        print(f"omega {omegas[k]:.4f}: cold {cold[k][1]} steps, cold+DIIS {rd.iterations}, "
              f"learned warm+DIIS {r.iterations}, |dE| {dE[-1]:.2e}")
    # the packaged driver: same warm result, and the cold start with the same
    # acceleration as the fallback when the warm attempt is cut off before
    # converging
    k = held[0]
    H0 = ctx.operator(1, h0s[k])
    r, fb, spent = warm_flow(ctx, ws, omegas[k], H0, ctx.operator(1), ctx.operator(-1))
    assert not fb and r.iterations == n_warm[0]
    r, fb, spent = warm_flow(ctx, ws, omegas[k], H0, ctx.operator(1), ctx.operator(-1), max_iter=1)
    assert fb and spent == 1 and r.iterations == n_cold_diis[0] and r.e0 == e_cold_diis[0]
    ws.close()
    assert max(dE) <= 1e-6
    assert np.mean(n_warm) < np.mean(n_cold) and np.mean(n_warm) < np.mean(n_cold_diis)
