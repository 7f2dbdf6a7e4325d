"""Batched form: the grouped persistent GEMM launch, its autotuned tile
shapes, the per-channel launch path and hipGraph-captured BCH terms must all
give the plain path's results (FP64, same kernels => bitwise for the BCH
graphs; the GEMM tile shape changes the MFMA accumulation order, so those
are checked against the oracle to 1e-12)."""
import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.channels import Channels
from oracle.oracle import Oracle
from tests.pins.hamiltonians import model_hamiltonian
from tests.pins.layout import random_flat

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def case():
    print("[case] batched N=20 R=14", flush=True)
    ch = Channels(20, 14)
    ctx = qd.Context(20, 14)
    yield ch, Oracle(ch), ctx
    ctx.close()


def full(ctx, x, y):
    X, Y, Z = ctx.operator(-1, x), ctx.operator(1, y), ctx.operator(1)
    HCC, ECC = ctx.cc(1.0), ctx.cc(-1.0)
    qd.calc_cc(Y, HCC)
    qd.calc_cc(X, ECC)
    qd.commutator(X, Y, Z, HCC, ECC)
    return Z.download()


def test_per_channel_and_autotuned_launches(case):
    ch, o, ctx = case
    rng = np.random.default_rng(31)
    x, y = random_flat(ch, rng, -1), random_flat(ch, rng, 1)
    ref = o.commutator(x, -1, y, 1)
    grouped = full(ctx, x, y)
    ctx.set_launch_mode(1)
    per_channel = full(ctx, x, y)
    ctx.set_launch_mode(0)
    assert np.array_equal(grouped, per_channel)  # same tiles, same order inside each tile
    before, after = ctx.autotune(reps=2)
    print(f"autotune: GEMM launches {before:.3f} ms -> {after:.3f} ms; shapes {np.bincount(ctx.gemm_shapes())}")
    tuned = full(ctx, x, y)
    assert rel(grouped, ref) <= 1e-12 and rel(tuned, ref) <= 1e-12


def test_bch_graphs_bitwise(case):
    ch, o, ctx = case
    rng = np.random.default_rng(32)
    g = random_flat(ch, rng, -1, scale=0.02)
    h = random_flat(ch, rng, 1)
    G, H = ctx.operator(-1, g), ctx.operator(1, h)
    HS1, HS2 = ctx.operator(1), ctx.operator(1)
    n1 = qd.BCH_EXPAND(HS1, G, H)
    ctx.set_graphs(True)
    outs = []
    for _ in range(3):  # capture, then replays with eager work in between
        n2 = qd.BCH_EXPAND(HS2, G, H)
        outs.append(HS2.download())
        qd.add_arrays(H, 1.0, H, 0.0, H)
    ctx.set_graphs(False)
    ref = HS1.download()
    assert all(n == n1 for n in [n2]) and all(np.array_equal(v, ref) for v in outs)


def test_magnus_flow_with_graphs_matches_plain():
    ch = Channels(6, 6)
    h0 = model_hamiltonian(ch, 1)
    ctx = qd.Context(6, 6)
    H0 = ctx.operator(1, h0)
    HS, OM = ctx.operator(1), ctx.operator(-1)
    r1, t1 = qd.magnus_flow(ctx, H0, HS, OM)
    hs1 = HS.download()
    ctx.set_graphs(True)
    HS2, OM2 = ctx.operator(1), ctx.operator(-1)
    r2, t2 = qd.magnus_flow(ctx, H0, HS2, OM2)
    assert (r1.iterations, r1.bch_terms) == (r2.iterations, r2.bch_terms)
    assert np.array_equal(t1, t2) and np.array_equal(HS2.download(), hs1)
    ctx.close()


def test_gemm_shapes_replay_and_mfma_flops(case):
    """Autotuned shapes set on a fresh context give the tuned context's result
    bit for bit; the matrix-core flop count of the accumulating list at 64 x 64
    tiles equals the count from the block dimensions (every tile 64 x 64, K in
    steps of 4, the problems of matrix_mult222's Vpphh / Vphhh / Vppph sums);
    executed >= algorithmic in every list; bad shapes are rejected."""
    ch, o, ctx = case
    rng = np.random.default_rng(33)
    x, y = random_flat(ch, rng, -1), random_flat(ch, rng, 1)
    shapes = ctx.gemm_shapes()
    tuned = full(ctx, x, y)
    fresh = qd.Context(20, 14)
    try:
        fresh.set_gemm_shapes(shapes)
        assert np.array_equal(fresh.gemm_shapes(), shapes)
        assert np.array_equal(full(fresh, x, y), tuned)
        fresh.set_gemm_shapes(np.zeros_like(shapes))
        ex, alg = fresh.gemm_mfma_flops(), fresh.gemm_flops()
        for k in ("products", "pandya", "accumulate"):
            assert ex[k] >= alg[k] > 0
        cd = lambda a, b: -(-a // b)
        k4 = lambda k: 4 * cd(k, 4)
        want = 0.0
        for q in range(ch.nblock):
            h, b, p = int(ch.nh[q]), int(ch.nb[q]), int(ch.np_[q])
            for M, N, ks in ((p, h, (p, h)), (p, h, (p, h)), (b, h, (h, h)), (p, b, (p,)), (p, b, (p,))):
                if M * N:
                    want += cd(M, 64) * cd(N, 64) * 2.0 * 64 * 64 * sum(k4(k) for k in ks)
        assert ex["accumulate"] == want
        print(f"64x64 tiles: executed / algorithmic = " +
              ", ".join(f"{k} {ex[k] / alg[k]:.2f}" for k in ex))
        with pytest.raises(qd.QDError):
            fresh.set_gemm_shapes(shapes[:-1])
        bad = shapes.copy()
        bad[0] = 99
        with pytest.raises(qd.QDError):
            fresh.set_gemm_shapes(bad)
    finally:
        fresh.close()
