"""Precision form: the 2b-2b GEMMs in FP32 or in 2..4 BF16 mantissa slices
(FP32 MFMA accumulation per K stage of 32, FP64 across stages), everything
else FP64.  Errors are measured against the FP64 oracle (oracle/, the CPU
restatement of the reference) and, as a second line, against the plain FP64
GPU form on the same inputs, and must stay inside the stated bounds:

  commutator at N=20 R=14 (random X, Y), max |err| / max |[X,Y]_oracle|:
      fp64 <= 1e-12; fp32, bf16x3, bf16x4 <= 1e-6;  bf16x2 <= 1e-3
  Magnus flow at N=6 R=6 hw=1 (HF basis): |E_gs - E_gs(fp64)| <= 1e-6 Ha for
      fp32, bf16x3, bf16x4; relative error of ||Omega|| <= 1e-4 and
      absolute error of ||Gamma_od|| (small at convergence) <= 1e-5.

The FP64 mode must stay bitwise what it was, and the zero-body part (FP64
traces) is unaffected by the GEMM precision."""
import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.channels import SUB_PPHH, Channels
from oracle.oracle import Oracle
from tests.pins.layout import random_flat

pytestmark = pytest.mark.gpu

BOUND_COMM = {"fp32": 1e-6, "bf16x2": 1e-3, "bf16x3": 1e-6, "bf16x4": 1e-6}
BOUND_E = 1e-6


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def gamma_od_norm(ch, v):
    """Frobenius norm of the off-diagonal part: f_ph and every Vpphh block."""
    s = float(np.sum(v[ch.off_fph:ch.off_fph + ch.mp * ch.n] ** 2))
    for q in range(ch.nblock):
        k = int(ch.off[q, SUB_PPHH])
        s += float(np.sum(v[k:k + int(ch.np_[q]) * int(ch.nh[q])] ** 2))
    return s ** 0.5


def full(ctx, x, y):
    X, Y, Z = ctx.operator(-1, x), ctx.operator(1, y), ctx.operator(1)
    HCC, ECC = ctx.cc(1.0), ctx.cc(-1.0)
    qd.calc_cc(Y, HCC)
    qd.calc_cc(X, ECC)
    e0 = qd.commutator(X, Y, Z, HCC, ECC)
    return e0, Z.download()


def test_commutator_errors_per_precision():
    print("[case] precision N=20 R=14 commutator", flush=True)
    ch = Channels(20, 14)
    ctx = qd.Context(20, 14)
    rng = np.random.default_rng(41)
    x, y = random_flat(ch, rng, -1), random_flat(ch, rng, 1)
    oracle = Oracle(ch).commutator(x, -1, y, 1)
    e_ref, ref = full(ctx, x, y)
    errs, errs_gpu = {"fp64": rel(ref, oracle)}, {}
    for p in ("fp32", "bf16x2", "bf16x3", "bf16x4"):
        ctx.set_precision(p)
        assert ctx.precision() == p
        e0, z = full(ctx, x, y)
        errs[p] = rel(z, oracle)
        errs_gpu[p] = rel(z, ref)
        assert e0 == e_ref  # zero-body traces stay FP64 and do not use the GEMMs
    ctx.set_precision("fp64")
    e_again, again = full(ctx, x, y)
    ctx.close()
    print("commutator max-rel error vs the FP64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()})
    print("                    vs the FP64 GPU form:", {k: f"{v:.2e}" for k, v in errs_gpu.items()})
    assert np.array_equal(again, ref) and e_again == e_ref
    assert errs["fp64"] <= 1e-12
    for p in errs_gpu:
        assert errs[p] <= BOUND_COMM[p] and errs_gpu[p] <= BOUND_COMM[p], (p, errs[p], errs_gpu[p])
    assert errs["bf16x3"] < errs["bf16x2"]


def test_magnus_flow_errors_per_precision():
    print("[case] precision N=6 R=6 Magnus flow", flush=True)
    ch = Channels(6, 6)
    h0, _, _ = qd.hf_hamiltonian(6, 6, 1.0)
    ctx = qd.Context(6, 6)
    H0 = ctx.operator(1, h0)

    def flow(p):
        ctx.set_precision(p)
        HS, OM = ctx.operator(1), ctx.operator(-1)
        res, _ = qd.magnus_flow(ctx, H0, HS, OM)
        return res, qd.mat_2_norm(OM), gamma_od_norm(ch, HS.download())

    r64, om64, od64 = flow("fp64")
    out = {}
    for p in ("fp32", "bf16x2", "bf16x3", "bf16x4"):
        r, om, od = flow(p)
        out[p] = (abs(r.e0 - r64.e0), abs(om - om64) / om64, abs(od - od64), r.iterations)
    ctx.close()
    print(f"E_gs(fp64) = {r64.e0:.10f} in {r64.iterations} steps, ||Gamma_od|| = {od64:.3e}")
    for p, (de, dom, dod, it) in out.items():
        print(f"  {p}: |dE| = {de:.2e} Ha, rel ||Omega|| {dom:.2e}, abs ||Gamma_od|| {dod:.2e}, {it} steps")
    for p in ("fp32", "bf16x3", "bf16x4"):
        de, dom, dod, _ = out[p]
        assert de <= BOUND_E and dom <= 1e-4 and dod <= 1e-5, (p, out[p])


def test_fused_bf16_pandya_matches_product_and_combine():
    """In the grouped form the BF16-slice Pandya product runs fused with its
    combine (upper tiles of D only); the per-channel form keeps the product
    GEMM + combine pass.  Both cut the same slices and sum the same slice
    products in the same order (K stages of 32 from k = 0, FP32 within a
    stage, FP64 across stages), so the results are bitwise equal, and both
    stay inside the oracle bound, with the Pandya term on (commutator_cc:
    calc_cc-consistent sets) and with caller-supplied CC sets."""
    print("[case] fused BF16 Pandya vs product + combine, N=20 R=14", flush=True)
    ch = Channels(20, 14)
    ctx = qd.Context(20, 14)
    rng = np.random.default_rng(43)
    x, y = random_flat(ch, rng, -1), random_flat(ch, rng, 1)
    oracle = Oracle(ch).commutator(x, -1, y, 1)
    X, Y = ctx.operator(-1, x), ctx.operator(1, y)

    def both(mode, p):
        ctx.set_launch_mode(mode)
        ctx.set_precision(p)
        Z1, Z2 = ctx.operator(1), ctx.operator(1)
        HCC, ECC = ctx.cc(1.0), ctx.cc(-1.0)
        qd.calc_cc(Y, HCC)
        qd.calc_cc(X, ECC)
        qd.commutator(X, Y, Z1, HCC, ECC)
        qd.commutator_cc(X, Y, Z2, ctx.cc(1.0), ctx.cc(-1.0))
        return Z1.download(), Z2.download()

    for p in ("bf16x2", "bf16x3", "bf16x4"):
        f1, f2 = both(0, p)
        u1, u2 = both(1, p)
        d1, d2 = rel(f1, u1), rel(f2, u2)
        print(f"  {p}: fused vs unfused {d1:.2e} / {d2:.2e} (bitwise {np.array_equal(f1, u1)}), "
              f"vs oracle {rel(f1, oracle):.2e}", flush=True)
        assert np.array_equal(f1, u1) and np.array_equal(f2, u2), (p, d1, d2)
        for z in (f1, f2, u1, u2):
            assert rel(z, oracle) <= BOUND_COMM[p], p
    ctx.set_launch_mode(0)
    ctx.set_precision("fp64")
    ctx.close()
