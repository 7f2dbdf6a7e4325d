"""Benchmark of the IM-SRG(2) commutator hot path on one MI355X.

Workload (default): the commutator microbenchmark -- channel structure of
N=20 electrons in R=14 shells (210 spin-orbitals, all (M_L, M_S) blocks), two
seeded random Hermitian operators X, Y with one-body entries N(0,1) and
two-body block entries N(0,1)/sqrt(block dim).  One step = one full FP64
[X, Y] (0/1/2-body, Pandya part with calc_cc-filled CC inputs) with X, Y
already resident in HBM.  The reported value is the 2b-2b commutator rate:
the reference's dgemm flop count for commutator_222 at this channel structure
(matrix_mult222 + construct_cc_ints) divided by the time of the whole step.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6e12   # MI355X dense FP64 matrix, FLOP/s
PEAK_HBM = 8.0e12          # MI355X HBM3E, B/s
CLASS_NAMES = ["gemm_products", "gemm_accumulate", "twobody_gather", "onebody", "calc_cc", "other"]


def random_operator(ch, rng, herm):
    """Seeded random operator in the blocked layout (same recipe as the tests)."""
    from tests.pins.layout import random_flat
    return random_flat(ch, rng, herm)


def cpu_baseline(x, y, ch, flops_ref, budget_s=20.0):
    """The FP64 CPU oracle (C + OpenMP) on the host cores, one commutator at a
    time until ~budget_s of CPU work; returns the same metric in TFLOP/s."""
    from oracle.oracle import Oracle
    o = Oracle(ch)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o.commutator(x, 1, y, 1)  # warm
    n, t0 = 0, time.perf_counter()
    while True:
        o.commutator(x, 1, y, 1)
        n += 1
        dt = time.perf_counter() - t0
        if dt > budget_s or n >= 20:
            break
    return dict(value=flops_ref * n / dt / 1e12, unit="TFLOP/s", cores=threads, kind="port",
                sample=f"{n} full [X,Y] at N=20 R=14 (calc_cc + all pieces) by the C oracle, "
                       f"{dt / n * 1e3:.1f} ms each, OMP_NUM_THREADS={threads}")


def breakdown(ctx, X, Y, HCC, ECC, reps=10):
    """Device time of each reference entry point on its own (HIP events)."""
    import qd_imsrg_amd as qd
    Z = ctx.operator(-1)
    pieces = {
        "calc_cc": lambda: qd.calc_cc(Y, HCC),
        "commutator_111": lambda: qd.commutator_111(X, Y, Z),
        "commutator_121": lambda: qd.commutator_121(X, Y, Z),
        "commutator_221": lambda: qd.commutator_221(X, Y, Z),
        "commutator_122": lambda: qd.commutator_122(X, Y, Z),
        "commutator_222": lambda: qd.commutator_222(X, Y, Z, HCC, ECC),
        "commutator_222_no_pandya": lambda: qd.commutator_222(X, Y, Z, None, None),
    }
    res = {}
    for name, fn in pieces.items():
        fn()
        ctx.sync()
        ctx.timing(True)
        for _ in range(reps):
            fn()
        ms, n = ctx.timing_read()
        ctx.timing(False)
        res[name] = {CLASS_NAMES[i]: ms[i] / reps for i in range(6) if n[i]}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nbody", type=int, default=20)
    ap.add_argument("--nshell", type=int, default=14)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--breakdown", action="store_true",
                    help="also time each reference entry point (110..222, calc_cc) on its own")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(the multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import qd_imsrg_amd as qd
    from oracle.channels import Channels  # layout sizes and the reference flop count only

    ch = Channels(a.nbody, a.nshell)
    fl = ch.gemm_flops_2b2b()
    rng = np.random.default_rng(a.seed)
    x = random_operator(ch, rng, 1)
    y = random_operator(ch, rng, 1)

    ctx = qd.Context(a.nbody, a.nshell)
    X, Y, Z = ctx.operator(1, x), ctx.operator(1, y), ctx.operator(-1)
    HCC, ECC = ctx.cc(1.0), ctx.cc(1.0)

    def step():
        qd.calc_cc(Y, HCC)
        qd.calc_cc(X, ECC)
        qd.commutator(X, Y, Z, HCC, ECC)

    for _ in range(a.warmup):
        step()
    ctx.sync()
    ctx.timing(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    ctx.sync()
    t1 = time.perf_counter()
    ms_cls, n_cls = ctx.timing_read()
    ctx.timing(False)
    ms_step = (t1 - t0) * 1e3 / a.steps
    value = fl["total"] / (ms_step * 1e-3) / 1e12

    gf = ctx.gemm_flops()
    per_launch = {CLASS_NAMES[i]: (ms_cls[i] / n_cls[i] if n_cls[i] else 0.0) for i in range(6)}
    # dominant kernel class by device time
    dom = int(np.argmax(ms_cls))
    n2b = ch.n2b()
    if dom == 0:
        achieved = (gf["products"] + gf["pandya"]) / (per_launch["gemm_products"] * 1e-3)
        roof = dict(kernel="k_gemm (ladder intermediates + Pandya products)", bound="mfma",
                    achieved=achieved / 1e12, peak=PEAK_FP64_MFMA / 1e12, unit="TFLOP/s")
    elif dom == 1:
        achieved = gf["accumulate"] / (per_launch["gemm_accumulate"] * 1e-3)
        roof = dict(kernel="k_gemm (accumulating ladder products)", bound="mfma",
                    achieved=achieved / 1e12, peak=PEAK_FP64_MFMA / 1e12, unit="TFLOP/s")
    else:
        # two-body gather: one write and two operand reads per output element,
        # plus its ladder-intermediate and Pandya-product reads (8 + 8*6 bytes)
        byts = {2: 8.0 * n2b * (3 + 4 + 8), 3: 8.0 * 3 * (ch.neq - n2b), 4: 8.0 * 2 * ch.cc_size}.get(dom, 0.0)
        t = per_launch[CLASS_NAMES[dom]] * 1e-3
        roof = dict(kernel=CLASS_NAMES[dom], bound="hbm", achieved=byts / t / 1e9 if t else 0.0,
                    peak=PEAK_HBM / 1e9, unit="GB/s")
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None

    out = {
        "metric": "IMSRG(2) time-to-E_gs (|dE|<=1e-6 Ha vs FP64) and 2b-2b commutator TFLOP/s",
        "value": value, "unit": "TFLOP/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded random Hermitian X, Y)",
        "config": {"workload": f"commutator microbenchmark: full FP64 [X,Y] incl. calc_cc, channel structure "
                               f"N={a.nbody} R={a.nshell} (M={ch.M}, neq={ch.neq})",
                   "flops_2b2b_reference_count": fl["total"], "parallelism": "single GPU"},
        "kernel_ms_per_step": {CLASS_NAMES[i]: ms_cls[i] / a.steps for i in range(6)},
        "gemm_flops_per_step": gf,
        "roofline": roof,
    }
    if a.breakdown:
        out["breakdown_ms"] = breakdown(ctx, X, Y, HCC, ECC)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, y, ch, fl["total"])
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
