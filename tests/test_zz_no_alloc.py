"""The allocation-free entry point qd_mbpt2_noalloc, called through the C ABI
and checked against qd_mbpt2 (the same kernels and partition, so bit for
bit): two calls in a row, the second with a smaller channel structure, then
the first again after the scratch holds the smaller call's contents; and the
size check, exercised with a lowered cap (QD_NOALLOC_CAP, read at context
creation)."""
import ctypes as C
import os

import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.channels import Channels
from tests.pins.hamiltonians import model_hamiltonian

pytestmark = pytest.mark.gpu


def both(ctx, op):
    a, b = C.c_double(), C.c_double()
    assert ctx.L.qd_mbpt2(ctx.h, op.h, C.byref(a)) == 0
    rc = ctx.L.qd_mbpt2_noalloc(ctx.h, op.h, C.byref(b))
    assert rc == 0, ctx.L.qd_last_error().decode()
    return a.value, b.value


def test_noalloc_mbpt2_matches_mbpt2():
    print("[case] mbpt2_noalloc N=20 R=14 then N=6 R=6", flush=True)
    big, small = qd.Context(20, 14), qd.Context(6, 6)
    hb = big.operator(1, model_hamiltonian(Channels(20, 14), 11))
    hs = small.operator(1, model_hamiltonian(Channels(6, 6), 12))
    a, b = both(big, hb)
    assert a == b and a < 0.0
    a2, b2 = both(small, hs)
    assert a2 == b2 and a2 < 0.0 and a2 != a
    a3, b3 = both(big, hb)
    assert (a3, b3) == (a, b)
    big.close()
    small.close()


def test_noalloc_mbpt2_rejects_oversized(monkeypatch):
    print("[case] mbpt2_noalloc cap 8, N=6 R=6 (11 partials)", flush=True)
    monkeypatch.setenv("QD_NOALLOC_CAP", "8")
    ctx = qd.Context(6, 6)
    monkeypatch.delenv("QD_NOALLOC_CAP")
    h = ctx.operator(1, model_hamiltonian(Channels(6, 6), 12))
    out = C.c_double(123.0)
    assert ctx.L.qd_mbpt2_noalloc(ctx.h, h.h, C.byref(out)) == -1
    assert "holds 8" in ctx.L.qd_last_error().decode() and out.value == 123.0
    ref = qd.MBPT2(h)
    assert ref < 0.0
    ctx.close()
