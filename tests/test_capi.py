"""C-ABI library checks that need no GPU: it loads, exports every symbol the
header declares, and its host channel builder equals the oracle's."""
import os
import re

import numpy as np
import pytest

from oracle.channels import Channels
from qd_imsrg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "qdimsrg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qd_[a-z0-9_]+)\s*\(", txt)))


def test_library_loads_and_exports_header():
    L = _lib.load()
    names = header_functions()
    assert len(names) >= 40
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def _table(N, R, which):
    L = _lib.load()
    n = L.qd_layout_table(N, R, which, None, 0)
    assert n >= 0, L.qd_last_error()
    out = np.zeros(n, dtype=np.int64)
    assert L.qd_layout_table(N, R, which, out.ctypes.data_as(_lib.i64p), n) == n
    return out


@pytest.mark.parametrize("N,R", [(2, 3), (6, 6), (12, 8), (20, 14)])
def test_layout_matches_oracle(N, R):
    ch = Channels(N, R)
    assert list(_table(N, R, 6)) == [ch.neq, ch.cc_size, ch.nblock, ch.ncc]
    assert np.array_equal(_table(N, R, 0).reshape(-1, 3), np.stack([ch.nh, ch.np_, ch.nb], 1))
    assert np.array_equal(_table(N, R, 1).reshape(-1, 6), ch.off)
    assert np.array_equal(_table(N, R, 2).reshape(-1, 3), ch.iarr)
    assert np.array_equal(_table(N, R, 3).reshape(-1, 2), ch.cc_map)
    assert np.array_equal(_table(N, R, 4).reshape(-1, 2), np.stack([ch.cc_r, ch.cc_nb], 1))
    pairs = np.concatenate([np.concatenate([b["qnhh"], b["qnph"], b["qnpp"]]) for b in ch.blocks])
    assert np.array_equal(_table(N, R, 5).reshape(-1, 2), pairs)


def test_layout_rejects_open_shell():
    L = _lib.load()
    assert L.qd_layout_table(5, 4, 6, None, 0) == -3
    assert b"shell" in L.qd_last_error()


@pytest.mark.parametrize("N,R", [(6, 6), (20, 14)])
def test_layout_view_matches_oracle(N, R):
    """The package's host layout view (library tables) against the oracle's
    channel builder: dims, offsets, single-particle quantum numbers, the
    reference flop count (C2: 10.02 GF) and the synthetic operator recipe."""
    from qd_imsrg_amd.layout import Layout, random_operator
    from tests.pins.layout import random_flat
    lay, ch = Layout(N, R), Channels(N, R)
    assert lay.neq == ch.neq and lay.nblock == ch.nblock and lay.ncc == ch.ncc and lay.cc_size == ch.cc_size
    assert np.array_equal(lay.off, ch.off) and np.array_equal(lay.np_, ch.np_)
    assert np.array_equal(lay.nh, ch.nh) and np.array_equal(lay.nb, ch.nb)
    assert np.array_equal(lay.ml, ch.ml) and np.array_equal(lay.ms, ch.ms)
    assert lay.reference_gemm_flops() == ch.gemm_flops_2b2b() and lay.n2b() == ch.n2b()
    if (N, R) == (20, 14):
        assert lay.reference_gemm_flops()["total"] == 10019392114
    a = random_operator(lay, np.random.default_rng(5), -1)
    b = random_flat(ch, np.random.default_rng(5), -1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("N,R", [(6, 6), (12, 8)])
def test_cc_gather_table_reproduces_calc_cc(N, R):
    """The host-built calc_cc gather (layout table 8: per CC element in device
    order its code and its reference-order index) applied to a random
    operator on the CPU equals the oracle's calc_cc bit for bit, for both
    herm factors; the reference indices are a permutation."""
    from oracle.oracle import Oracle
    from tests.pins.layout import random_flat
    ch = Channels(N, R)
    o = Oracle(ch)
    t = _table(N, R, 8).reshape(-1, 2)
    code, ref = t[:, 0], t[:, 1]
    assert len(code) == ch.cc_size
    assert np.array_equal(np.sort(ref), np.arange(ch.cc_size))
    v = random_flat(ch, np.random.default_rng(11), 1)
    k = code & 0x3FFFFFFF
    zero = k == 0x3FFFFFFF
    assert zero.any() and (~zero).any()
    for h in (1, -1):
        val = np.where(zero, 0.0, v[np.where(zero, 0, k)])
        val = np.where((code >> 30) & 1, h * val, val)
        val = np.where((code >> 31) & 1, -val, val)
        got = np.zeros(ch.cc_size)
        got[ref] = val
        assert np.array_equal(got, o.calc_cc(v, h))
