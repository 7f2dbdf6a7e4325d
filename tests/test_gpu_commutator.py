"""GPU parity: every commutator entry point of the HIP library against the CPU
oracle on the same seeded inputs.  Tolerance: max |gpu - oracle| <=
1e-12 * max |oracle| (FP64; the MFMA accumulation order differs from the
oracle's).  Sizes: N=6/R=6 (the CPU-runnable configuration), N=12/R=10 (other
channel sizes and box mixes) and N=20/R=14 (the commutator microbenchmark's
channel structure, where the largest class-pair boxes are split into slabs)."""
import sys

import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.channels import Channels
from oracle.oracle import Oracle
from tests.pins.layout import random_flat

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


class Case:
    def __init__(self, N, R):
        print(f"[case] N={N} R={R}", flush=True)
        self.ch = Channels(N, R)
        self.o = Oracle(self.ch)
        self.ctx = qd.Context(N, R)
        assert self.ctx.neq == self.ch.neq and self.ctx.cc_size == self.ch.cc_size

    def ops(self, seed, h1, h2, dense_f=False):
        rng = np.random.default_rng(seed)
        x = random_flat(self.ch, rng, h1, dense_f=dense_f)
        y = random_flat(self.ch, rng, h2, dense_f=dense_f)
        return x, y, self.ctx.operator(h1, x), self.ctx.operator(h2, y)


@pytest.fixture(scope="module", params=[(6, 6), (12, 10), (20, 14)], ids=["N6R6", "N12R10", "N20R14"])
def case(request):
    c = Case(*request.param)
    yield c
    c.ctx.close()


@pytest.mark.parametrize("h1,h2", [(-1, 1), (1, 1), (1, -1)])
def test_zero_body(case, h1, h2):
    x, y, X, Y = case.ops(1, h1, h2)
    a, b = qd.commutator_110(X, Y), qd.commutator_220(X, Y)
    ra, rb = case.o.commutator_110(x, h1, y, h2), case.o.commutator_220(x, h1, y, h2)
    assert abs(a - ra) <= TOL * max(abs(ra), 1) and abs(b - rb) <= TOL * max(abs(rb), 1)


@pytest.mark.parametrize("h1,h2", [(-1, 1), (1, 1)])
def test_one_body_pieces(case, h1, h2):
    x, y, X, Y = case.ops(2, h1, h2)
    o, ctx = case.o, case.ctx
    # 111 overwrites, 121 and 221 add (BCH_EXPAND call order)
    Z = ctx.operator(1)
    ref = o.zeros()
    qd.commutator_111(X, Y, Z); o.commutator_111(x, h1, y, h2, ref)
    one = slice(1, case.ch.off[0, 0])
    assert rel(Z.download()[one], ref[one]) <= TOL
    qd.commutator_121(X, Y, Z); o.commutator_121(x, h1, y, h2, ref)
    assert rel(Z.download()[one], ref[one]) <= TOL
    qd.commutator_221(X, Y, Z); o.commutator_221(x, h1, y, h2, ref)
    assert rel(Z.download()[one], ref[one]) <= TOL


@pytest.mark.parametrize("h1,h2", [(-1, 1), (1, -1)])
def test_two_body_pieces(case, h1, h2):
    x, y, X, Y = case.ops(3, h1, h2)
    o, ctx = case.o, case.ctx
    two = slice(case.ch.off[0, 0], None)
    Z = ctx.operator(1)
    ref = o.zeros()
    qd.commutator_122(X, Y, Z); o.commutator_122(x, h1, y, h2, ref)
    assert rel(Z.download()[two], ref[two]) <= TOL
    # 222 with caller-supplied CC sets (oracle calc_cc uploaded)
    hcc, ecc = o.calc_cc(y, h2), o.calc_cc(x, h1)
    HCC, ECC = ctx.cc(float(h2)), ctx.cc(float(h1))
    HCC.upload(hcc[:case.ch.cc_size]); ECC.upload(ecc[:case.ch.cc_size])
    qd.commutator_222(X, Y, Z, HCC, ECC); o.commutator_222(x, h1, y, h2, ref, hcc, h2, ecc, h1)
    assert rel(Z.download()[two], ref[two]) <= TOL


def test_calc_cc(case):
    x, _, X, _ = case.ops(4, -1, 1)
    C = case.ctx.cc(-1.0)
    qd.calc_cc(X, C)
    assert np.array_equal(C.download(), case.o.calc_cc(x, -1)[:case.ch.cc_size])


def test_cc_transfer_round_trip(case):
    # the device keeps CC rows in its own order; upload and download permute
    # them, so a round trip returns the reference layout bit for bit
    v = np.random.default_rng(11).standard_normal(case.ch.cc_size)
    C = case.ctx.cc(1.0)
    C.upload(v)
    assert np.array_equal(C.download(), v)


@pytest.mark.parametrize("pandya", [True, False])
@pytest.mark.parametrize("h1,h2", [(-1, 1), (1, 1)])
def test_full_commutator(case, h1, h2, pandya):
    x, y, X, Y = case.ops(5, h1, h2)
    ctx = case.ctx
    Z = ctx.operator(-h1 * h2)
    if pandya:
        HCC, ECC = ctx.cc(float(h2)), ctx.cc(float(h1))
        qd.calc_cc(Y, HCC); qd.calc_cc(X, ECC)
        e0 = qd.commutator(X, Y, Z, HCC, ECC)
    else:
        e0 = qd.commutator(X, Y, Z)
    ref = case.o.commutator(x, h1, y, h2, pandya=pandya)
    got = Z.download()
    assert rel(got, ref) <= TOL
    assert e0 == got[0]


@pytest.mark.parametrize("h1,h2", [(-1, 1), (1, 1)])
def test_commutator_cc_equals_calc_cc_then_commutator(case, h1, h2):
    # the fused entry overlaps its two calc_cc launches with the ladder
    # products and, its CC sets being calc_cc of the operands, runs the paired
    # box tasks (one grid writes both orientations of a class-pair box, the
    # transposed outputs as h_Z times the direct ones); the separate calls
    # with caller-supplied CC sets run every box on its own.  CC sets and E0
    # are bit-identical, the two-body outputs equal to rounding; the serial
    # schedule (timing events on) gives the overlapped one's bits.
    x, y, X, Y = case.ops(9, h1, h2)
    ctx = case.ctx
    Z1, Z2, Z3 = ctx.operator(-h1 * h2), ctx.operator(-h1 * h2), ctx.operator(-h1 * h2)
    H1, E1 = ctx.cc(float(h2)), ctx.cc(float(h1))
    H2, E2 = ctx.cc(float(h2)), ctx.cc(float(h1))
    qd.calc_cc(Y, H1); qd.calc_cc(X, E1)
    e1 = qd.commutator(X, Y, Z1, H1, E1)
    e2 = qd.commutator_cc(X, Y, Z2, H2, E2)
    assert np.array_equal(H1.download(), H2.download()) and np.array_equal(E1.download(), E2.download())
    z1, z2 = Z1.download(), Z2.download()
    assert rel(z2, z1) <= 1e-14 and e1 == e2
    ctx.timing(True)
    e3 = qd.commutator_cc(X, Y, Z3, H2, E2)
    ctx.timing_read()
    ctx.timing(False)
    assert np.array_equal(Z3.download(), z2) and e3 == e2
    ref = case.o.commutator(x, h1, y, h2, pandya=True)
    assert rel(z2, ref) <= TOL and rel(z1, ref) <= TOL


def test_commutator_cc_graph_replays(case):
    # with graphs on, the call is captured once and replayed; every replay,
    # with eager work between them, equals the eager result bit for bit
    x, y, X, Y = case.ops(10, -1, 1)
    ctx = case.ctx
    Z0, Z = ctx.operator(1), ctx.operator(1)
    H0, E0 = ctx.cc(1.0), ctx.cc(-1.0)
    H, E = ctx.cc(1.0), ctx.cc(-1.0)
    ref = qd.commutator_cc(X, Y, Z0, H0, E0)
    ctx.set_graphs(True)
    try:
        for rep in range(4):
            W = ctx.operator(1)
            qd.commutator(Y, X, W)  # eager work between the replays
            e = qd.commutator_cc(X, Y, Z, H, E)
            assert e == ref and np.array_equal(Z.download(), Z0.download()), rep
            assert np.array_equal(H.download(), H0.download()) and np.array_equal(E.download(), E0.download())
            Z.upload(np.zeros(case.ch.neq))
    finally:
        ctx.set_graphs(False)


def test_dense_one_body(case):
    x, y, X, Y = case.ops(6, -1, 1, dense_f=True)
    Z = case.ctx.operator(1)
    HCC, ECC = case.ctx.cc(1.0), case.ctx.cc(-1.0)
    qd.calc_cc(Y, HCC); qd.calc_cc(X, ECC)
    qd.commutator(X, Y, Z, HCC, ECC)
    assert rel(Z.download(), case.o.commutator(x, -1, y, 1)) <= TOL


def test_deterministic(case):
    _, _, X, Y = case.ops(7, -1, 1)
    HCC, ECC = case.ctx.cc(1.0), case.ctx.cc(-1.0)
    qd.calc_cc(Y, HCC); qd.calc_cc(X, ECC)
    Z1, Z2 = case.ctx.operator(1), case.ctx.operator(1)
    qd.commutator(X, Y, Z1, HCC, ECC)
    qd.commutator(X, Y, Z2, HCC, ECC)
    assert np.array_equal(Z1.download(), Z2.download())


def test_white_mbpt2_norm_add(case):
    x, y, X, Y = case.ops(8, 1, 1)
    o, ctx = case.o, case.ctx
    E = ctx.operator(-1)
    qd.build_white(E, X)
    assert rel(E.download(), o.build_white(x, 1)) <= 1e-15
    assert abs(qd.MBPT2(X) - o.mbpt2(x)) <= 1e-12 * abs(o.mbpt2(x))
    assert abs(qd.mat_2_norm(X) - o.norm(x)) <= 1e-13 * o.norm(x)
    Z = ctx.operator(1)
    qd.add_arrays(X, 0.5, Y, -2.0, Z)
    assert np.array_equal(Z.download(), o.add(x, 0.5, y, -2.0))


def test_bch_expand_pandya(case):
    """BCH_EXPAND with calc_cc-filled CC inputs (every nested commutator gets
    its Pandya term) against the oracle's restatement."""
    from oracle.flow import bch_expand_pandya
    rng = np.random.default_rng(10)
    g = random_flat(case.ch, rng, -1, scale=0.02)
    h = random_flat(case.ch, rng, 1)
    G, H, HS = case.ctx.operator(-1, g), case.ctx.operator(1, h), case.ctx.operator(1)
    HCC, OCC = case.ctx.cc(1.0), case.ctx.cc(-1.0)
    qd.calc_cc(H, HCC)
    qd.calc_cc(G, OCC)
    nt = qd.BCH_EXPAND(HS, G, H, HCC, OCC)
    ref, nref = bch_expand_pandya(case.o, g, h)
    no_pandya, _ = case.o.bch_expand(g, -1, h)
    assert nt == nref
    assert rel(HS.download(), ref) <= 1e-11
    assert rel(no_pandya, ref) > 1e-6  # the Pandya term is really there


def test_bch_expand(case):
    rng = np.random.default_rng(9)
    g = random_flat(case.ch, rng, -1, scale=0.02)
    h = random_flat(case.ch, rng, 1)
    G, H, HS = case.ctx.operator(-1, g), case.ctx.operator(1, h), case.ctx.operator(1)
    nt = qd.BCH_EXPAND(HS, G, H)
    ref, nref = case.o.bch_expand(g, -1, h)
    assert nt == nref
    assert rel(HS.download(), ref) <= 1e-11


def generator_shaped(ch, v):
    """v with only fph and the Vpphh sub-blocks kept (the shape of a White
    generator and of the Magnus Omega built from them)."""
    w = np.zeros_like(v)
    w[ch.off_fph:ch.off_fph + ch.mp * ch.n] = v[ch.off_fph:ch.off_fph + ch.mp * ch.n]
    for q in range(ch.nblock):
        o, nh, np_ = int(ch.off[q][2]), int(ch.nh[q]), int(ch.np_[q])
        w[o:o + np_ * nh] = v[o:o + np_ * nh]
    return w


@pytest.mark.parametrize("n,r", [(6, 6), (20, 14)])
def test_bch_skips_zero_subblocks(monkeypatch, n, r):
    # BCH_EXPAND scans G for sub-block types that are zero in every block and
    # skips the GEMM products that read them (QD_ZSKIP=0 never skips): with a
    # generator-shaped G (fph and Vpphh only) both give the same result and
    # the oracle's, with the same term count; with the Pandya term too
    outs = {}
    for zs in ("1", "0"):
        monkeypatch.setenv("QD_ZSKIP", zs)
        c = Case(n, r)
        try:
            rng = np.random.default_rng(31)
            g = generator_shaped(c.ch, random_flat(c.ch, rng, -1, scale=0.02))
            h = random_flat(c.ch, rng, 1)
            G, H = c.ctx.operator(-1, g), c.ctx.operator(1, h)
            from oracle.flow import bch_expand_pandya
            res = []
            for pandya in (False, True):
                HS = c.ctx.operator(1)
                if pandya:
                    HCC, OCC = c.ctx.cc(1.0), c.ctx.cc(-1.0)
                    qd.calc_cc(H, HCC)
                    qd.calc_cc(G, OCC)
                    nt = qd.BCH_EXPAND(HS, G, H, HCC, OCC)
                    ref, nref = bch_expand_pandya(c.o, g, h)
                else:
                    nt = qd.BCH_EXPAND(HS, G, H)
                    ref, nref = c.o.bch_expand(g, -1, h)
                assert nt == nref
                assert rel(HS.download(), ref) <= 1e-11
                res.append((nt, HS.download()))
            outs[zs] = res
        finally:
            c.ctx.close()
    for (n1, a), (n0, b) in zip(outs["1"], outs["0"]):
        assert n1 == n0 and np.array_equal(a, b)


@pytest.mark.parametrize("cap", [40, 100])
def test_box_slabs(monkeypatch, cap):
    # a small cap on the LDS grid of the two-body boxes splits most boxes of
    # N=6/R=6 into column slabs (bra terms) and row slabs (ket terms)
    monkeypatch.setenv("QD_BOX_CAP", str(cap))
    c = Case(6, 6)
    try:
        for h1, h2 in [(-1, 1), (1, 1)]:
            x, y, X, Y = c.ops(21, h1, h2)
            Z = c.ctx.operator(-h1 * h2)
            HCC, ECC = c.ctx.cc(float(h2)), c.ctx.cc(float(h1))
            qd.calc_cc(Y, HCC); qd.calc_cc(X, ECC)
            qd.commutator(X, Y, Z, HCC, ECC)
            assert rel(Z.download(), c.o.commutator(x, h1, y, h2, pandya=True)) <= TOL
            two = slice(c.ch.off[0, 0], None)
            Z = c.ctx.operator(1)
            ref = c.o.zeros()
            qd.commutator_122(X, Y, Z); c.o.commutator_122(x, h1, y, h2, ref)
            assert rel(Z.download()[two], ref[two]) <= TOL
            # paired boxes (the ones that fit the cap both ways) beside slabs
            Z = c.ctx.operator(-h1 * h2)
            qd.commutator_cc(X, Y, Z, HCC, ECC)
            assert rel(Z.download(), c.o.commutator(x, h1, y, h2, pandya=True)) <= TOL
    finally:
        c.ctx.close()


def test_box_pairs_switch(monkeypatch):
    # QD_BOX_PAIRS=0: every class-pair box on its own; the paired form (the
    # transposed box's outputs as h_Z times the direct ones) agrees to 1e-14
    outs = {}
    for pairs in ("1", "0"):
        monkeypatch.setenv("QD_BOX_PAIRS", pairs)
        c = Case(6, 6)
        try:
            x, y, X, Y = c.ops(23, -1, 1)
            Z, Z2 = c.ctx.operator(-1), c.ctx.operator(1)
            HCC, ECC = c.ctx.cc(1.0), c.ctx.cc(-1.0)
            qd.commutator_cc(X, Y, Z, HCC, ECC)
            qd.commutator_122(X, Y, Z2)
            outs[pairs] = (Z.download(), Z2.download())
            assert rel(outs[pairs][0], c.o.commutator(x, -1, y, 1, pandya=True)) <= TOL
        finally:
            c.ctx.close()
    for k in range(2):
        assert rel(outs["1"][k], outs["0"][k]) <= 1e-14


@pytest.mark.parametrize("n,r", [(6, 6), (20, 14)])
def test_staged_pandya_reads(monkeypatch, n, r):
    # QD_DSTAGE=1: whole boxes copy their Pandya sub-blocks of D (contiguous
    # in run panels) into LDS and sum the four terms from there; QD_DSTAGE=0
    # reads each term from D.  Same terms, same order: bitwise equal, both
    # against the oracle (paired boxes with calc_cc inputs, and every box
    # direct with caller-supplied CC sets)
    outs = {}
    for st in ("1", "0"):
        monkeypatch.setenv("QD_DSTAGE", st)
        c = Case(n, r)
        try:
            x, y, X, Y = c.ops(29, -1, 1)
            ref = c.o.commutator(x, -1, y, 1, pandya=True)
            Z1, Z2 = c.ctx.operator(1), c.ctx.operator(1)
            HCC, ECC = c.ctx.cc(1.0), c.ctx.cc(-1.0)
            qd.commutator_cc(X, Y, Z1, HCC, ECC)
            qd.commutator(X, Y, Z2, HCC, ECC)
            outs[st] = (Z1.download(), Z2.download())
            for z in outs[st]:
                assert rel(z, ref) <= TOL
        finally:
            c.ctx.close()
    for k in range(2):
        assert np.array_equal(outs["1"][k], outs["0"][k])


@pytest.mark.parametrize("wide", ["0", "1"])
def test_direct_square_ladder_terms(monkeypatch, wide):
    # QD_LDIRECT=1 (the default): the box pass forms the ladder terms of the
    # HHHH, PHPH and PPPP sub-blocks from the products at (I, J) and (J, I);
    # QD_LDIRECT=0: k_ladder_sq forms them into a buffer first.  Both against
    # the oracle for every herm pair, with paired boxes (commutator_cc; the
    # narrow form writes the paired outputs in a second phase) and every box
    # on its own (commutator), and against each other to rounding
    monkeypatch.setenv("QD_BOX_WIDE", wide)
    outs = {}
    for ld in ("1", "0"):
        monkeypatch.setenv("QD_LDIRECT", ld)
        c = Case(6, 6)
        try:
            res = []
            for h1, h2 in [(-1, 1), (1, 1), (1, -1), (-1, -1)]:
                x, y, X, Y = c.ops(37, h1, h2)
                ref = c.o.commutator(x, h1, y, h2, pandya=True)
                HCC, ECC = c.ctx.cc(float(h2)), c.ctx.cc(float(h1))
                for paired in (True, False):
                    Z = c.ctx.operator(-h1 * h2)
                    if paired:
                        qd.commutator_cc(X, Y, Z, HCC, ECC)
                    else:
                        qd.commutator(X, Y, Z, HCC, ECC)
                    assert rel(Z.download(), ref) <= TOL, (h1, h2, paired)
                    res.append(Z.download())
            outs[ld] = res
        finally:
            c.ctx.close()
    for a, b in zip(outs["1"], outs["0"]):
        assert rel(a, b) <= 1e-14


def test_without_gather_tables(monkeypatch):
    # QD_GATHER_TABLES=0: calc_cc and the 221 traces resolve the pair tables
    # per element, the path of operators whose flat index needs more than
    # 30 bits; calc_cc bit for bit and the full commutator against the oracle
    monkeypatch.setenv("QD_GATHER_TABLES", "0")
    c = Case(6, 6)
    try:
        for h1, h2 in [(-1, 1), (1, 1)]:
            x, y, X, Y = c.ops(23, h1, h2)
            Z = c.ctx.operator(-h1 * h2)
            HCC, ECC = c.ctx.cc(float(h2)), c.ctx.cc(float(h1))
            qd.calc_cc(Y, HCC); qd.calc_cc(X, ECC)
            assert np.array_equal(HCC.download(), c.o.calc_cc(y, h2)[:c.ch.cc_size])
            qd.commutator(X, Y, Z, HCC, ECC)
            assert rel(Z.download(), c.o.commutator(x, h1, y, h2, pandya=True)) <= TOL
    finally:
        c.ctx.close()


@pytest.mark.parametrize("cap", ["8000", "100"])
def test_wide_box_form(monkeypatch, cap):
    # QD_BOX_WIDE=1: 1024-thread tasks with grids of up to 8000 elements (the
    # form chosen by itself at R = 20); with a cap of 100 most boxes take the
    # slab path of the wide form
    monkeypatch.setenv("QD_BOX_WIDE", "1")
    monkeypatch.setenv("QD_BOX_CAP", cap)
    c = Case(6, 6)
    try:
        for h1, h2 in [(-1, 1), (1, 1)]:
            x, y, X, Y = c.ops(25, h1, h2)
            Z = c.ctx.operator(-h1 * h2)
            HCC, ECC = c.ctx.cc(float(h2)), c.ctx.cc(float(h1))
            qd.commutator_cc(X, Y, Z, HCC, ECC)
            assert rel(Z.download(), c.o.commutator(x, h1, y, h2, pandya=True)) <= TOL
            qd.commutator(X, Y, Z, HCC, ECC)
            assert rel(Z.download(), c.o.commutator(x, h1, y, h2, pandya=True)) <= TOL
            two = slice(c.ch.off[0, 0], None)
            Z = c.ctx.operator(1)
            ref = c.o.zeros()
            qd.commutator_122(X, Y, Z); c.o.commutator_122(x, h1, y, h2, ref)
            assert rel(Z.download()[two], ref[two]) <= TOL
    finally:
        c.ctx.close()
