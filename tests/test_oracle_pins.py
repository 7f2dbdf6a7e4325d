"""Pins for the CPU oracle (no GPU).

Chain: Fock-space brute force  ->  textbook dense IM-SRG(2) commutator  ->
oracle (the reference's blocked algorithm).  Each link is checked on inputs
where a dropped term, a wrong sign, a swapped index or a transposed operand
shows up as an O(1) difference.
"""
import numpy as np
import pytest

from oracle.channels import Channels
from oracle.oracle import Oracle
from tests.pins.fock import Fock
from tests.pins.layout import dense_to_flat, flat_to_dense, random_flat
from tests.pins.textbook import commutator, random_operator


@pytest.mark.parametrize("M,n,seed", [(6, 2, 1), (7, 3, 2), (8, 3, 3)])
def test_textbook_matches_fock_space(M, n, seed):
    """General (non-hermitian) operators: textbook == exact Fock commutator
    projected to its normal-ordered 0/1/2-body parts."""
    rng = np.random.default_rng(seed)
    A, B = random_operator(M, rng), random_operator(M, rng)
    F = Fock(M, n)
    XA, XB = F.build(A), F.build(B)
    E0, f, G = F.extract(XA @ XB - XB @ XA)
    nocc = np.array([1] * n + [0] * (M - n))
    c0, c1, c2 = commutator(A, B, nocc)
    assert abs(E0 - c0) < 1e-12 * max(1, abs(c0))
    assert np.abs(f - c1).max() < 1e-12 * np.abs(c1).max()
    assert np.abs(G - c2).max() < 1e-12 * np.abs(c2).max()
    # the Pandya term is not negligible on these inputs
    _, _, c2np = commutator(A, B, nocc, pandya=False)
    assert np.abs(G - c2np).max() > 1e-2 * np.abs(c2).max()


def test_fock_roundtrip():
    rng = np.random.default_rng(7)
    A = random_operator(6, rng)
    F = Fock(6, 2)
    e, f, G = F.extract(F.build(A))
    assert abs(e - A[0]) < 1e-13 and np.abs(f - A[1]).max() < 1e-13 and np.abs(G - A[2]).max() < 1e-13


@pytest.fixture(scope="module")
def c1():
    ch = Channels(6, 6)
    return ch, Oracle(ch)


@pytest.mark.parametrize("h1,h2", [(-1, 1), (1, 1), (-1, -1), (1, -1)])
@pytest.mark.parametrize("pandya", [True, False])
def test_oracle_matches_textbook_c1(c1, h1, h2, pandya):
    ch, o = c1
    rng = np.random.default_rng(100 + 10 * h1 + h2)
    v1, v2 = random_flat(ch, rng, h1), random_flat(ch, rng, h2)
    r = o.commutator(v1, h1, v2, h2, pandya=pandya)
    nocc = np.array([1] * ch.n + [0] * ch.mp)
    ref = dense_to_flat(ch, *commutator(flat_to_dense(ch, v1, h1), flat_to_dense(ch, v2, h2), nocc, pandya))
    assert np.abs(r - ref).max() <= 1e-13 * np.abs(ref).max()
    # the result carries the implied hermiticity -h1*h2 exactly like the textbook tensor
    _, _, G = flat_to_dense(ch, r, -h1 * h2)
    _, _, Gt = commutator(flat_to_dense(ch, v1, h1), flat_to_dense(ch, v2, h2), nocc, pandya)
    assert np.abs(G - Gt).max() <= 1e-12 * np.abs(Gt).max()


def test_oracle_dense_one_body_c1(c1):
    """Random one-body entries off the (m_l, m_s) diagonal: stored blocks still
    match the textbook (the reference's formulas need no symmetry of fph)."""
    ch, o = c1
    rng = np.random.default_rng(11)
    v1, v2 = random_flat(ch, rng, -1, dense_f=True), random_flat(ch, rng, 1, dense_f=True)
    r = o.commutator(v1, -1, v2, 1)
    nocc = np.array([1] * ch.n + [0] * ch.mp)
    ref = dense_to_flat(ch, *commutator(flat_to_dense(ch, v1, -1), flat_to_dense(ch, v2, 1), nocc))
    assert np.abs(r - ref).max() <= 1e-13 * np.abs(ref).max()


def test_oracle_pieces_sum_to_full(c1):
    """110/220/111/121/221/122/222 called one by one in BCH order == full."""
    ch, o = c1
    rng = np.random.default_rng(3)
    v1, v2 = random_flat(ch, rng, -1), random_flat(ch, rng, 1)
    hcc, ecc = o.calc_cc(v2, 1), o.calc_cc(v1, -1)
    r = o.zeros()
    r[0] = o.commutator_110(v1, -1, v2, 1) + o.commutator_220(v1, -1, v2, 1)
    o.commutator_111(v1, -1, v2, 1, r)
    o.commutator_121(v1, -1, v2, 1, r)
    o.commutator_221(v1, -1, v2, 1, r)
    o.commutator_122(v1, -1, v2, 1, r)
    o.commutator_222(v1, -1, v2, 1, r, hcc, 1.0, ecc, -1.0)
    full = o.commutator(v1, -1, v2, 1)
    assert np.array_equal(r, full)


def test_oracle_algebraic_identities(c1):
    ch, o = c1
    rng = np.random.default_rng(4)
    x, y, z = random_flat(ch, rng, -1), random_flat(ch, rng, 1), random_flat(ch, rng, 1)
    zero = o.zeros()
    assert np.abs(o.commutator(x, -1, zero, 1)).max() == 0.0
    # antisymmetry: [x, y] = -[y, x]
    a, b = o.commutator(x, -1, y, 1), o.commutator(y, 1, x, -1)
    assert np.abs(a + b).max() <= 1e-13 * np.abs(a).max()
    # bilinearity in the second slot
    lin = o.commutator(x, -1, 2.0 * y - 0.5 * z, 1)
    ref = 2.0 * a - 0.5 * o.commutator(x, -1, z, 1)
    assert np.abs(lin - ref).max() <= 1e-12 * np.abs(ref).max()
    # [x, N] = 0 for the number operator N = n + sum_p {p+ p}
    num = o.zeros()
    num[0] = ch.n
    num[ch.off_fhh:ch.off_fhh + ch.n ** 2] = np.eye(ch.n).reshape(-1)
    num[ch.off_fpp:ch.off_fpp + ch.mp ** 2] = np.eye(ch.mp).reshape(-1)
    assert np.abs(o.commutator(x, -1, num, 1)).max() <= 1e-13 * np.abs(x).max()


def test_velem_antisymmetry_and_blocks(c1):
    ch, o = c1
    rng = np.random.default_rng(9)
    v = random_flat(ch, rng, 1)
    for _ in range(200):
        a, b, c, d = rng.integers(0, ch.M, 4)
        x = o.velem(v, 1, a, b, c, d)
        assert o.velem(v, 1, b, a, c, d) == -x
        assert o.velem(v, 1, a, b, d, c) == -x
        assert o.velem(v, 1, c, d, a, b) == x  # hermitian
        if ch.ml[a] + ch.ml[b] != ch.ml[c] + ch.ml[d]:
            assert x == 0.0


def test_white_generator_closed_form(c1):
    """eta_ph = f_ph / (f_pp - f_hh - <ph|ph>) and
    eta_pphh = V / Epstein-Nesbet denominator, checked on the dense tensors."""
    ch, o = c1
    rng = np.random.default_rng(12)
    h = random_flat(ch, rng, 1)
    eta = o.build_white(h, 1)
    _, f, G = flat_to_dense(ch, h, 1)
    _, fe, Ge = flat_to_dense(ch, eta, -1)
    n, M = ch.n, ch.M
    for a in range(n, M):
        for i in range(n):
            assert np.isclose(fe[a, i], f[a, i] / (f[a, a] - f[i, i] - G[a, i, a, i]), rtol=1e-14, atol=0)
    for q, b in enumerate(ch.blocks):
        for (p1, p2) in b["qnpp"]:
            for (h1, h2) in b["qnhh"]:
                den = (f[p1, p1] + f[p2, p2] - f[h1, h1] - f[h2, h2] + G[p1, p2, p1, p2] + G[h1, h2, h1, h2]
                       - G[p1, h1, p1, h1] - G[p1, h2, p1, h2] - G[p2, h1, p2, h1] - G[p2, h2, p2, h2])
                assert np.isclose(Ge[p1, p2, h1, h2], G[p1, p2, h1, h2] / den, rtol=1e-13, atol=0)


def test_mbpt2_closed_form(c1):
    ch, o = c1
    rng = np.random.default_rng(13)
    h = random_flat(ch, rng, 1)
    _, f, G = flat_to_dense(ch, h, 1)
    n = ch.n
    e2 = 0.0
    for a in range(n, ch.M):
        for b in range(a + 1, ch.M):
            for i in range(n):
                for j in range(i + 1, n):
                    if G[a, b, i, j] != 0.0:
                        e2 += G[a, b, i, j] ** 2 / (f[i, i] + f[j, j] - f[a, a] - f[b, b])
    assert np.isclose(o.mbpt2(h), e2, rtol=1e-13)


@pytest.mark.parametrize("N,R,neq,nblk,ncc", [(6, 6, 18482, 63, 105), (20, 14, 5083094, 159, 265),
                                               (20, 16, 12970043, 183, 305)])
def test_channel_counts(N, R, neq, nblk, ncc):
    ch = Channels(N, R)
    assert (ch.neq, ch.nblock, ch.ncc) == (neq, nblk, ncc)
    assert ch.nblock == 12 * R - 9 and ch.ncc == 20 * (R - 1) + 5


def test_diis_weights_minimise_residual():
    """Pulay weights: sum to 1 and minimise |sum c_i r_i| (closed form
    c = B^-1 1 / (1^T B^-1 1) for a non-singular Gram matrix B)."""
    from oracle.flow import diis_weights
    rng = np.random.default_rng(12)
    R = rng.standard_normal((4, 30))
    B = R @ R.T
    c = diis_weights(B)
    x = np.linalg.solve(B, np.ones(4))
    assert abs(c.sum() - 1) < 1e-13 and np.allclose(c, x / x.sum(), rtol=1e-10, atol=1e-12)
    best = np.linalg.norm(c @ R)
    for _ in range(20):  # any other affine combination has a larger residual
        d = rng.standard_normal(4)
        d -= d.mean()
        assert np.linalg.norm((c + 0.1 * d) @ R) > best


def test_diis_flow_reaches_the_same_energy():
    """The Pulay-accelerated flow converges to the plain flow's E_gs (the
    decoupled fixed point, 1e-6 Ha) in fewer iterations (N=6 R=6 hw=1)."""
    from oracle.flow import magnus_flow
    from tests.pins.hamiltonians import model_hamiltonian
    ch = Channels(6, 6)
    o = Oracle(ch)
    h0 = model_hamiltonian(ch, 1)
    a = magnus_flow(o, h0)
    b = magnus_flow(o, h0, diis=6)
    assert a["crit"] <= 1e-9 and b["crit"] <= 1e-9
    assert abs(a["e0"] - b["e0"]) <= 1e-6
    assert b["iterations"] < a["iterations"]
