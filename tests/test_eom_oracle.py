"""Pins of the oracle's TDA matrix and EOM-IM-SRG mat-vec, no GPU.

* The TDA matrix d_ij f_ab - d_ab f_ji + <aj|V|ib> (src/operators.f90:67-90)
  equals <Phi_i^a| H |Phi_j^b> - E0 d computed by brute force in Fock space.
* matvec_prod (HQ_comm_1b / HQ_comm_2b, src/commutators.f90:1162-1260) equals
  the ph and pphh elements of the textbook IM-SRG(2) commutator [H, Q] of
  the dense tensors (tests/pins/textbook.py, itself pinned by Fock-space
  brute force), with Q the ladder operator of the given (M_L, M_S) transfer.
* For a diagonal one-body H the EOM matrix is diagonal with the ph and pphh
  excitation energies; for a decoupled hermitian H it is symmetric.
"""
import numpy as np
import pytest

from oracle.channels import Channels
from oracle.observables import eom_matrix, n_ph, tda_labels, tda_matrix
from oracle.oracle import Oracle
from tests.pins import textbook
from tests.pins.fock import Fock
from tests.pins.layout import flat_to_dense, random_flat


def ladder_dense(ch, ml, ms, x):
    """Dense (E0, f, G) of the ladder operator wrapped in x (unwrap order)."""
    M, n = ch.M, ch.n
    f = np.zeros((M, M))
    G = np.zeros((M, M, M, M))
    g = 0
    for a in range(n, M):
        for i in range(n):
            if ch.ml[a] - ch.ml[i] == ml and ch.ms[a] - ch.ms[i] == ms:
                f[a, i] = x[g]
                g += 1
    for a in range(n, M):
        for b in range(a + 1, M):
            for i in range(n):
                for j in range(i + 1, n):
                    if ch.ml[a] + ch.ml[b] - ch.ml[i] - ch.ml[j] == ml and \
                            ch.ms[a] + ch.ms[b] - ch.ms[i] - ch.ms[j] == ms:
                        v = x[g]
                        G[a, b, i, j], G[b, a, i, j], G[a, b, j, i], G[b, a, j, i] = v, -v, -v, v
                        g += 1
    assert g == len(x)
    return 0.0, f, G


def unwrap_out(ch, ml, ms, C1, C2):
    out = []
    n, M = ch.n, ch.M
    for a in range(n, M):
        for i in range(n):
            if ch.ml[a] - ch.ml[i] == ml and ch.ms[a] - ch.ms[i] == ms:
                out.append(C1[a, i])
    for a in range(n, M):
        for b in range(a + 1, M):
            for i in range(n):
                for j in range(i + 1, n):
                    if ch.ml[a] + ch.ml[b] - ch.ml[i] - ch.ml[j] == ml and \
                            ch.ms[a] + ch.ms[b] - ch.ms[i] - ch.ms[j] == ms:
                        out.append(C2[a, b, i, j])
    return np.array(out)


@pytest.mark.parametrize("N,R,ml,ms", [(2, 3, 0, 0), (2, 3, 1, 2), (6, 4, 0, 0), (6, 4, -1, 0), (6, 4, 2, -2)])
def test_eom_matvec_equals_textbook_commutator(N, R, ml, ms):
    ch = Channels(N, R)
    o = Oracle(ch)
    rng = np.random.default_rng(100 * N + R + 7 * ml + ms)
    h = random_flat(ch, rng, 1)
    nx = o.eom_dim(ml, ms)
    assert nx > 0
    x = rng.standard_normal(nx)
    y = o.eom_matvec(h, 1, ml, ms, x)
    H = flat_to_dense(ch, h, 1)
    Q = ladder_dense(ch, ml, ms, x)
    nocc = np.array([1.0] * ch.n + [0.0] * ch.mp)
    _, C1, C2 = textbook.commutator(H, Q, nocc)
    ref = unwrap_out(ch, ml, ms, C1, C2)
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


def test_eom_diagonal_one_body_closed_form():
    ch = Channels(6, 4)
    o = Oracle(ch)
    e = 1.0 + 0.5 * np.arange(ch.M)
    h = np.zeros(ch.neq)
    n, mp = ch.n, ch.mp
    h[ch.off_fhh:ch.off_fhh + n * n] = np.diag(e[:n]).T.reshape(-1)
    h[ch.off_fpp:ch.off_fpp + mp * mp] = np.diag(e[n:]).T.reshape(-1)
    A = eom_matrix(o, h, 1, 0, 0)
    expect = [e[a] - e[i] for a in range(n, ch.M) for i in range(n)
              if ch.ml[a] == ch.ml[i] and ch.ms[a] == ch.ms[i]]
    expect += [e[a] + e[b] - e[i] - e[j] for a in range(n, ch.M) for b in range(a + 1, ch.M)
               for i in range(n) for j in range(i + 1, n)
               if ch.ml[a] + ch.ml[b] == ch.ml[i] + ch.ml[j] and ch.ms[a] + ch.ms[b] == ch.ms[i] + ch.ms[j]]
    assert np.array_equal(A, np.diag(np.diag(A))) and np.allclose(np.diag(A), expect, rtol=0, atol=1e-14)


def test_eom_matrix_symmetric_for_decoupled_h():
    """f_ph = 0 and Gamma_pphh = 0: [H, Q] on the excitation space is the
    projection of H - E0, a symmetric matrix."""
    ch = Channels(6, 4)
    o = Oracle(ch)
    h = random_flat(ch, np.random.default_rng(5), 1)
    h[ch.off_fph:ch.off_fph + ch.mp * ch.n] = 0.0
    for q in range(ch.nblock):
        k = int(ch.off[q, 2])
        h[k:k + int(ch.np_[q]) * int(ch.nh[q])] = 0.0
    A = eom_matrix(o, h, 1, 1, 0)
    assert np.abs(A - A.T).max() <= 1e-13 * np.abs(A).max()
    assert n_ph(ch, 1, 0) > 0


def excited(fk, i, a):
    """Fock-space vector of a_a^+ a_i |Phi>."""
    st = np.array([fk.phi])
    amp = np.array([1.0])
    st, amp = fk._apply(i, False, st, amp)
    st, amp = fk._apply(a, True, st, amp)
    v = np.zeros(fk.dim)
    v[st] = amp
    return v


@pytest.mark.parametrize("ml,ms", [(1, 0), (-1, 2), (1, -2)])
def test_tda_matrix_matches_fock_space(ml, ms):
    ch = Channels(2, 2)
    o = Oracle(ch)
    h = random_flat(ch, np.random.default_rng(11), 1)
    fk = Fock(ch.M, ch.n)
    Hf = fk.build(flat_to_dense(ch, h, 1))
    lab = tda_labels(ch, ml, ms, ch.M)
    A = tda_matrix(o, h, 1, ml, ms, ch.M)
    assert len(lab) > 0
    vecs = [excited(fk, i, a) for (i, a) in lab]
    B = np.array([[u @ Hf @ w for w in vecs] for u in vecs]) - h[0] * np.eye(len(lab))
    assert np.abs(A - B).max() <= 1e-12 * np.abs(B).max()
    # the cut keeps only particles below it (1-based a <= cutshell)
    cut = ch.n + 2
    assert all(a < cut for _, a in tda_labels(ch, ml, ms, cut))


def test_sector_matrix_matches_fock_space():
    """The K-particle sector builder (tests/pins/sector.py) against the full
    Fock-space matrix of fock.py (M = 6, every K)."""
    from tests.pins.sector import sector_matrix
    ch = Channels(2, 2)
    op = flat_to_dense(ch, random_flat(ch, np.random.default_rng(21), 1), 1)
    Hf = Fock(ch.M, ch.n).build(op)
    for K in range(1, ch.M):
        dets, H = sector_matrix(op, ch.n, K)
        assert np.abs(H - Hf[np.ix_(dets, dets)]).max() <= 1e-12 * np.abs(Hf).max()


def decoupled_dense(ch, seed):
    h = random_flat(ch, np.random.default_rng(seed), 1)
    h[ch.off_fph:ch.off_fph + ch.mp * ch.n] = 0.0
    for q in range(ch.nblock):
        k = int(ch.off[q, 2])
        h[k:k + int(ch.np_[q]) * int(ch.nh[q])] = 0.0
    return h


@pytest.mark.parametrize("N,mode,ml,ms", [(2, 1, 0, 1), (2, 1, 1, -1), (2, 1, -2, 1), (6, 2, 0, 1), (6, 2, 1, -1),
                                          (6, 2, -1, 1)])
def test_eom_attached_removed_spectrum_matches_sector(N, mode, ml, ms):
    """For a decoupled H (f_ph = 0, Gamma_pphh = 0) H|Phi> = E0|Phi>, so the
    EOM matrix of matvec_1p_attached / matvec_1h_removed is H - E0 on the
    span of the determinants a+|Phi>, a+ b+ i|Phi> (attached) or i|Phi>,
    a+ j i|Phi> (removed): its eigenvalues equal those of that block of the
    N +- 1 particle sector Hamiltonian (basis signs drop out)."""
    from tests.pins.sector import sector_matrix
    ch = Channels(N, 3)
    o = Oracle(ch)
    h = decoupled_dense(ch, 30 + mode)
    nx = o.eom13_dim(mode, ml, ms)
    assert nx > 1
    A = np.zeros((nx, nx))
    for k in range(nx):
        e = np.zeros(nx)
        e[k] = 1.0
        A[:, k] = o.eom13_matvec(h, 1, mode, ml, ms, e)
    n, M = ch.n, ch.M
    phi = (1 << n) - 1
    if mode == 1:
        basis = [phi | (1 << a) for a in range(n, M) if ch.ml[a] == ml and ch.ms[a] == ms]
        basis += [(phi & ~(1 << i)) | (1 << a) | (1 << b) for a in range(n, M) for b in range(a + 1, M)
                  for i in range(n) if ch.ml[a] + ch.ml[b] - ch.ml[i] == ml and ch.ms[a] + ch.ms[b] - ch.ms[i] == ms]
        K = n + 1
    else:
        basis = [phi & ~(1 << i) for i in range(n) if -ch.ml[i] == ml and -ch.ms[i] == ms]
        basis += [(phi & ~(1 << a) & ~(1 << b)) | (1 << i) for a in range(n) for b in range(a + 1, n)
                  for i in range(n, M) if -ch.ml[a] - ch.ml[b] + ch.ml[i] == ml and -ch.ms[a] - ch.ms[b] + ch.ms[i] == ms]
        K = n - 1
    assert len(basis) == nx
    dets, H = sector_matrix(flat_to_dense(ch, h, 1), n, K)
    idx = {d: i for i, d in enumerate(dets)}
    sel = [idx[b] for b in basis]
    P = H[np.ix_(sel, sel)] - h[0] * np.eye(nx)
    ev_a = np.sort(np.linalg.eigvals(A).real)
    ev_p = np.sort(np.linalg.eigvalsh(P))
    assert np.abs(ev_a - ev_p).max() <= 1e-11 * max(1.0, np.abs(ev_p).max())
