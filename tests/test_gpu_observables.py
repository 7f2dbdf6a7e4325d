"""The consumers of H(s) on the GPU against the oracle: the TDA spectrum
(src/operators.f90:7-124), the EOM-IM-SRG mat-vec (HQ_comm_1b/2b,
src/commutators.f90:1162-1260) and the Lanczos excited states
(src/EOM_IMSRG.f90:132-280) at N=6 R=6.

Tolerances: TDA eigenvalues and the mat-vec 1e-12 relative (FP64 sums in a
different order); Lanczos eigenvalues of a decoupled H 1e-9 Ha against a dense
eigensolver of the oracle's EOM matrix, their ph weights 1e-6."""
import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.channels import Channels
from oracle.observables import eom_matrix, eom_states, tda_spectrum
from oracle.oracle import Oracle
from tests.pins.layout import random_flat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1():
    print("[case] observables N=6 R=6 hw=1", flush=True)
    ch = Channels(6, 6)
    h0, _, _ = qd.hf_hamiltonian(6, 6, 1.0)
    ctx = qd.Context(6, 6)
    H0, HS, OM = ctx.operator(1, h0), ctx.operator(1), ctx.operator(-1)
    r, _ = qd.magnus_flow(ctx, H0, HS, OM)
    assert r.converged
    yield ch, Oracle(ch), ctx, H0, HS
    ctx.close()


@pytest.mark.parametrize("ml,ms,cut", [(0, 0, 20), (1, 0, 42), (-2, 2, 30), (3, 0, 12)])
def test_tda_spectrum_matches_oracle(c1, ml, ms, cut):
    ch, o, ctx, H0, HS = c1
    for H in (H0, HS):
        got = qd.tda_spectrum(H, ml, ms, cut)
        ref = tda_spectrum(o, H.download(), 1, ml, ms, cut)
        assert got.size == ref.size
        if ref.size:
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("ml,ms", [(0, 0), (1, 0), (-2, 2), (0, -2)])
def test_eom_matvec_matches_oracle(c1, ml, ms):
    ch, o, ctx, H0, HS = c1
    h = random_flat(ch, np.random.default_rng(77), 1)
    H = ctx.operator(1, h)
    n1, n2 = qd.eom_dim(ctx, ml, ms)
    assert n1 + n2 == o.eom_dim(ml, ms) > 0
    x = np.random.default_rng(78).standard_normal(n1 + n2)
    got = qd.eom_matvec(H, ml, ms, x)
    ref = o.eom_matvec(h, 1, ml, ms, x)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def decoupled(ch, v):
    """H(s) with its off-diagonal part (f_ph, every Vpphh) set to zero."""
    w = v.copy()
    w[ch.off_fph:ch.off_fph + ch.mp * ch.n] = 0.0
    for q in range(ch.nblock):
        k = int(ch.off[q, 2])
        w[k:k + int(ch.np_[q]) * int(ch.nh[q])] = 0.0
    return w


@pytest.mark.parametrize("ml,ms", [(0, 0), (1, 0)])
def test_eom_excited_states(c1, ml, ms):
    ch, o, ctx, H0, HS = c1
    hs = HS.download()
    # exactly decoupled: the EOM matrix is symmetric, Lanczos == dense eigensolver
    hd = decoupled(ch, hs)
    D = ctx.operator(1, hd)
    eig, w, it = qd.calculate_excited_states(D, ml, ms, 10)
    ref, wref, A = eom_states(o, hd, 1, ml, ms, 10)
    print(f"(ML, MS) = ({ml}, {ms}): {it} Lanczos steps, dE = {np.round(eig[:4], 6)}")
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    assert np.abs(eig - ref).max() <= 1e-9 and np.all(np.diff(eig) >= 0)
    assert np.abs(w - wref).max() <= 1e-6
    # the flowed H(s) itself (off-diagonal part ~1e-5 left): close to the
    # eigenvalues of the symmetric part of its EOM matrix
    eig2, _, _ = qd.calculate_excited_states(HS, ml, ms, 10)
    Ah = eom_matrix(o, hs, 1, ml, ms)
    ref2 = np.linalg.eigvalsh(0.5 * (Ah + Ah.T))[:10]
    assert np.abs(eig2 - ref2).max() <= 1e-6


@pytest.mark.parametrize("mode,ml,ms", [(1, 0, 1), (1, 2, -1), (2, 0, 1), (2, -1, -1)])
def test_eom_attached_removed(c1, mode, ml, ms):
    """matvec_1p_attached / matvec_1h_removed against the oracle (1e-12) and
    the Lanczos eigenvalues of a decoupled H against a dense eigensolver of
    the oracle's matrix (1e-9 Ha), one-index weights 1e-6."""
    ch, o, ctx, H0, HS = c1
    h = random_flat(ch, np.random.default_rng(90 + mode), 1)
    H = ctx.operator(1, h)
    n1, n3 = qd.eom13_dim(ctx, mode, ml, ms)
    assert n1 + n3 == o.eom13_dim(mode, ml, ms) > 10
    x = np.random.default_rng(91).standard_normal(n1 + n3)
    got = qd.eom13_matvec(H, mode, ml, ms, x)
    ref = o.eom13_matvec(h, 1, mode, ml, ms, x)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    hd = decoupled(ch, HS.download())
    D = ctx.operator(1, hd)
    f = qd.calculate_1p_attached if mode == 1 else qd.calculate_1h_removed
    nev = 5
    eig, w, it = f(D, ml, ms, nev)
    nx = n1 + n3
    A = np.zeros((nx, nx))
    for k in range(nx):
        e = np.zeros(nx)
        e[k] = 1.0
        A[:, k] = o.eom13_matvec(hd, 1, mode, ml, ms, e)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    lam, U = np.linalg.eigh(A)
    print(f"mode {mode} (ML, MS) = ({ml}, {ms}): {it} Lanczos steps, dE = {np.round(eig, 6)}")
    assert np.abs(eig - lam[:nev]).max() <= 1e-9
    assert np.abs(w - (U[:n1, :nev] ** 2).sum(axis=0)).max() <= 1e-6
