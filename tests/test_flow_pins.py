"""Pins of the oracle's BCH_EXPAND and s_evolve (src/magnus_IMSRG.f90:194-318)
by what the mathematics fixes, independent of the oracle's code path.

BCH_EXPAND: for one-body Omega and H the nested commutators are matrix
commutators of the M x M one-body matrices and the zero-body part of [A, B]
is the hole trace of the matrix commutator, so the series, the literal
coefficients cof(k) (src/magnus_IMSRG.f90:265-267), the 1e-5 truncation rule
(:316) and the E0 accumulation are recomputed here with dense matrices; a
two-level rotation has the closed form cos/sin 2 theta, and the full series
is exp(Omega) F exp(-Omega).  A wrong cof index, a sign, a dropped term or a
term-count slip fails one of these.

s_evolve: the first step is Omega = stp * eta(H(0)) and H(s) = BCH(Omega,
H(0)); a forced oversized step reproduces the reject-and-halve sequence with
the widened constants (:220-228) and an ||eta_0|| that is never updated; a
one-body Hamiltonian flows to the exact decoupled energy (the sum of the N
lowest eigenvalues of f)."""
import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle.channels import Channels
from oracle.flow import BCH_COF, R4_0P7, R4_1EM3, R4_1EM5, magnus_flow
from oracle.oracle import Oracle
from tests.pins.hamiltonians import model_hamiltonian
from tests.pins.layout import dense_to_flat, flat_to_dense, random_flat

EXACT = [1.0 / math.factorial(k) for k in range(1, 40)]  # 1/k!, k = 1, 2, ...


@pytest.fixture(scope="module")
def c1():
    ch = Channels(6, 6)
    return ch, Oracle(ch)


def one_body_flat(ch, E0, f):
    return dense_to_flat(ch, E0, f, np.zeros((ch.M,) * 4))


def mat_norm(ch, f):
    """mat_2_norm of a one-body operator: fhh, fpp, fph (no hp block, no E0),
    src/ME_general.f90:2112-2138."""
    n = ch.n
    return np.sqrt((f[:n, :n] ** 2).sum() + (f[n:, n:] ** 2).sum() + (f[n:, :n] ** 2).sum())


def dense_bch(ch, Om, E0, F, cof):
    """sum_k cof(k) ad_Om^k F with the reference's stopping rule, in dense
    matrices: returns (E0, F, terms)."""
    n = ch.n
    X, hs_f, hs_e = F.copy(), F.copy(), E0
    for k in range(1, 12):
        prev_norm = mat_norm(ch, hs_f)
        X = Om @ X - X @ Om
        hs_f = hs_f + cof[k - 1] * X
        hs_e = hs_e + cof[k - 1] * np.trace(X[:n, :n])
        if mat_norm(ch, X) * BCH_COF[k - 1] / prev_norm < R4_1EM5:
            return hs_e, hs_f, k
    return hs_e, hs_f, 11


def channel_mask(ch):
    return (ch.ml[:, None] == ch.ml[None, :]) & (ch.ms[:, None] == ch.ms[None, :])


def test_bch_of_zero_omega_is_h(c1):
    ch, o = c1
    h = random_flat(ch, np.random.default_rng(1), 1)
    hs, nt = o.bch_expand(o.zeros(), -1, h)
    assert nt == 1 and np.array_equal(hs, h)


def test_bch_of_commuting_operators_is_h(c1):
    """[Omega, N] = 0 for the one-body identity N: one term, H unchanged."""
    ch, o = c1
    om = random_flat(ch, np.random.default_rng(2), -1)
    N = one_body_flat(ch, float(ch.n), np.eye(ch.M))
    hs, nt = o.bch_expand(om, -1, N)
    assert nt == 1 and np.abs(hs - N).max() <= 1e-13


@pytest.mark.parametrize("theta", [0.05, 0.3, 0.8])
def test_bch_two_level_rotation_closed_form(c1, theta):
    """Omega rotates one hole i into one particle a (Omega_ai = theta =
    -Omega_ia) on a diagonal f: ad^k f alternates between sigma_x and sigma_z
    in the (i, a) plane with weight d (2 theta)^k, d = (e_i - e_a)/2, so the
    series is known term by term and sums to the rotation; E0 gains
    (e_a - e_i) sin^2 theta."""
    ch, o = c1
    n, M = ch.n, ch.M
    e = 1.0 + np.arange(M) * 0.37
    F = np.diag(e)
    i = 2
    a = next(p for p in range(n, M) if ch.ml[p] == ch.ml[i] and ch.ms[p] == ch.ms[i])
    Om = np.zeros((M, M))
    Om[a, i], Om[i, a] = theta, -theta
    h = one_body_flat(ch, float(e[:n].sum()), F)
    om = one_body_flat(ch, 0.0, Om)
    om[ch.off_fhh:ch.off_fhh + n * n] = 0.0
    hs, nt = o.bch_expand(om, -1, h)
    E0, Fs, f = flat_to_dense(ch, hs, 1)
    # term count: the closed-form norms of ad^k f and of the partial sums
    d = 0.5 * (e[i] - e[a])
    partial_x, partial_z = 0.0, d      # sigma_x and sigma_z weights of the (i, a) block
    rest = np.delete(e, [i, a])
    expect = 11
    for k in range(1, 12):
        prev = np.sqrt((rest ** 2).sum() + (0.5 * (e[i] + e[a]) + partial_z) ** 2
                       + (0.5 * (e[i] + e[a]) - partial_z) ** 2 + partial_x ** 2)
        w = d * (2 * theta) ** k * (1, -1, -1, 1)[k % 4 - 1 if k % 4 else 3]
        if k % 2:
            partial_x += BCH_COF[k - 1] * w
            norm_k = abs(w)
        else:
            partial_z += BCH_COF[k - 1] * w
            norm_k = np.sqrt(2) * abs(w)
        if norm_k * BCH_COF[k - 1] / prev < R4_1EM5:
            expect = k
            break
    assert nt == expect
    # the literal-coefficient series, term by term
    assert abs(Fs[a, i] - partial_x) <= 1e-13 and abs(Fs[i, i] - (0.5 * (e[i] + e[a]) + partial_z)) <= 1e-13
    assert abs(E0 - (e[:n].sum() + partial_z - d)) <= 1e-12
    # against the exact rotation: the omitted terms and the truncated
    # coefficients (relative deviations <= 2e-10 from 1/k!) bound the error
    U = expm(Om)
    R = U @ F @ U.T
    omitted = sum(abs(d) * (2 * theta) ** k * EXACT[k - 1] for k in range(nt + 1, 30))
    tol = omitted + 2e-10 * abs(d) * np.exp(2 * theta) + 1e-13
    assert np.abs(Fs - R).max() <= tol
    assert abs(E0 - (e[:n].sum() + (e[a] - e[i]) * np.sin(theta) ** 2)) <= tol


@pytest.mark.parametrize("seed,scale", [(3, 0.05), (4, 0.2), (5, 0.6)])
def test_bch_one_body_series_and_term_count(c1, seed, scale):
    """Random one-body Omega (anti-hermitian) and H (hermitian, dense within
    the (m_l, m_s) channels): the oracle's BCH equals the dense matrix series
    with the literal coefficients and the same term count; it differs from
    the exact-coefficient series only by the coefficients' truncation, and
    from exp(Omega) F exp(-Omega) by at most the omitted terms."""
    ch, o = c1
    n, M = ch.n, ch.M
    rng = np.random.default_rng(seed)
    mask = channel_mask(ch)
    A = rng.standard_normal((M, M)) * mask * scale
    Om = A - A.T
    B = rng.standard_normal((M, M)) * mask
    F = 0.5 * (B + B.T) + np.diag(np.arange(M) * 0.5)
    E0 = 3.0
    h, om = one_body_flat(ch, E0, F), one_body_flat(ch, 0.0, Om)
    hs, nt = o.bch_expand(om, -1, h)
    e_lit, f_lit, k_lit = dense_bch(ch, Om, E0, F, BCH_COF)
    assert nt == k_lit
    got_e, got_f, _ = flat_to_dense(ch, hs, 1)
    scale_f = np.abs(F).max()
    assert np.abs(got_f - f_lit).max() <= 1e-12 * scale_f and abs(got_e - e_lit) <= 1e-12 * scale_f
    e_ex, f_ex, _ = dense_bch(ch, Om, E0, F, EXACT)
    assert np.abs(f_lit - f_ex).max() <= 3e-10 * np.abs(f_ex).max()
    U = expm(Om)
    R = U @ F @ U.T
    assert np.abs(got_f - R).max() <= 20 * R4_1EM5 * mat_norm(ch, F)
    assert abs(got_e - (E0 + np.trace(R[:n, :n]) - np.trace(F[:n, :n]))) <= 20 * R4_1EM5 * mat_norm(ch, F)


def test_bch_coefficients_are_the_reference_literals():
    """cof(k) as written at src/magnus_IMSRG.f90:265-267: 1/k! to the
    printed digits (relative deviations below 2e-10, largest at k = 11)."""
    rel = [abs(c - x) / x for c, x in zip(BCH_COF, EXACT[:11])]
    assert BCH_COF[0] == 1.0 and BCH_COF[1] == 0.5
    assert max(rel) < 2e-10 and rel[10] > 1e-10


def test_first_step_is_euler_then_bch(c1):
    """One accepted step: Omega = stp * eta(H(0)) (MAGNUS_EXPAND as shipped,
    euler_step) and H(s) = BCH(Omega, H(0))."""
    ch, o = c1
    h0 = model_hamiltonian(ch, 1)
    r = magnus_flow(o, h0, max_iter=1)
    eta0 = o.build_white(h0, 1)
    assert np.array_equal(r["omega"], o.add(o.zeros(), 1.0, eta0, R4_0P7))
    hs, nt = o.bch_expand(r["omega"], -1, h0)
    assert np.array_equal(r["hs"], hs) and r["bch_terms"] == nt
    assert r["iterations"] == 1 and abs(r["s"] - R4_0P7) == 0.0


def test_forced_rejections_halve_the_step(c1):
    """An oversized first step is rejected while ||eta|| > 1.5 ||eta_0||:
    s falls back, stp -> stp/2 + 1e-3 (REAL(4) 1e-3 widened), and ||eta_0||
    (column 3 of the step line) is the initial norm throughout."""
    ch, o = c1
    h0 = model_hamiltonian(ch, 1)
    stp0 = 40.0
    r = magnus_flow(o, h0, stp=stp0)
    tr = r["trace"]
    nrm1 = o.norm(o.build_white(h0, 1))
    assert r["rejected"] >= 2 and np.all(tr[:, 3] == nrm1)
    stp, s_acc, it = stp0, 0.0, 0
    for row in tr:
        assert row[0] == it and row[1] == s_acc + stp
        if 1.5 * nrm1 < row[4]:
            stp = stp / 2.0 + R4_1EM3
        else:
            s_acc, it = row[1], it + 1
    assert R4_1EM3 == np.float32(1e-3) and it == r["iterations"] and len(tr) == it + r["rejected"]


def test_one_body_flow_reaches_exact_decoupling(c1):
    """With no two-body part the IM-SRG(2) flow is exact: the White
    generator rotates f until the hole and particle spaces decouple, and E0
    becomes the sum of the N lowest eigenvalues of f (stopping rule |dE/ds|,
    since MBPT2 vanishes without a two-body part)."""
    ch, o = c1
    n, M = ch.n, ch.M
    rng = np.random.default_rng(8)
    mask = channel_mask(ch)
    e = 1.0 * (2 * ch.qn[:, 0] + np.abs(ch.qn[:, 1]) + 1)
    B = rng.standard_normal((M, M)) * mask * 0.15
    F = np.diag(e) + 0.5 * (B + B.T)
    h0 = one_body_flat(ch, float(np.trace(F[:n, :n])), F)
    r = magnus_flow(o, h0, stop_rule="deds", conv_deds=1e-12)
    exact = np.sort(np.linalg.eigvalsh(F))[:n].sum()
    _, Om, _ = flat_to_dense(ch, r["omega"], -1)
    # H(s) is the literal-coefficient series of the final Omega ...
    e_lit, f_lit, _ = dense_bch(ch, Om, h0[0], F, BCH_COF)
    _, fs, _ = flat_to_dense(ch, r["hs"], 1)
    assert np.abs(fs - f_lit).max() <= 1e-12 * np.abs(F).max() and abs(r["e0"] - e_lit) <= 1e-11
    # ... and that Omega is the decoupling rotation: exp(Omega) f exp(-Omega)
    # has a vanishing ph block and its hole trace is the exact energy (to
    # second order in the residual ph block)
    U = expm(Om)
    R = U @ F @ U.T
    assert np.abs(R[n:, :n]).max() <= 1e-5
    assert abs(np.trace(R[:n, :n]) - exact) <= 1e-9
    # the flow's own E0 carries the 1e-5 truncation of each BCH series
    assert abs(r["e0"] - exact) <= 1e-5 * mat_norm(ch, F) * 0.1
