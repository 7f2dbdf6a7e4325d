"""Input side (host, no GPU): the analytic Coulomb element v_int and the
Hartree-Fock H(0), product (libqdimsrg.so) against oracle and closed forms."""
import math

import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.channels import Channels
from oracle.hf import hf_hamiltonian
from oracle.vint import v_int, v_int_exact

S = math.sqrt(math.pi / 2)
Z, P, M_ = (0, 0, 1), (0, 1, 1), (0, -1, 1)


@pytest.mark.parametrize("args,ratio", [((Z, Z, Z, Z), 1.0), ((Z, P, P, Z), 0.75), ((Z, P, Z, P), 0.25),
                                        ((P, M_, M_, P), 11 / 16), ((P, M_, P, M_), 3 / 16)])
def test_v_int_closed_forms(args, ratio):
    """Textbook 2D oscillator Coulomb elements in units of sqrt(pi/2) (hw=1);
    v_int(i,j,k,l) = <ij|V|lk>."""
    assert abs(qd.v_int(*args) - ratio * S) < 1e-15
    assert abs(v_int(*args) - ratio * S) < 1e-15


def test_v_int_scales_with_sqrt_hw():
    a = ((1, 2, 1), (0, -1, -1), (0, 1, -1), (1, 0, 1))
    assert math.isclose(qd.v_int(*a, 0.28), math.sqrt(0.28) * qd.v_int(*a, 1.0), rel_tol=1e-14)


def test_v_int_spin_selection():
    assert qd.v_int((0, 0, 1), (0, 0, -1), (0, 0, 1), (0, 0, -1)) == 0.0
    assert qd.v_int((0, 0, 1), (0, 0, -1), (0, 0, -1), (0, 0, 1)) != 0.0


def test_v_int_matches_exact_arithmetic():
    """binary64 sums against the same sum in exact rationals (50 digits), on a
    seeded sample of m_l-conserving elements up to shell 6."""
    rng = np.random.default_rng(21)
    orbs = [((R - abs(m)) // 2, m) for R in range(7) for m in range(-R, R + 1, 2)]
    done = 0
    while done < 40:
        a, b, c, d = (orbs[i] for i in rng.integers(0, len(orbs), 4))
        if a[1] + b[1] != c[1] + d[1]:
            continue
        q = [(x[0], x[1], 1) for x in (a, b, d, c)]
        ex = float(v_int_exact(*q))
        assert abs(qd.v_int(*q) - ex) <= 1e-13 * max(abs(ex), 1e-3)
        # the oracle sums in binary64 like the reference: cancellation leaves ~1e-14 absolute
        assert abs(v_int(*q) - ex) <= 2e-13
        done += 1


def test_hf_two_electrons_one_shell():
    """N=2, R=1: E_HF = 2 hw + sqrt(pi hw / 2)."""
    for hw in (1.0, 0.5):
        h0, eig, _ = qd.hf_hamiltonian(2, 1, hw)
        assert abs(h0[0] - (2 * hw + math.sqrt(math.pi * hw / 2))) < 1e-14


@pytest.mark.parametrize("N,R,hw", [(2, 3, 1.0), (6, 6, 1.0), (6, 5, 0.5), (12, 5, 1.0)])
def test_hf_matches_oracle(N, R, hw):
    h0, eig, it = qd.hf_hamiltonian(N, R, hw)
    ch = Channels(N, R)
    v, hf = hf_hamiltonian(ch, hw)
    assert it == hf["iterations"]
    assert np.abs(eig - hf["eig"]).max() <= 1e-12 * np.abs(hf["eig"]).max()
    assert np.abs(h0 - v).max() <= 1e-12 * np.abs(v).max()
    # HF basis: the Fock matrix is diagonal with the orbital energies on it
    n, mp = ch.n, ch.mp
    fhh = h0[ch.off_fhh:ch.off_fhh + n * n].reshape(n, n).T
    fpp = h0[ch.off_fpp:ch.off_fpp + mp * mp].reshape(mp, mp).T
    fph = h0[ch.off_fph:ch.off_fph + mp * n]
    assert np.abs(fph).max() < 1e-12
    assert np.allclose(np.diag(fhh), eig[:n], atol=1e-12) and np.allclose(np.diag(fpp), eig[n:], atol=1e-12)
    # E0 = sum_h <h|t|h> + 1/2 sum_hh' <hh'|V|hh'> (independent of the HF_E2 form);
    # the two forms agree to the SCF tolerance (1e-6 on the orbital energies)
    C = hf["C"]
    t = np.diag(hw * (2 * ch.qn[:, 0] + np.abs(ch.qn[:, 1]) + 1).astype(float))
    thf = C.T @ t @ C
    e_alt = np.trace(thf[:n, :n]) + 0.5 * np.einsum("ijij->", hf["G"][:n, :n, :n, :n])
    assert abs(h0[0] - e_alt) < 1e-6


def test_hf_energy_literature_n6():
    """N=6, hw=1, R=6: E_HF = 20.7203 (the R -> large limit is 20.719)."""
    h0, _, _ = qd.hf_hamiltonian(6, 6, 1.0)
    assert abs(h0[0] - 20.7203) < 1e-4
