"""Seeded synthetic inputs in the blocked layout (test data, not physics).

``model_hamiltonian`` is a hermitian operator shaped like a quantum-dot
Hamiltonian in an oscillator basis: diagonal one-body energies
hw*(2n+|m|+1) plus a weak random two-body interaction on every stored block,
with E0 = sum of hole energies.  White-generator denominators stay well away
from zero, so the Magnus flow converges.
"""
from __future__ import annotations

import numpy as np

from tests.pins.layout import random_flat


def model_hamiltonian(ch, seed: int, strength: float = 0.05, hw: float = 1.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    v = random_flat(ch, rng, 1, scale=strength)
    n, mp = ch.n, ch.mp
    e = hw * (2 * ch.qn[:, 0] + np.abs(ch.qn[:, 1]) + 1).astype(float)
    v[ch.off_fhh:ch.off_fhh + n * n] = np.diag(e[:n]).T.reshape(-1)
    v[ch.off_fpp:ch.off_fpp + mp * mp] = np.diag(e[n:]).T.reshape(-1)
    fph = 0.02 * rng.standard_normal((mp, n)) * (
        (ch.ml[n:, None] == ch.ml[None, :n]) & (ch.ms[n:, None] == ch.ms[None, :n]))
    v[ch.off_fph:ch.off_fph + mp * n] = fph.T.reshape(-1)
    v[0] = e[:n].sum()
    return v
