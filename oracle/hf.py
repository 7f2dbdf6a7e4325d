"""Hartree-Fock input H(0) -- oracle side.  TEST INFRASTRUCTURE ONLY.

Restates construct_two_particle_HF_BASIS (src/HFQDmod.f90:9-254), with the
interaction from the analytic v_int table instead of ME.bin (the file is not
in the reference tree; get_ints, src/ME_general.f90:746-798, builds the same
antisymmetrised elements from v_int):
  * T = hw (2n + |m| + 1) (calc_h0, src/ME_general.f90:800-821)
  * SCF: rho = C_occ C_occ^T with the first N states occupied (calcDen,
    :868-891); V_ik = sum_jl rho_lj <j i|V|l k>_A (construct_V,
    src/HFQDmod.f90:260-290); F = T + V diagonalised per (m_l, m_s) block,
    eigenvectors assigned to the block's states in states order
    (eigvecs_to_normal); stop when sum_{i<N} |e_old - e| <= 1e-6 (:196-221)
  * F_HF = C^T F C (:228-229); two-body elements transformed per block
    (new_VMAT / HF_Vmatrix, src/ME_general.f90:1362-1492, 2179-2202)
  * E0 = sum_h e_h - sum_{h<h'} <hh'|V|hh'> (HF_E2, src/HFQDmod.f90:292-323)
Eigenvector phases are fixed by making the largest-magnitude component of
every eigenvector positive (LAPACK's choice is unspecified; the energies do
not depend on it).
"""
from __future__ import annotations

import numpy as np

from .channels import Channels
from .vint import spatial_table


def antisym_ho(ch: Channels, hw: float, table=None):
    """Dense <pq|V|rs>_A in the oscillator spin-orbital basis [M]*4."""
    orbs, T = table if table is not None else spatial_table(ch.R, hw)
    index = {o: k for k, o in enumerate(orbs)}
    sp = np.array([index[(int(n), int(m))] for n, m, _ in ch.qn])
    s = ch.qn[:, 2]
    D = T[np.ix_(sp, sp, sp, sp)]
    same = (s[:, None] == s[None, :]).astype(float)
    direct = D * same[:, None, :, None] * same[None, :, None, :]        # s_p = s_r, s_q = s_s
    G = direct - direct.transpose(0, 1, 3, 2)
    return G


def canonical_sign(C):
    k = np.argmax(np.abs(C), axis=0)
    sg = np.sign(C[k, np.arange(C.shape[1])])
    sg[sg == 0] = 1
    return C * sg


def hartree_fock(ch: Channels, hw: float, G=None, tol: float = 1e-6, max_iter: int = 1000):
    n, M = ch.n, ch.M
    if G is None:
        G = antisym_ho(ch, hw)
    Tm = np.diag(hw * (2 * ch.qn[:, 0] + np.abs(ch.qn[:, 1]) + 1).astype(float))
    C = np.eye(M)
    eig = np.zeros(M)
    chan = {}
    for i in range(M):
        chan.setdefault((int(ch.ml[i]), int(ch.ms[i])), []).append(i)
    it = 0
    crit = 1.0
    while crit > tol and it < max_iter:
        it += 1
        rho = C[:, :n] @ C[:, :n].T
        V = np.einsum("lj,jilk->ik", rho, G, optimize=True)
        F = Tm + V
        eold = eig.copy()
        Cn = np.zeros((M, M))
        for idx in chan.values():
            w, U = np.linalg.eigh(F[np.ix_(idx, idx)])
            U = canonical_sign(U)
            for k, st in enumerate(idx):
                Cn[idx, st] = U[:, k]
                eig[st] = w[k]
        crit = np.abs(eold[:n] - eig[:n]).sum()
        C = Cn
        Flast = F
    FHF = C.T @ Flast @ C
    GHF = np.einsum("pa,qb,pqrs,rc,sd->abcd", C, C, G, C, C, optimize=True)
    return dict(C=C, eig=eig, F=FHF, G=GHF, iterations=it)


def hf_hamiltonian(ch: Channels, hw: float, G=None):
    """H(0) in the HF basis as a flat vectorize-order vector (+ HF data)."""
    from tests.pins.layout import dense_to_flat  # the layout packer of the pins
    hf = hartree_fock(ch, hw, G)
    n = ch.n
    e0 = hf["eig"][:n].sum() - 0.5 * np.einsum("ijij->", hf["G"][:n, :n, :n, :n])
    v = dense_to_flat(ch, e0, hf["F"], hf["G"])
    return v, hf
