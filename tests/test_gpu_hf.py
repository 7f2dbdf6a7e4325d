"""Hartree-Fock on the device (qd_hf_hamiltonian_device: the Fock build of
each SCF iteration, src/HFQDmod.f90:196-221 and 260-290, and the two-body
basis transform new_VMAT, src/ME_general.f90:1362-1492) against the host
build, which tests/test_hf.py pins to the oracle.

Both run the same SCF; the device sums the mean field and T V T^T in its own
order, so H(0) agrees to FP64 rounding: 1e-11 relative to max |H(0)|, HF
energies 1e-11 Ha, the same SCF iteration count.  Sizes: the CPU-runnable
case, the batched form's N=20 R=16 hw=0.5 (whose SCF takes the restart from
the oscillator basis after a 2-cycle) and the precision form's N=42 R=20."""
import time

import numpy as np
import pytest

import qd_imsrg_amd as qd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,R,hw", [(6, 6, 1.0), (20, 16, 0.5), (42, 20, 0.28)])
def test_device_hf_matches_host(N, R, hw):
    print(f"[case] device HF N={N} R={R} hw={hw}", flush=True)
    t0 = time.perf_counter()
    h_ref, e_ref, it_ref = qd.hf_hamiltonian(N, R, hw)
    t1 = time.perf_counter()
    ctx = qd.Context(N, R)
    try:
        t2 = time.perf_counter()
        H, eig, it = qd.hf_hamiltonian_device(ctx, hw)
        t3 = time.perf_counter()
        got = H.download()
        print(f"host {t1 - t0:.2f} s, device {t3 - t2:.2f} s; {it} SCF iterations; E_HF = {got[0]:.10f}")
        assert it == it_ref
        assert np.abs(eig - e_ref).max() <= 1e-11
        scale = np.abs(h_ref).max()
        err = np.abs(got - h_ref).max()
        print(f"max |dH(0)| = {err:.3e} (scale {scale:.3e})")
        assert err <= 1e-11 * scale
        assert abs(got[0] - h_ref[0]) <= 1e-11 * abs(h_ref[0])
        # the device build overwrites whatever the operator held
        H.upload(np.full(got.size, 7.0))
        H2, _, _ = qd.hf_hamiltonian_device(ctx, hw, H)
        assert np.array_equal(H2.download(), got)
    finally:
        ctx.close()


def test_device_hf_rejects_antihermitian():
    ctx = qd.Context(2, 3)
    try:
        with pytest.raises(qd.QDError):
            qd.hf_hamiltonian_device(ctx, 1.0, ctx.operator(-1))
    finally:
        ctx.close()
