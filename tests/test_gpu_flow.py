"""GPU Magnus flow (s_evolve) against the oracle's flow on a seeded model
Hamiltonian at N=6/R=6: same accepted/rejected step sequence, same BCH term
counts, E0 per step and final Omega / H(s) to 1e-10 relative."""
import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.channels import Channels
from oracle.flow import magnus_flow
from oracle.oracle import Oracle
from tests.pins.hamiltonians import model_hamiltonian

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,R,seed", [(6, 6, 1), (2, 4, 2)])
def test_magnus_flow_matches_oracle(N, R, seed):
    print(f"[case] magnus flow N={N} R={R} seed={seed}", flush=True)
    ch = Channels(N, R)
    o = Oracle(ch)
    h0 = model_hamiltonian(ch, seed)
    ref = magnus_flow(o, h0)
    ctx = qd.Context(N, R)
    H0, HS, OM = ctx.operator(1, h0), ctx.operator(1), ctx.operator(-1)
    res, trace = qd.magnus_flow(ctx, H0, HS, OM)
    assert res.converged
    assert (res.iterations, res.rejected, res.bch_terms) == (ref["iterations"], ref["rejected"], ref["bch_terms"])
    assert np.allclose(trace[:, :2], ref["trace"][:, :2], rtol=0, atol=1e-12)
    assert np.abs(trace[:, 2] - ref["trace"][:, 2]).max() <= 1e-10 * abs(ref["e0"])
    hs, om = HS.download(), OM.download()
    assert np.abs(hs - ref["hs"]).max() <= 1e-10 * np.abs(ref["hs"]).max()
    assert np.abs(om - ref["omega"]).max() <= 1e-10 * np.abs(ref["omega"]).max()
    # the flow lowers E0 below the reference-state energy
    assert res.e0 < h0[0]
    ctx.close()


def test_warm_start_from_converged_omega():
    """Restarting from the converged Omega needs (almost) no further steps."""
    ch = Channels(6, 6)
    h0 = model_hamiltonian(ch, 1)
    ctx = qd.Context(6, 6)
    H0, HS, OM = ctx.operator(1, h0), ctx.operator(1), ctx.operator(-1)
    cold, _ = qd.magnus_flow(ctx, H0, HS, OM)
    e_cold = cold.e0
    HS2 = ctx.operator(1)
    warm, _ = qd.magnus_flow(ctx, H0, HS2, OM, warm_start=True)
    assert warm.converged and warm.iterations <= 1
    assert abs(warm.e0 - e_cold) < 1e-8
    ctx.close()


@pytest.mark.parametrize("warm", [False, True])
def test_diis_flow_matches_oracle(warm):
    """Pulay-accelerated flow (qd_flow_params.diis = 6) against the oracle's
    restatement, cold and from a perturbed converged Omega: same iteration
    count, E_gs to 1e-10 relative, H(s) and Omega to 1e-9 relative (the
    weights come from a small solve on rounded dot products, so the
    trajectories agree to rounding, not bitwise)."""
    print(f"[case] diis flow N=6 R=6 warm={warm}", flush=True)
    ch = Channels(6, 6)
    o = Oracle(ch)
    h0 = model_hamiltonian(ch, 1)
    om0 = None
    if warm:
        om0 = magnus_flow(o, h0)["omega"]
        om0 = om0 * (1 + 0.01 * np.random.default_rng(5).standard_normal(om0.size))
    ref = magnus_flow(o, h0, omega0=om0, diis=6)
    ctx = qd.Context(6, 6)
    H0, HS = ctx.operator(1, h0), ctx.operator(1)
    OM = ctx.operator(-1, om0) if warm else ctx.operator(-1)
    res, _ = qd.magnus_flow(ctx, H0, HS, OM, warm_start=warm, diis=6)
    print(f"iterations {res.iterations} (oracle {ref['iterations']}), E {res.e0:.12f}")
    assert res.converged and res.iterations == ref["iterations"]
    assert abs(res.e0 - ref["e0"]) <= 1e-10 * abs(ref["e0"])
    assert np.abs(HS.download() - ref["hs"]).max() <= 1e-9 * np.abs(ref["hs"]).max()
    assert np.abs(OM.download() - ref["omega"]).max() <= 1e-9 * np.abs(ref["omega"]).max()
    ctx.close()
