"""GPU Magnus flow (s_evolve) against the oracle's flow on a seeded model
Hamiltonian at N=6/R=6: same accepted/rejected step sequence, same BCH term
counts, E0 per step and final Omega / H(s) to 1e-10 relative; as shipped
(zero CC inputs) and with calc_cc-filled CC inputs (the Pandya term)."""
import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.channels import Channels
from oracle.flow import magnus_flow
from oracle.oracle import Oracle
from tests.pins.hamiltonians import model_hamiltonian

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,R,seed,pandya", [(6, 6, 1, False), (2, 4, 2, False), (6, 6, 1, True)])
def test_magnus_flow_matches_oracle(N, R, seed, pandya):
    print(f"[case] magnus flow N={N} R={R} seed={seed} pandya={pandya}", flush=True)
    ch = Channels(N, R)
    o = Oracle(ch)
    h0 = model_hamiltonian(ch, seed)
    ref = magnus_flow(o, h0, pandya=pandya)
    ctx = qd.Context(N, R)
    H0, HS, OM = ctx.operator(1, h0), ctx.operator(1), ctx.operator(-1)
    res, trace = qd.magnus_flow(ctx, H0, HS, OM, pandya=pandya)
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


def test_gs_ode_flow_matches_oracle():
    """The ODE path (gs_IMSRG: Adams PECE over dGam, output every 0.1 until
    |dE/E| < 1e-7) on the GPU against the oracle's restatement at N=6 R=6
    hw=1 (HF basis): same number of outputs, integrator steps and RHS
    evaluations, E_gs to 1e-10 relative, H(s) to 1e-8 relative (the step
    and order decisions are taken on reductions whose rounding differs)."""
    from oracle.adams import gs_flow
    print("[case] gs ode flow N=6 R=6", flush=True)
    ch = Channels(6, 6)
    o = Oracle(ch)
    h0, _, _ = qd.hf_hamiltonian(6, 6, 1.0)
    ref = gs_flow(o, h0)
    ctx = qd.Context(6, 6)
    H0, HS = ctx.operator(1, h0), ctx.operator(1)
    res, trace = qd.gs_flow(ctx, H0, HS)
    print(f"outputs {res.outputs} (oracle {ref['outputs']}), steps {res.nsteps} ({ref['nsteps']}), "
          f"nfe {res.nfe} ({ref['nfe']}), E {res.e0:.12f} ({ref['e0']:.12f})")
    assert res.converged and res.iflag == 2
    assert (res.outputs, res.nsteps, res.nfe) == (ref["outputs"], ref["nsteps"], ref["nfe"])
    assert abs(res.e0 - ref["e0"]) <= 1e-10 * abs(ref["e0"])
    assert np.abs(trace[:, 2] - ref["trace"][:, 2]).max() <= 1e-10 * abs(ref["e0"])
    assert np.abs(HS.download() - ref["hs"]).max() <= 1e-8 * np.abs(ref["hs"]).max()
    ctx.close()


def test_ode_integrate_entry_point():
    """qd_ode_integrate (the reference's ode call) over two output points
    equals the same two calls of the oracle's integrator."""
    from oracle.adams import GS_ABS, GS_REL, Adams
    ch = Channels(6, 6)
    o = Oracle(ch)
    h0 = model_hamiltonian(ch, 3)

    def dgam(_t, y):
        return o.commutator(o.build_white(y, 1), -1, y, 1, pandya=True)
    a = Adams(dgam, ch.neq)
    y, t, rel, ab, flag = h0.copy(), 0.0, GS_REL, GS_ABS, 1
    ctx = qd.Context(6, 6)
    ode = qd.ODE(ctx)
    Y = ctx.operator(1, h0)
    tg, rg, ag, fg = 0.0, GS_REL, GS_ABS, 1
    for tout in (0.25, 0.5):
        y, t, rel, ab, flag = a.ode(y, t, tout, rel, ab, flag)
        tg, rg, ag, fg = ode.integrate(Y, tg, tout, rg, ag, fg)
        assert (fg, tg) == (flag, t) == (2, tout)
        assert np.abs(Y.download() - y).max() <= 1e-9 * np.abs(y).max()
    st = ode.stats()
    assert (st["nsteps"], st["nfe"]) == (a.nsteps, a.nfe)
    ode.close()
    ctx.close()


@pytest.mark.parametrize("pandya", [False, True])
def test_deds_stop_rule_matches_oracle(pandya):
    """The stopping rule |dE/ds| < 1e-8 Ha (dE/ds = zero-body part of
    [eta(s), H(s)]) at N=6 R=6 hw=1 (HF basis): same accepted and rejected
    steps as the oracle's flow under the same rule, E_gs to 1e-10 relative,
    and the final |dE/ds| below the threshold on both sides."""
    print(f"[case] deds rule N=6 R=6 pandya={pandya}", flush=True)
    ch = Channels(6, 6)
    o = Oracle(ch)
    h0, _, _ = qd.hf_hamiltonian(6, 6, 1.0)
    ref = magnus_flow(o, h0, pandya=pandya, stop_rule="deds", conv_deds=1e-8)
    ctx = qd.Context(6, 6)
    H0, HS, OM = ctx.operator(1, h0), ctx.operator(1), ctx.operator(-1)
    res, _ = qd.magnus_flow(ctx, H0, HS, OM, pandya=pandya, stop_rule="deds", conv_deds=1e-8)
    print(f"iterations {res.iterations} ({ref['iterations']}), dE/ds {res.deds:.3e} ({ref['deds']:.3e})")
    assert res.converged and abs(res.deds) <= 1e-8 and abs(ref["deds"]) <= 1e-8
    assert (res.iterations, res.rejected) == (ref["iterations"], ref["rejected"])
    assert abs(res.e0 - ref["e0"]) <= 1e-10 * abs(ref["e0"])
    assert abs(res.mbpt2 - ref["mbpt2"]) <= 1e-10 * max(abs(ref["mbpt2"]), 1e-12) + 1e-18
    ctx.close()


def test_gs_flow_continues_past_step_limit(monkeypatch):
    """qd_gs_flow with the integrator's step limit per call cut to 3 (the
    reference's is 500): ode returns 4 before output points, the loop calls
    it again with that flag, as gs_IMSRG does, and the flow matches the
    oracle's loop under the same limit (outputs, steps, RHS evaluations,
    E_gs to 1e-10)."""
    import oracle.adams as ad
    print("[case] gs flow maxnum=3 N=2 R=3", flush=True)
    ch = Channels(2, 3)
    o = Oracle(ch)
    h0 = model_hamiltonian(ch, 4)
    monkeypatch.setattr(ad, "MAXNUM", 3)
    ref = ad.gs_flow(o, h0)
    assert 4 in ref["flags"]
    ctx = qd.Context(2, 3)
    H0, HS = ctx.operator(1, h0), ctx.operator(1)
    res, trace = qd.gs_flow(ctx, H0, HS, maxnum=3)
    assert res.converged and res.iflag == 2
    assert (res.outputs, res.nsteps, res.nfe) == (ref["outputs"], ref["nsteps"], ref["nfe"])
    assert abs(res.e0 - ref["e0"]) <= 1e-10 * abs(ref["e0"])
    ctx.close()


def test_graph_cache_drops_destroyed_operands():
    """A graph captured on an operand set is dropped when one of its
    operators or CC sets is destroyed."""
    ch = Channels(6, 6)
    ctx = qd.Context(6, 6)
    ctx.set_graphs(True)
    rng = np.random.default_rng(9)
    from tests.pins.layout import random_flat
    X, Y = ctx.operator(-1, random_flat(ch, rng, -1)), ctx.operator(1, random_flat(ch, rng, 1))
    HCC, ECC = ctx.cc(1.0), ctx.cc(-1.0)
    for _ in range(3):
        Z = ctx.operator(1)
        qd.commutator_cc(X, Y, Z, HCC, ECC)
        assert ctx.graph_count() == 1
        Z.close()
        assert ctx.graph_count() == 0
    Z = ctx.operator(1)
    qd.commutator_cc(X, Y, Z, HCC, ECC)
    HCC.close()
    assert ctx.graph_count() == 0
    ctx.close()


def test_export_import_device_operator(tmp_path):
    """export_hamiltonian of a device operator, read back into another one:
    bit for bit, and equal to the host-vector export of the same values."""
    ch = Channels(6, 6)
    h0, _, _ = qd.hf_hamiltonian(6, 6, 1.0)
    ctx = qd.Context(6, 6)
    H, G = ctx.operator(1, h0), ctx.operator(1)
    p1, p2 = str(tmp_path / "ham_dev"), str(tmp_path / "ham_host")
    qd.export_hamiltonian(H, p1)
    qd.export_hamiltonian_host(6, 6, h0, p2)
    assert open(p1, "rb").read() == open(p2, "rb").read()
    qd.import_hamiltonian(G, p1)
    assert np.array_equal(G.download(), h0)
    ctx.close()
