"""The ODE oracle (oracle/adams.py, a restatement of the Shampine-Gordon
Adams PECE code the reference uses, src/adams_ode.f90) pinned by problems
with closed-form solutions and by the bookkeeping the algorithm promises."""
import math

import numpy as np
import pytest

from oracle.adams import MAXNUM, Adams


def run(f, y0, t0, outs, rel, ab):
    a = Adams(f, len(y0))
    y, t, flag = np.array(y0, dtype=float), t0, 1
    res = []
    for tout in outs:
        y, t, rel, ab, flag = a.ode(y, t, tout, rel, ab, flag)
        assert flag == 2 and t == tout
        res.append(y.copy())
    return a, res


@pytest.mark.parametrize("tol", [1e-6, 1e-9])
def test_exponential_decay(tol):
    outs = [0.5 * i for i in range(1, 21)]
    a, res = run(lambda t, y: -y, [1.0, 2.0], 0.0, outs, tol, tol)
    for t, y in zip(outs, res):
        assert np.abs(y - np.array([1.0, 2.0]) * math.exp(-t)).max() <= 50 * tol
    assert max(k for _, _, k in a.log) >= 4  # smooth problem: the order is raised


def test_harmonic_oscillator_energy_and_phase():
    outs = [1.0 * i for i in range(1, 31)]
    a, res = run(lambda t, y: np.array([y[1], -y[0]]), [1.0, 0.0], 0.0, outs, 1e-10, 1e-10)
    for t, y in zip(outs, res):
        assert abs(y[0] - math.cos(t)) <= 1e-7 and abs(y[1] + math.sin(t)) <= 1e-7
    # PECE: two evaluations per successful step (plus the start); the order
    # climbs on this smooth problem (phase 1 doubles h until an estimate fails)
    assert a.nfe <= 2 * a.nsteps + 2 * 20 and max(k for _, _, k in a.log) >= 8


def test_gaussian_and_linear_system():
    outs = [0.25 * i for i in range(1, 13)]
    _, res = run(lambda t, y: -2.0 * t * y, [1.0], 0.0, outs, 1e-9, 1e-12)
    for t, y in zip(outs, res):
        assert abs(y[0] - math.exp(-t * t)) <= 1e-7
    A = np.array([[-1.0, 2.0, 0.0], [-2.0, -1.0, 0.5], [0.0, 0.0, -3.0]])
    from scipy.linalg import expm
    y0 = np.array([1.0, -1.0, 2.0])
    _, res = run(lambda t, y: A @ y, y0, 0.0, [0.5, 1.0, 2.0], 1e-10, 1e-10)
    for t, y in zip([0.5, 1.0, 2.0], res):
        assert np.abs(y - expm(A * t) @ y0).max() <= 1e-7


def test_too_many_steps_flag():
    """iflag 4 after MAXNUM steps without reaching tout (src/adams_ode.f90:510)."""
    a = Adams(lambda t, y: np.array([math.cos(50.0 * t)]), 1)
    y, t, rel, ab, flag = a.ode(np.array([0.0]), 0.0, 1000.0, 1e-12, 1e-12, 1)
    assert flag == 4 and a.nsteps == MAXNUM and 0.0 < t < 1000.0
    assert abs(y[0] - math.sin(50.0 * t) / 50.0) <= 1e-8


def test_bad_input_flag():
    a = Adams(lambda t, y: -y, 1)
    assert a.ode(np.array([1.0]), 1.0, 1.0, 1e-6, 1e-6, 1)[4] == 6      # t == tout
    assert a.ode(np.array([1.0]), 0.0, 1.0, -1e-6, 1e-6, 1)[4] == 6     # negative tolerance


def test_gs_loop_flag3_without_progress_ends_it():
    """gs_IMSRG's loop (src/gs_IMSRG.f90:100-120) passes the flag back into
    the next ode call.  Tolerances below what the arithmetic supports make
    the first call return 3 with raised tolerances before any step: y and t
    are unchanged, so the loop's test |dE/E| sees 0 and the loop ends there
    (a quirk of the reference kept as is)."""
    from oracle.adams import gs_flow
    from oracle.channels import Channels
    from oracle.oracle import Oracle
    from tests.pins.hamiltonians import model_hamiltonian
    ch = Channels(2, 3)
    o = Oracle(ch)
    h0 = model_hamiltonian(ch, 4)
    tight = gs_flow(o, h0, relerr=1e-18, abserr=1e-18)
    assert tight["flags"] == [3] and tight["outputs"] == 1 and tight["s"] == 0.0
    assert tight["relerr"] > 1e-18 and np.array_equal(tight["hs"], h0)


def test_gs_loop_forced_too_many_steps_continues(monkeypatch):
    """With the step limit per call cut to 3 (the reference's is 500), ode
    returns 4 before the output point; the loop's next call continues and the
    flow ends where the unlimited one does."""
    import oracle.adams as ad
    from oracle.channels import Channels
    from oracle.oracle import Oracle
    from tests.pins.hamiltonians import model_hamiltonian
    ch = Channels(2, 3)
    o = Oracle(ch)
    h0 = model_hamiltonian(ch, 4)
    ref = ad.gs_flow(o, h0)
    monkeypatch.setattr(ad, "MAXNUM", 3)
    cut = ad.gs_flow(o, h0)
    assert 4 in cut["flags"] and cut["flags"][-1] == 2
    assert abs(cut["e0"] - ref["e0"]) <= 1e-6 * abs(ref["e0"])
