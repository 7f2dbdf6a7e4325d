"""The Fortran driver programs of the reference's job scripts, run on the GPU
(build_job_script.py:112 "./run_magnus N hwd0 R", :223 "./gs_decouple N hwd0
R 0 1 20"): their stdout lines in the reference's edit descriptors, the TRAD
summary and the exported H(s), checked against the library's own flows."""
import os
import subprocess

import numpy as np
import pytest

import qd_imsrg_amd as qd

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "qd_imsrg_amd", "bin")


def run(prog, args, cwd):
    print(f"[case] {prog} {' '.join(args)}", flush=True)
    r = subprocess.run([os.path.join(BIN, prog), *args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout.splitlines()


def fixed(line, widths):
    out, k = [], 0
    for w in widths:
        out.append(line[k:k + w])
        k += w
    assert k == len(line.rstrip("\n")), (line, widths)
    return out


def test_run_magnus_lines_match_library(tmp_path):
    """'(I5,5(f18.9))' per step: iteration, s, E0, ||eta_0||, ||eta||, crit;
    spectrum_magnus.dat with the TDA block of (ML, MS) = (1, 0) below state
    20 (one line at the start, two per accepted step); then the EOM block of
    calculate_excited_states (10 states) and nothing after it (the reference
    stops at src/magnus_IMSRG.f90:144)."""
    lines = run("run_magnus", ["6", "1.0d0", "6", "1", "0", "20"], str(tmp_path))
    k = next(i for i, l in enumerate(lines) if not l.strip())
    steps = [l for l in lines[:k] if len(l) == 95]
    assert steps and all(len(l) == 95 or l.strip() == "faillll" for l in lines[:k]), lines[:3]
    rows = np.array([[float(x) for x in fixed(l, [5] + [18] * 5)] for l in steps])
    h0, _, _ = qd.hf_hamiltonian(6, 6, 1.0)
    ctx = qd.Context(6, 6)
    H0, HS, OM = ctx.operator(1, h0), ctx.operator(1), ctx.operator(-1)
    res, trace = qd.magnus_flow(ctx, H0, HS, OM)
    assert rows.shape[0] == trace.shape[0] == res.iterations + res.rejected
    assert np.abs(rows - trace).max() <= 1e-9 + 1e-9 * np.abs(trace).max()
    assert abs(rows[-1, 2] - res.e0) <= 1e-9
    # spectrum_magnus.dat
    spec = (tmp_path / "spectrum_magnus.dat").read_text().splitlines()
    nlev = 6 * 36 + 2
    assert len(spec) == 1 + 2 * res.iterations and all(len(l) == 12 * nlev for l in spec)
    last = np.array([float(x) for x in fixed(spec[-1], [12] * nlev)])
    tda = qd.tda_spectrum(HS, 1, 0, 20)
    assert 0 < tda.size < nlev - 2
    assert abs(last[1] - res.e0) <= 1e-7 and np.abs(last[2:2 + tda.size] - (tda + res.e0)).max() <= 1e-6
    assert np.all(last[2 + tda.size:] == 0.0)
    first = np.array([float(x) for x in fixed(spec[0], [12] * nlev)])
    assert first[0] == 0.0 and abs(first[1] - h0[0]) <= 1e-7
    # the EOM block
    tail = lines[k:]
    assert any("EXECUTING EOM CALCULATION FOR EXCITED STATES: Ml= 1 Ms= 0" in l for l in tail)
    assert any(l.startswith("Ground State Energy: ") and abs(float(l[21:]) - res.e0) <= 1e-8 for l in tail)
    eom = [l for l in tail if len(l) == 48]
    assert len(eom) == 10
    de, _, _ = qd.calculate_excited_states(HS, 1, 0, 10)
    got = np.array([[float(x) for x in fixed(l, [16] * 3)] for l in eom])
    assert np.abs(got[:, 0] - de).max() <= 1e-8 and np.abs(got[:, 1] - (de + res.e0)).max() <= 1e-8
    ctx.close()


def test_gs_decouple_outputs_match_library(tmp_path):
    """'(I5,3(e14.6))' per output point; '(I4,I5,3(f15.9),f15.4)' appended to
    ../output/TRAD_6_1.00.dat; H(s) in ../hamiltonians/ham_6_1.00_6."""
    work = tmp_path / "run"
    for d in ("run", "output", "hamiltonians"):
        (tmp_path / d).mkdir()
    lines = run("gs_decouple", ["6", "1.0d0", "6", "0", "1", "20"], str(work))
    assert lines[0].strip().startswith("Hartree-Fock Energy:")
    outs = [l for l in lines[1:] if len(l) == 47]
    rows = np.array([[float(x) for x in fixed(l, [5, 14, 14, 14])] for l in outs])
    h0, _, _ = qd.hf_hamiltonian(6, 6, 1.0)
    ctx = qd.Context(6, 6)
    H0, HS = ctx.operator(1, h0), ctx.operator(1)
    res, trace = qd.gs_flow(ctx, H0, HS)
    assert rows.shape[0] == res.outputs
    assert np.array_equal(rows[:, 0], trace[:, 0])
    # E14.6 prints 0.dddddd: half a unit of the sixth digit is up to 5e-6 relative
    err = np.abs(rows[:, 1:3] - trace[:, 1:3]) / np.abs(trace[:, 1:3])
    assert float(err.max()) <= 5.01e-6
    trad = (tmp_path / "output" / "TRAD_6_1.00.dat").read_text().splitlines()
    assert len(trad) == 1 and len(trad[0]) == 4 + 5 + 45 + 15
    f = fixed(trad[0], [4, 5, 15, 15, 15, 15])
    assert (int(f[0]), int(f[1])) == (6, 42)
    assert abs(float(f[2]) - h0[0]) <= 1e-9 and abs(float(f[4]) - res.e0) <= 1e-9
    ham = qd.import_hamiltonian_host(6, 6, str(tmp_path / "hamiltonians" / "ham_6_1.00_6"))
    hs = HS.download()
    assert np.abs(ham - hs).max() <= 1e-10 * np.abs(hs).max()
    # MS = 1: one particle attached and one removed (src/gs_IMSRG.f90:124-128)
    for mode, key, fname in ((1, "PARTICLE ATTACHED", "EOM_particle_attached.dat"),
                             (2, "PARTICLE REMOVED", "EOM_particle_removed.dat")):
        assert any(f"EXECUTING EOM CALCULATION FOR {key} STATES: Ml= 0 Ms= 1" in l for l in lines)
        de, w, _ = (qd.calculate_1p_attached if mode == 1 else qd.calculate_1h_removed)(HS, 0, 1, 5)
        rec = (tmp_path / "output" / fname).read_text().splitlines()
        assert len(rec) == 1 and len(rec[0]) == 4 * 5 + 8 + 4 * 17
        f = fixed(rec[0], [5] * 4 + [8] + [17] * 4)
        assert [int(x) for x in f[:4]] == [6, 2, 0, 1] and float(f[4]) == 1.0
        assert abs(float(f[6]) - de[0]) <= 1e-8 and abs(float(f[8]) - w[0]) <= 1e-8
    ctx.close()
