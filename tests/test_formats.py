"""On-disk formats, no GPU: ME.bin (the interaction file construct_two_particle_
HF_BASIS reads, src/HFQDmod.f90:97-129) and export/import_hamiltonian
(src/ME_general.f90:207-361).

The ME.bin pins build a file independently here (records from qd_v_int in
the reference's stream layout) and require the Hartree-Fock H(0) read from it
to equal the one built from v_int directly.  The export pins parse the file
with numpy in the reference's record order, derived here from the oracle's
channel tables."""
import os

import numpy as np
import pytest

import qd_imsrg_amd as qd
from oracle.channels import Channels

REC = np.dtype([("q", "<i4", 8), ("v", "<f8")])


def orbitals(nshell):
    """Spatial orbitals (n, m) in nm_index order (src/ME_general.f90:2140-2149)."""
    return [((R - abs(m)) // 2, m) for R in range(nshell) for m in range(-R, R + 1, 2)]


def own_me_bin(path, nshell):
    """ME.bin written by this test: Vdir = <ab|V|cd> = v_int(a, b, d, c) at
    hw = 1 for every m-conserving quadruple."""
    orb = orbitals(nshell)
    recs = []
    for a in orb:
        for b in orb:
            for c in orb:
                for d in orb:
                    if a[1] + b[1] != c[1] + d[1]:
                        continue
                    v = qd.v_int((a[0], a[1], 1), (b[0], b[1], 1), (d[0], d[1], 1), (c[0], c[1], 1), 1.0)
                    recs.append(((*a, *b, *c, *d), v))
    arr = np.zeros(len(recs), dtype=REC)
    for i, (q, v) in enumerate(recs):
        arr[i] = (q, v)
    arr.tofile(path)
    return len(recs)


@pytest.mark.parametrize("hw", [1.0, 0.5])
def test_hf_from_own_me_bin_equals_v_int(tmp_path, hw):
    p = str(tmp_path / "ME.bin")
    own_me_bin(p, 3)
    ref, eig_ref, _ = qd.hf_hamiltonian(2, 3, hw)
    got, eig, _ = qd.hf_hamiltonian(2, 3, hw, me_bin=p)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.abs(eig - eig_ref).max() <= 1e-13


def test_library_writer_round_trip_and_shell_cut(tmp_path):
    """The library's writer: records parse in the reference layout with
    Vdir = <ab|V|cd>; a file for 8 shells read at 6 shells (records beyond
    the basis skipped, src/HFQDmod.f90:113-116) gives the 6-shell H(0)
    bit for bit."""
    p8 = str(tmp_path / "ME8.bin")
    n = qd.me_bin_write(8, p8)
    arr = np.fromfile(p8, dtype=REC)
    assert arr.size == n and os.path.getsize(p8) == 40 * n
    rng = np.random.default_rng(3)
    for r in arr[rng.choice(n, 50, replace=False)]:
        q = r["q"]
        v = qd.v_int((q[0], q[1], 1), (q[2], q[3], 1), (q[6], q[7], 1), (q[4], q[5], 1), 1.0)
        assert abs(r["v"] - v) <= 1e-15 * max(1.0, abs(v))
    ref, _, _ = qd.hf_hamiltonian(6, 6, 1.0)
    got, _, _ = qd.hf_hamiltonian(6, 6, 1.0, me_bin=p8)
    assert np.array_equal(got, ref)


def test_truncated_me_bin_is_an_error(tmp_path):
    p = str(tmp_path / "ME.bin")
    qd.me_bin_write(3, p)
    with open(p, "ab") as f:
        f.write(b"\0" * 17)
    with pytest.raises(qd.QDError, match="truncated"):
        qd.hf_hamiltonian(2, 3, 1.0, me_bin=p)
    with pytest.raises(qd.QDError, match="cannot open"):
        qd.hf_hamiltonian(2, 3, 1.0, me_bin=str(tmp_path / "missing.bin"))


def reference_record_order(ch):
    """Flat vectorize-order index of each record of export_hamiltonian, in
    file order: E0, fhh, fph, fpp row by row, then per block Vhhhh, Vpppp,
    Vpphh, Vphph, Vphhh, Vppph row by row (src/ME_general.f90:218-279)."""
    n, mp = ch.n, ch.mp
    out = [0]

    def mat(off, rows, cols):
        out.extend(off + i + j * rows for i in range(rows) for j in range(cols))
    mat(ch.off_fhh, n, n)
    mat(ch.off_fph, mp, n)
    mat(ch.off_fpp, mp, mp)
    for q in range(ch.nblock):
        nh, np_, nb = int(ch.nh[q]), int(ch.np_[q]), int(ch.nb[q])
        o = ch.off[q]
        mat(o[0], nh, nh)
        mat(o[1], np_, np_)
        mat(o[2], np_, nh)
        mat(o[3], nb, nb)
        mat(o[5], nb, nh)
        mat(o[4], np_, nb)
    return np.array(out)


@pytest.mark.parametrize("N,R", [(2, 3), (6, 6)])
def test_export_hamiltonian_layout_and_round_trip(tmp_path, N, R):
    ch = Channels(N, R)
    vec = np.random.default_rng(N * 100 + R).standard_normal(ch.neq)
    p = str(tmp_path / f"ham_{N}_1.00_{R}")
    qd.export_hamiltonian_host(N, R, vec, p)
    raw = np.fromfile(p, dtype=np.dtype([("a", "<i4"), ("v", "<f8"), ("b", "<i4")]))
    assert raw.size == ch.neq and np.all(raw["a"] == 8) and np.all(raw["b"] == 8)
    order = reference_record_order(ch)
    assert sorted(order) == list(range(ch.neq))
    assert np.array_equal(raw["v"], vec[order])
    back = qd.import_hamiltonian_host(N, R, p)
    assert np.array_equal(back, vec)


def test_import_rejects_wrong_size(tmp_path):
    ch = Channels(2, 3)
    p = str(tmp_path / "ham")
    qd.export_hamiltonian_host(2, 3, np.zeros(ch.neq), p)
    with pytest.raises(qd.QDError, match="needs"):
        qd.import_hamiltonian_host(6, 6, p)
