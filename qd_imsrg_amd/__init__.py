"""MI355X IM-SRG(2) commutator / BCH engine -- Python host mirror of the C ABI.

Mirrors the reference's operator API (src/commutators.f90, src/IMSRG_tools.f90,
src/magnus_IMSRG.f90, src/gs_IMSRG.f90): ``Context`` is the channel structure,
``Operator`` a device-resident ``type(full_ham)``, ``CCMat`` a device-resident
``type(cc_mat)``; the module functions keep the reference's procedure names and
argument order.  Host vectors are the reference's ``vectorize`` vectors.
All compute runs in libqdimsrg.so on the GPU; nothing here computes physics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FlowParams, FlowResult, QDError, check

# QD_PREC_* of qdimsrg.h: arithmetic of the 2b-2b GEMMs
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x2": 2, "bf16x3": 3, "bf16x4": 4}

__all__ = ["Context", "Operator", "CCMat", "QDError", "commutator_110", "commutator_220", "commutator_111",
           "commutator_121", "commutator_221", "commutator_122", "commutator_222", "calc_cc", "build_white",
           "add_arrays", "copy_arrays", "set_to_zero", "mat_2_norm", "MBPT2", "commutator", "BCH_EXPAND",
           "MAGNUS_EXPAND", "euler_step", "dGam", "magnus_flow"]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_lib.dp)


class Context:
    """Channel structure of N electrons in R shells (reindex + build_cross_coupled)."""

    def __init__(self, nbody: int, nshell: int):
        self.L = _lib.load()
        h = C.c_void_p()
        check(self.L.qd_context_create(int(nbody), int(nshell), C.byref(h)))
        self.h = h
        info = _lib.Info()
        check(self.L.qd_context_info(self.h, C.byref(info)))
        self.nbody, self.nshell, self.msp = info.nbody, info.nshell, info.msp
        self.nblock, self.ncc, self.neq = info.nblock, info.ncc, info.neq
        self.cc_size, self.n2b = info.cc_size, info.n2b

    def table(self, which: int) -> np.ndarray:
        n = self.L.qd_context_table(self.h, which, None, 0)
        if n < 0:
            check(int(n))
        out = np.zeros(max(int(n), 1), dtype=np.int64)
        check(int(self.L.qd_context_table(self.h, which, out.ctypes.data_as(_lib.i64p), int(n))) - int(n))
        return out[:n]

    def sync(self):
        check(self.L.qd_context_sync(self.h))

    def operator(self, herm: int = 1, vec: np.ndarray | None = None) -> "Operator":
        op = Operator(self, herm)
        if vec is not None:
            op.upload(vec)
        return op

    def cc(self, herm: float = 1.0) -> "CCMat":
        return CCMat(self, herm)

    def timing(self, on: bool):
        check(self.L.qd_timing_enable(self.h, int(bool(on))))

    def timing_read(self):
        ms = np.zeros(7)
        n = np.zeros(7, dtype=np.int64)
        check(self.L.qd_timing_read(self.h, _dp(ms), n.ctypes.data_as(_lib.i64p), 7))
        return ms, n

    def gemm_flops(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(self.L.qd_gemm_flops(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(products=a.value, pandya=b.value, accumulate=c.value)

    def set_launch_mode(self, mode: int):
        """0: grouped persistent GEMM launch; 1: one launch per block problem."""
        check(self.L.qd_set_launch_mode(self.h, int(mode)))

    def set_graphs(self, on: bool):
        """Capture each BCH term in a hipGraph and replay it."""
        check(self.L.qd_set_graphs(self.h, int(bool(on))))

    def graph_count(self) -> int:
        n = int(self.L.qd_graph_count(self.h))
        if n < 0:
            check(n)
        return n

    def set_precision(self, prec):
        """Arithmetic of the 2b-2b GEMMs: "fp64" (default), "fp32", or
        "bf16x2" / "bf16x3" / "bf16x4" (BF16 mantissa slices); or the
        QD_PREC_* integer."""
        check(self.L.qd_set_precision(self.h, PRECISIONS[prec] if isinstance(prec, str) else int(prec)))

    def precision(self) -> str:
        v = C.c_int()
        check(self.L.qd_get_precision(self.h, C.byref(v)))
        return {i: k for k, i in PRECISIONS.items()}[v.value]

    def autotune(self, reps: int = 3):
        a, b = C.c_double(), C.c_double()
        check(self.L.qd_autotune(self.h, int(reps), C.byref(a), C.byref(b)))
        return a.value, b.value

    def gemm_shapes(self) -> np.ndarray:
        n = self.L.qd_gemm_shapes(self.h, None, 0)
        out = np.zeros(max(n, 1), dtype=np.int32)
        self.L.qd_gemm_shapes(self.h, out.ctypes.data_as(_lib.ip), n)
        return out[:n]

    def set_gemm_shapes(self, shapes):
        """Tile shape of every GEMM problem (as gemm_shapes reports them)."""
        a = np.ascontiguousarray(shapes, dtype=np.int32)
        check(self.L.qd_set_gemm_shapes(self.h, a.ctypes.data_as(_lib.ip), int(a.size)))

    def gemm_mfma_flops(self):
        """Matrix-core flops the GEMM launches of one commutator execute at the
        current shapes and precision (tile padding included)."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(self.L.qd_gemm_mfma_flops(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(products=a.value, pandya=b.value, accumulate=c.value)

    def close(self):
        if getattr(self, "h", None):
            self.L.qd_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Operator:
    """Device-resident full_ham (allocate_everything: zeroed)."""

    def __init__(self, ctx: Context, herm: int = 1):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.L.qd_op_create(ctx.h, int(herm), C.byref(h)))
        self.h = h

    @property
    def herm(self) -> int:
        x = C.c_int()
        check(self.ctx.L.qd_op_get_herm(self.h, C.byref(x)))
        return x.value

    @herm.setter
    def herm(self, v: int):
        check(self.ctx.L.qd_op_set_herm(self.h, int(v)))

    def upload(self, vec: np.ndarray):
        v = np.ascontiguousarray(vec, dtype=np.float64)
        check(self.ctx.L.qd_op_upload(self.h, _dp(v), v.size))

    def download(self) -> np.ndarray:
        v = np.empty(self.ctx.neq)
        check(self.ctx.L.qd_op_download(self.h, _dp(v), v.size))
        return v

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.L.qd_op_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CCMat:
    """Device-resident cc_mat (copy_cc: zeroed)."""

    def __init__(self, ctx: Context, herm: float = 1.0):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.L.qd_cc_create(ctx.h, float(herm), C.byref(h)))
        self.h = h

    def set_herm(self, herm: float):
        check(self.ctx.L.qd_cc_set_herm(self.h, float(herm)))

    def upload(self, x: np.ndarray):
        v = np.ascontiguousarray(x, dtype=np.float64)
        check(self.ctx.L.qd_cc_upload(self.h, _dp(v), v.size))

    def download(self) -> np.ndarray:
        v = np.empty(self.ctx.cc_size)
        check(self.ctx.L.qd_cc_download(self.h, _dp(v), v.size))
        return v

    def zero(self):
        check(self.ctx.L.qd_cc_zero(self.ctx.h, self.h))

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            self.ctx.L.qd_cc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- reference procedure names (src/commutators.f90) ----------------------
def _L(op):
    return op.ctx.L, op.ctx.h


def commutator_110(r1: Operator, r2: Operator) -> float:
    L, h = _L(r1)
    c = C.c_double()
    check(L.qd_commutator_110(h, r1.h, r2.h, C.byref(c)))
    return c.value


def commutator_220(r1: Operator, r2: Operator) -> float:
    L, h = _L(r1)
    c = C.c_double()
    check(L.qd_commutator_220(h, r1.h, r2.h, C.byref(c)))
    return c.value


def commutator_111(r1, r2, r3):
    L, h = _L(r1); check(L.qd_commutator_111(h, r1.h, r2.h, r3.h))


def commutator_121(r1, r2, r3):
    L, h = _L(r1); check(L.qd_commutator_121(h, r1.h, r2.h, r3.h))


def commutator_221(r1, r2, r3):
    L, h = _L(r1); check(L.qd_commutator_221(h, r1.h, r2.h, r3.h))


def commutator_122(r1, r2, r3):
    L, h = _L(r1); check(L.qd_commutator_122(h, r1.h, r2.h, r3.h))


def commutator_222(r1, r2, r3, HCC: CCMat | None, ETACC: CCMat | None):
    L, h = _L(r1)
    check(L.qd_commutator_222(h, r1.h, r2.h, r3.h, HCC.h if HCC else None, ETACC.h if ETACC else None))


def calc_cc(H: Operator, HCC: CCMat):
    L, h = _L(H); check(L.qd_calc_cc(h, H.h, HCC.h))


def build_white(gen: Operator, rec: Operator):
    L, h = _L(gen); check(L.qd_build_white(h, gen.h, rec.h))


def add_arrays(r1, s1, r2, s2, r3):
    L, h = _L(r1); check(L.qd_add_arrays(h, r1.h, float(s1), r2.h, float(s2), r3.h))


def copy_arrays(r1, r2):
    L, h = _L(r1); check(L.qd_copy_arrays(h, r1.h, r2.h))


def set_to_zero(r):
    L, h = _L(r); check(L.qd_set_to_zero(h, r.h))


def mat_2_norm(r) -> float:
    L, h = _L(r)
    x = C.c_double()
    check(L.qd_mat_2_norm(h, r.h, C.byref(x)))
    return x.value


def MBPT2(r) -> float:
    L, h = _L(r)
    x = C.c_double()
    check(L.qd_mbpt2(h, r.h, C.byref(x)))
    return x.value


def commutator(r1, r2, r3, HCC: CCMat | None = None, ETACC: CCMat | None = None, e0: bool = True):
    """Full [r1, r2] -> r3 in one pass; CC sets None = zero (Magnus as shipped).
    Returns the zero-body part (a device-to-host copy and wait), or None with
    e0=False, which leaves the call asynchronous."""
    L, h = _L(r1)
    v = C.c_double()
    check(L.qd_commutator(h, r1.h, r2.h, r3.h, HCC.h if HCC else None, ETACC.h if ETACC else None,
                          C.byref(v) if e0 else None))
    return v.value if e0 else None


def commutator_cc(r1, r2, r3, HCC: CCMat, ETACC: CCMat, e0: bool = True):
    """calc_cc(r2, HCC), calc_cc(r1, ETACC), then the full [r1, r2] -> r3 (the
    dGam sequence in one call; the CC fills overlap the ladder products).
    e0 as for commutator()."""
    L, h = _L(r1)
    v = C.c_double()
    check(L.qd_commutator_cc(h, r1.h, r2.h, r3.h, HCC.h, ETACC.h, C.byref(v) if e0 else None))
    return v.value if e0 else None


def BCH_EXPAND(HS, G, H, HCC=None, OCC=None) -> int:
    L, h = _L(HS)
    nt = C.c_int()
    check(L.qd_bch_expand(h, HS.h, G.h, H.h, HCC.h if HCC else None, OCC.h if OCC else None, C.byref(nt)))
    return nt.value


def MAGNUS_EXPAND(DG, G, ETA):
    L, h = _L(DG); check(L.qd_magnus_expand(h, DG.h, G.h, ETA.h))


def euler_step(G, DG, s: float, stp: float) -> float:
    L, h = _L(G)
    x = C.c_double(s)
    check(L.qd_euler_step(h, G.h, DG.h, C.byref(x), float(stp)))
    return x.value


def dGam(H, DER, HCC, ETACC):
    L, h = _L(H); check(L.qd_dgam(h, H.h, DER.h, HCC.h, ETACC.h))


def magnus_flow(ctx: Context, H0: Operator, HS: Operator, OMEGA: Operator, pandya: bool = False,
                warm_start: bool = False, verbose: int = 0, max_iter: int = 100000,
                conv: float | None = None, trace_len: int = 4096, diis: int = 0,
                stop_rule: str = "mbpt2", conv_deds: float = 1e-8, tda=(0, 0, 0), spectrum_path: str | None = None):
    """s_evolve (src/magnus_IMSRG.f90:194-248) on the GPU; returns (result, trace).
    verbose: 1 prints the reference's per-iteration line to stdout, 2 to stderr.
    diis: depth of the Pulay acceleration (0 = the reference's iteration).
    stop_rule: "mbpt2" (the reference: |MBPT2(H(s))| <= conv) or "deds"
    (|dE/ds| <= conv_deds, dE/ds = zero-body part of [eta(s), H(s)]).
    spectrum_path: write the reference's spectrum_magnus.dat there, with the
    TDA block tda = (M_L, M_S, cut)."""
    p = FlowParams()
    ctx.L.qd_flow_params_default(C.byref(p))
    p.pandya, p.warm_start, p.verbose, p.max_iter = int(pandya), int(warm_start), int(verbose), int(max_iter)
    p.diis = int(diis)
    p.stop_rule = {"mbpt2": 0, "deds": 1}[stop_rule]
    p.conv_deds = float(conv_deds)
    p.tda_ml, p.tda_ms, p.tda_cut = (int(x) for x in tda)
    path = None if spectrum_path is None else str(spectrum_path).encode()
    p.spectrum_path = path
    if conv is not None:
        p.conv = float(conv)
    res = FlowResult()
    trace = np.zeros((trace_len, 6))
    check(ctx.L.qd_magnus_flow(ctx.h, H0.h, HS.h, OMEGA.h, C.byref(p), C.byref(res), _dp(trace), trace_len))
    return res, trace[:res.trace_rows]


class ODE:
    """ode(dGam, ...) of src/adams_ode.f90:7 with the integrator's state on
    the device; ``integrate`` advances operator ``y`` from t to tout."""

    def __init__(self, ctx: Context):
        self.L = ctx.L
        h = C.c_void_p()
        check(self.L.qd_ode_create(ctx.h, C.byref(h)))
        self.h = h
        self.t, self.relerr, self.abserr, self.iflag = 0.0, 0.0, 0.0, 1

    def integrate(self, y: Operator, t: float, tout: float, relerr: float, abserr: float, iflag: int):
        tt, re, ab, fl = C.c_double(t), C.c_double(relerr), C.c_double(abserr), C.c_int(iflag)
        check(self.L.qd_ode_integrate(self.h, y.h, C.byref(tt), float(tout), C.byref(re), C.byref(ab), C.byref(fl)))
        return tt.value, re.value, ab.value, fl.value

    def stats(self):
        ns, nfe, k, h = C.c_int(), C.c_int64(), C.c_int(), C.c_double()
        check(self.L.qd_ode_stats(self.h, C.byref(ns), C.byref(nfe), C.byref(k), C.byref(h)))
        return dict(nsteps=ns.value, nfe=nfe.value, order=k.value, h=h.value)

    def close(self):
        if getattr(self, "h", None):
            self.L.qd_ode_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gs_flow(ctx: Context, H0: Operator, HS: Operator, verbose: int = 0, max_outputs: int = 100000,
            trace_len: int = 4096, **overrides):
    """gs_IMSRG's decoupling loop (src/gs_IMSRG.f90:100-120) on the GPU;
    returns (result, trace rows (output, s, E0, crit))."""
    p = _lib.GsParams()
    ctx.L.qd_gs_params_default(C.byref(p))
    p.verbose, p.max_outputs = int(verbose), int(max_outputs)
    for k, v in overrides.items():
        setattr(p, k, v)
    res = _lib.GsResult()
    trace = np.zeros((trace_len, 4))
    check(ctx.L.qd_gs_flow(ctx.h, H0.h, HS.h, C.byref(p), C.byref(res), _dp(trace), trace_len))
    return res, trace[:min(res.outputs, trace_len)]


# ---- input side (host) ------------------------------------------------------
def v_int(qi, qj, qk, ql, hw: float = 1.0) -> float:
    """<i j|V|l k> of the analytic oscillator Coulomb interaction (host)."""
    L = _lib.load()
    arr = [(C.c_int * 3)(*map(int, q)) for q in (qi, qj, qk, ql)]
    return L.qd_v_int(*[C.cast(a, _lib.ip) for a in arr], float(hw))


def _neq(nbody: int, nshell: int) -> int:
    L = _lib.load()
    n = L.qd_layout_table(int(nbody), int(nshell), 6, None, 0)
    if n < 0:
        check(int(n))
    tab = np.zeros(n, dtype=np.int64)
    L.qd_layout_table(int(nbody), int(nshell), 6, tab.ctypes.data_as(_lib.i64p), n)
    return int(tab[0])


def hf_hamiltonian(nbody: int, nshell: int, hw: float, me_bin: str | None = None):
    """H(0) in the HF basis (vectorize order) and the HF energies (host).
    me_bin: read the interaction from an ME.bin file (the reference's
    ../TBMEfiles/ME.bin) instead of evaluating v_int."""
    L = _lib.load()
    neq = _neq(nbody, nshell)
    M = nshell * (nshell + 1)
    h0, eig, it = np.zeros(neq), np.zeros(M), C.c_int()
    if me_bin is None:
        check(L.qd_hf_hamiltonian(int(nbody), int(nshell), float(hw), _dp(h0), neq, _dp(eig), C.byref(it)))
    else:
        check(L.qd_hf_hamiltonian_me(int(nbody), int(nshell), float(hw), str(me_bin).encode(), _dp(h0), neq,
                                     _dp(eig), C.byref(it)))
    return h0, eig, it.value


def hf_hamiltonian_device(ctx: "Context", hw: float, out: "Operator | None" = None):
    """H(0) in the HF basis built on the device (Fock builds and the two-body
    transform) into the hermitian operator out (a new one when None); returns
    (operator, HF energies, SCF iteration count)."""
    L = _lib.load()
    H = ctx.operator(1) if out is None else out
    M = ctx.nshell * (ctx.nshell + 1)
    eig, it = np.zeros(M), C.c_int()
    check(L.qd_hf_hamiltonian_device(ctx.h, float(hw), H.h, _dp(eig), C.byref(it)))
    return H, eig, it.value


def me_bin_write(nshell: int, path: str) -> int:
    """Write ME.bin (<ab|V|cd> at hw = 1 from v_int) for nshell shells;
    returns the record count."""
    L = _lib.load()
    n = C.c_int64()
    check(L.qd_me_bin_write(int(nshell), str(path).encode(), C.byref(n)))
    return n.value


# ---- on-disk formats (src/ME_general.f90:207-361) ---------------------------
def export_hamiltonian(op: Operator, path: str):
    """export_hamiltonian: the reference's unformatted per-scalar dump."""
    check(op.ctx.L.qd_export_hamiltonian(op.h, str(path).encode()))


def import_hamiltonian(op: Operator, path: str):
    check(op.ctx.L.qd_import_hamiltonian(op.h, str(path).encode()))


def export_hamiltonian_host(nbody: int, nshell: int, vec: np.ndarray, path: str):
    L = _lib.load()
    v = np.ascontiguousarray(vec, dtype=np.float64)
    check(L.qd_export_hamiltonian_host(int(nbody), int(nshell), _dp(v), v.size, str(path).encode()))


def import_hamiltonian_host(nbody: int, nshell: int, path: str) -> np.ndarray:
    L = _lib.load()
    v = np.zeros(_neq(nbody, nshell))
    check(L.qd_import_hamiltonian_host(int(nbody), int(nshell), _dp(v), v.size, str(path).encode()))
    return v


# ---- consumers of H(s): TDA spectrum and EOM-IM-SRG ---------------------------
def tda_spectrum(h: Operator, ml: int, ms: int, cut: int) -> np.ndarray:
    """Eigenvalues of the TDA block (src/operators.f90:7-124), relative to E0."""
    L, ctx = h.ctx.L, h.ctx
    cap = max(ctx.nbody * ctx.msp, 1)
    eig, d = np.zeros(cap), C.c_int()
    check(L.qd_tda_spectrum(ctx.h, h.h, int(ml), int(ms), int(cut), _dp(eig), cap, C.byref(d)))
    return eig[:d.value].copy()


def eom_dim(ctx: Context, ml: int, ms: int):
    a, b = C.c_int64(), C.c_int64()
    check(ctx.L.qd_eom_dim(ctx.h, int(ml), int(ms), C.byref(a), C.byref(b)))
    return a.value, b.value


def eom_matvec(h: Operator, ml: int, ms: int, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    check(h.ctx.L.qd_eom_matvec(h.ctx.h, h.h, int(ml), int(ms), _dp(x), _dp(y), x.size))
    return y


def calculate_excited_states(h: Operator, ml: int, ms: int, nev: int):
    """EOM-IM-SRG excitation energies dE (ascending), the ph weight of each
    state and the Lanczos steps (calculate_excited_states,
    src/EOM_IMSRG.f90:10-47)."""
    eig, w, it = np.zeros(nev), np.zeros(nev), C.c_int()
    check(h.ctx.L.qd_eom_excited_states(h.ctx.h, h.h, int(ml), int(ms), int(nev), _dp(eig), _dp(w), C.byref(it)))
    return eig, w, it.value


def eom13_dim(ctx: Context, mode: int, ml: int, ms: int):
    a, b = C.c_int64(), C.c_int64()
    check(ctx.L.qd_eom13_dim(ctx.h, int(mode), int(ml), int(ms), C.byref(a), C.byref(b)))
    return a.value, b.value


def eom13_matvec(h: Operator, mode: int, ml: int, ms: int, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    check(h.ctx.L.qd_eom13_matvec(h.ctx.h, h.h, int(mode), int(ml), int(ms), _dp(x), _dp(y), x.size))
    return y


def _attached_removed(h, mode, ml, ms, nev):
    eig, w, it = np.zeros(nev), np.zeros(nev), C.c_int()
    check(h.ctx.L.qd_eom_attached_removed(h.ctx.h, h.h, int(mode), int(ml), int(ms), int(nev), _dp(eig), _dp(w),
                                          C.byref(it)))
    return eig, w, it.value


def calculate_1p_attached(h: Operator, ml: int, ms: int, nev: int):
    """Particle-attached EOM energies dE, the 1p weight of each state and the
    Lanczos steps (calculate_1p_attached, src/EOM_IMSRG.f90:50-87)."""
    return _attached_removed(h, 1, ml, ms, nev)


def calculate_1h_removed(h: Operator, ml: int, ms: int, nev: int):
    """Hole-removed EOM energies (calculate_1h_removed, src/EOM_IMSRG.f90:92-128)."""
    return _attached_removed(h, 2, ml, ms, nev)
