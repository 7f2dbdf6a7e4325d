"""ctypes front end of the C oracle (oracle/imsrg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's cpu_baseline leg, always as the checker / CPU baseline, never as a
product code path.  Build with ``make -C oracle`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .channels import Channels

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
i64p = C.POINTER(C.c_int64)
dp = C.POINTER(C.c_double)


class _Tab(C.Structure):
    _fields_ = [("n", C.c_int64), ("M", C.c_int64), ("mp", C.c_int64), ("nblock", C.c_int64),
                ("neq", C.c_int64),
                ("nh", i64p), ("np", i64p), ("nb", i64p), ("off", i64p),
                ("qnhh", i64p), ("qnpp", i64p), ("qnph", i64p),
                ("hh_start", i64p), ("pp_start", i64p), ("ph_start", i64p),
                ("iarr", i64p),
                ("ncc", C.c_int64), ("cc_size", C.c_int64),
                ("cc_r", i64p), ("cc_nb", i64p), ("cc_off", i64p), ("cc_ab", i64p), ("cc_hp", i64p),
                ("cc_ab_start", i64p), ("cc_hp_start", i64p), ("cc_map", i64p),
                ("ml", i64p), ("ms", i64p),
                ("off_fhh", C.c_int64), ("off_fpp", C.c_int64), ("off_fph", C.c_int64)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        T = C.POINTER(_Tab)
        sig = {
            "o_velem": (C.c_double, [T, dp, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int64]),
            "o_felem": (C.c_double, [T, dp, C.c_int, C.c_int64, C.c_int64]),
            "o_comm110": (C.c_double, [T, dp, C.c_int, dp, C.c_int]),
            "o_comm220": (C.c_double, [T, dp, C.c_int, dp, C.c_int]),
            "o_comm111": (None, [T, dp, C.c_int, dp, C.c_int, dp]),
            "o_comm121": (None, [T, dp, C.c_int, dp, C.c_int, dp]),
            "o_comm221": (None, [T, dp, C.c_int, dp, C.c_int, dp]),
            "o_comm122": (None, [T, dp, C.c_int, dp, C.c_int, dp]),
            "o_comm222": (None, [T, dp, C.c_int, dp, C.c_int, dp, dp, C.c_double, dp, C.c_double]),
            "o_calc_cc": (None, [T, dp, C.c_int, dp]),
            "o_build_white": (None, [T, dp, dp, C.c_int]),
            "o_mbpt2": (C.c_double, [T, dp]),
            "o_norm": (C.c_double, [T, dp]),
            "o_add": (None, [T, dp, C.c_double, dp, C.c_double, dp]),
            "o_commutator": (C.c_double, [T, dp, C.c_int, dp, C.c_int, dp, dp, C.c_double, dp, C.c_double]),
            "o_bch_expand": (C.c_int, [T, dp, dp, C.c_int, dp, dp, C.c_double, dp, C.c_double, dp, dp, dp]),
            "o_eom_matvec": (C.c_int64, [T, dp, C.c_int, C.c_int, C.c_int, dp, dp]),
            "o_eom13_matvec": (C.c_int64, [T, dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(dp)


class Oracle:
    """Binds one channel structure; every method takes/returns flat numpy vectors."""

    def __init__(self, ch: Channels):
        self.ch = ch
        self._arrs = {k: np.ascontiguousarray(v, dtype=np.int64) for k, v in ch.to_ctypes_arrays().items()}
        a = self._arrs
        ptr = lambda k: a[k].ctypes.data_as(i64p)
        self.tab = _Tab(ch.n, ch.M, ch.mp, ch.nblock, ch.neq,
                        ptr("nh"), ptr("np"), ptr("nb"), ptr("off"),
                        ptr("qnhh"), ptr("qnpp"), ptr("qnph"),
                        ptr("hh_start"), ptr("pp_start"), ptr("ph_start"), ptr("iarr"),
                        ch.ncc, ch.cc_size,
                        ptr("cc_r"), ptr("cc_nb"), ptr("cc_off"), ptr("cc_ab"), ptr("cc_hp"),
                        ptr("cc_ab_start"), ptr("cc_hp_start"), ptr("cc_map"),
                        ptr("ml"), ptr("ms"), ch.off_fhh, ch.off_fpp, ch.off_fph)
        self.L = lib()
        self.t = C.byref(self.tab)

    def zeros(self):
        return np.zeros(self.ch.neq)

    def cc_zeros(self):
        return np.zeros(max(self.ch.cc_size, 1))

    # -- elementwise accessors ----------------------------------------
    def velem(self, v, h, a, b, c, d):
        return self.L.o_velem(self.t, _dp(v), h, a, b, c, d)

    def felem(self, v, h, a, b):
        return self.L.o_felem(self.t, _dp(v), h, a, b)

    # -- reference entry points ---------------------------------------
    def commutator_110(self, v1, h1, v2, h2):
        return self.L.o_comm110(self.t, _dp(v1), h1, _dp(v2), h2)

    def commutator_220(self, v1, h1, v2, h2):
        return self.L.o_comm220(self.t, _dp(v1), h1, _dp(v2), h2)

    def commutator_111(self, v1, h1, v2, h2, v3):
        self.L.o_comm111(self.t, _dp(v1), h1, _dp(v2), h2, _dp(v3))

    def commutator_121(self, v1, h1, v2, h2, v3):
        self.L.o_comm121(self.t, _dp(v1), h1, _dp(v2), h2, _dp(v3))

    def commutator_221(self, v1, h1, v2, h2, v3):
        self.L.o_comm221(self.t, _dp(v1), h1, _dp(v2), h2, _dp(v3))

    def commutator_122(self, v1, h1, v2, h2, v3):
        self.L.o_comm122(self.t, _dp(v1), h1, _dp(v2), h2, _dp(v3))

    def commutator_222(self, v1, h1, v2, h2, v3, hcc, hh, ecc, eh):
        self.L.o_comm222(self.t, _dp(v1), h1, _dp(v2), h2, _dp(v3), _dp(hcc), hh, _dp(ecc), eh)

    def calc_cc(self, v, h):
        cc = self.cc_zeros()
        self.L.o_calc_cc(self.t, _dp(v), h, _dp(cc))
        return cc

    def build_white(self, rec, hrec=1, gen=None):
        gen = self.zeros() if gen is None else gen
        self.L.o_build_white(self.t, _dp(gen), _dp(rec), hrec)
        return gen

    def mbpt2(self, v):
        return self.L.o_mbpt2(self.t, _dp(v))

    def norm(self, v):
        return self.L.o_norm(self.t, _dp(v))

    def add(self, r1, s1, r2, s2):
        out = self.zeros()
        self.L.o_add(self.t, _dp(r1), s1, _dp(r2), s2, _dp(out))
        return out

    def commutator(self, v1, h1, v2, h2, pandya: bool = True):
        """Full [r1, r2]; the CC inputs are calc_cc of the operands when
        ``pandya`` (dGam's usage) and zero otherwise (the Magnus driver's)."""
        if pandya:
            ecc, hcc = self.calc_cc(v1, h1), self.calc_cc(v2, h2)
        else:
            ecc = hcc = self.cc_zeros()
        out = self.zeros()
        self.L.o_commutator(self.t, _dp(v1), h1, _dp(v2), h2, _dp(out),
                            _dp(hcc), float(h2), _dp(ecc), float(h1))
        return out

    def eom_dim(self, ml, ms):
        """Length of the EOM vector of transfer (ml, ms): ph + pphh components."""
        return int(self.L.o_eom_matvec(self.t, None, 1, int(ml), int(ms), None, None))

    def eom_matvec(self, h, hh, ml, ms, x):
        """matvec_prod (src/EOM_IMSRG.f90:551-660): [H, Q] on the ph and pphh
        components of the ladder operator Q wrapped in x."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        n = self.L.o_eom_matvec(self.t, _dp(h), int(hh), int(ml), int(ms), _dp(x), _dp(y))
        assert n == x.size, (n, x.size)
        return y

    def eom13_dim(self, mode, ml, ms):
        """Length of the particle-attached (mode 1) or hole-removed (mode 2)
        EOM vector of (ml, ms)."""
        return int(self.L.o_eom13_matvec(self.t, None, 1, int(mode), int(ml), int(ms), None, None))

    def eom13_matvec(self, h, hh, mode, ml, ms, x):
        """matvec_1p_attached (mode 1) / matvec_1h_removed (mode 2),
        src/EOM_IMSRG.f90:662-812."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        n = self.L.o_eom13_matvec(self.t, _dp(h), int(hh), int(mode), int(ml), int(ms), _dp(x), _dp(y))
        assert n == x.size, (n, x.size)
        return y

    def bch_expand(self, g, hg, h, hcc=None, hh=1.0, occ=None, oh=-1.0):
        hcc = self.cc_zeros() if hcc is None else hcc
        occ = self.cc_zeros() if occ is None else occ
        hs, s1, s2, s3 = self.zeros(), self.zeros(), self.zeros(), self.zeros()
        nt = self.L.o_bch_expand(self.t, _dp(hs), _dp(g), hg, _dp(h), _dp(hcc), hh, _dp(occ), oh,
                                 _dp(s1), _dp(s2), _dp(s3))
        return hs, nt
