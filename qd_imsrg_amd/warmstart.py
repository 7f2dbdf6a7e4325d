"""Learned warm start of the Magnus flow (host driver over the library's
device network and basis).

A sweep of converged Magnus Omegas over oscillator frequencies gives the
training set.  Omega lives in each frequency's own Hartree-Fock basis, whose
orbital signs the HF step fixes canonically (largest-magnitude component
positive), so Omegas of neighbouring frequencies are comparable.  A PCA basis
of the training Omegas (host, FP64: Gram matrix of the centred set) is held
on the device; an FP32 network trained by the library's own kernels maps the
features (omega, log omega), standardised, to the standardised PCA
coefficients.  For a new frequency the predicted coefficients are expanded on
the device into Omega, and qd_magnus_flow continues from it (warm_start = 1)
to the same decoupling criterion as a cold start.  A warm flow that fails or
does not converge falls back to the cold start.

Rank: the centred span of n training Omegas has rank <= n - 1, so a
requested rank above that is capped (64 requested with 48 training Omegas
gives 47); directions whose Gram eigenvalue is below 1e-12 of the largest are
dropped as well.  No counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import check


def _fp(a):
    return a.ctypes.data_as(_lib.fp)


class MLP:
    """FP32 network nin -> layers x hidden (tanh) -> nout on a context's stream."""

    def __init__(self, ctx, nin: int, hidden: int, layers: int, nout: int, seed: int = 0):
        self.L, self.ctx = ctx.L, ctx
        h = C.c_void_p()
        check(self.L.qd_mlp_create(ctx.h, nin, hidden, layers, nout, C.c_uint64(seed), C.byref(h)))
        self.h = h
        self.widths = [nin] + [hidden] * layers + [nout]
        self.n = int(self.L.qd_mlp_num_params(self.h))

    def params(self) -> np.ndarray:
        p = np.zeros(self.n, dtype=np.float32)
        check(self.L.qd_mlp_get_params(self.h, _fp(p), self.n))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(self.L.qd_mlp_set_params(self.h, _fp(p), p.size))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.widths[-1]), dtype=np.float32)
        check(self.L.qd_mlp_forward(self.h, _fp(x), x.shape[0], _fp(y)))
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss, g = C.c_float(), np.zeros(self.n, dtype=np.float32)
        check(self.L.qd_mlp_loss_grad(self.h, _fp(x), _fp(t), x.shape[0], C.byref(loss), _fp(g)))
        return loss.value, g

    def train(self, x: np.ndarray, t: np.ndarray, steps: int, lr: float) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        trace = np.zeros(max(steps, 1), dtype=np.float32)
        check(self.L.qd_mlp_train(self.h, _fp(x), _fp(t), x.shape[0], int(steps), C.c_float(lr), _fp(trace)))
        return trace[:steps]

    def close(self):
        if getattr(self, "h", None):
            self.L.qd_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Basis:
    """Mean and orthonormal PCA directions of operators, on the device."""

    def __init__(self, ctx, mean: np.ndarray, vecs: np.ndarray):
        self.L, self.ctx = ctx.L, ctx
        mean = np.ascontiguousarray(mean, dtype=np.float64)
        vecs = np.ascontiguousarray(vecs, dtype=np.float64)
        assert mean.size == ctx.neq and (vecs.size == 0 or vecs.shape[1] == ctx.neq)
        self.rank = int(vecs.shape[0])
        h = C.c_void_p()
        check(self.L.qd_basis_create(ctx.h, self.rank, mean.ctypes.data_as(_lib.dp),
                                     vecs.ctypes.data_as(_lib.dp), C.byref(h)))
        self.h = h

    def expand(self, coef: np.ndarray, out):
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        assert coef.size == self.rank
        check(self.L.qd_basis_expand(self.h, coef.ctypes.data_as(_lib.dp), out.h))

    def close(self):
        if getattr(self, "h", None):
            self.L.qd_basis_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def features(omega) -> np.ndarray:
    w = np.atleast_1d(np.asarray(omega, dtype=np.float64))
    return np.stack([w, np.log(w)], axis=1)


def pca(X: np.ndarray, rank: int):
    """Mean, orthonormal directions (r x neq) and coefficients (n x r) of the
    rows of X, from the Gram matrix of the centred rows (FP64)."""
    mean = X.mean(axis=0)
    Xc = X - mean
    G = Xc @ Xc.T
    lam, U = np.linalg.eigh(G)
    order = np.argsort(lam)[::-1]
    lam, U = lam[order], U[:, order]
    keep = lam > 1e-12 * max(lam[0], 1e-300)
    r = int(min(rank, X.shape[0] - 1, int(keep.sum())))
    lam, U = lam[:r], U[:, :r]
    vecs = (U.T @ Xc) / np.sqrt(lam)[:, None]
    coef = Xc @ vecs.T
    return mean, vecs, coef


class WarmStart:
    """PCA basis + trained network for one channel structure (context)."""

    def __init__(self, ctx, rank: int = 64, hidden: int = 256, layers: int = 3, seed: int = 0):
        self.ctx, self.rank_req, self.hidden, self.layers, self.seed = ctx, rank, hidden, layers, seed
        self.basis = self.mlp = None

    def fit(self, omegas, omega_vecs, steps: int = 20000, lr: float = 1e-3) -> dict:
        X = np.asarray(omega_vecs, dtype=np.float64)
        t0 = time.perf_counter()
        mean, vecs, coef = pca(X, self.rank_req)
        t_pca = time.perf_counter() - t0
        self.rank = vecs.shape[0]
        self.basis = Basis(self.ctx, mean, vecs)
        f = features(omegas)
        self.f_mu, self.f_sd = f.mean(0), f.std(0) + 1e-300
        self.c_mu, self.c_sd = coef.mean(0), coef.std(0) + 1e-300
        x = (f - self.f_mu) / self.f_sd
        t = (coef - self.c_mu) / self.c_sd
        self.mlp = MLP(self.ctx, 2, self.hidden, self.layers, self.rank, seed=self.seed)
        t0 = time.perf_counter()
        # Adam with a step-decayed rate: half the steps at lr, then a quarter
        # each at lr/3 and lr/10 (the moments carry over between phases)
        parts = [(steps - 2 * (steps // 4), lr), (steps // 4, lr / 3), (steps // 4, lr / 10)]
        trace = np.concatenate([self.mlp.train(x, t, n, r) for n, r in parts if n > 0] or [np.zeros(0)])
        t_train = time.perf_counter() - t0
        # how well the basis reproduces the training set and the fit its coefficients
        fit = self.coefficients(omegas)
        return {"rank": self.rank, "rank_requested": self.rank_req, "pca_s": t_pca, "train_s": t_train,
                "steps": steps, "loss_first": float(trace[0]) if steps else None,
                "loss_last": float(trace[-1]) if steps else None,
                "train_coef_rel_err": float(np.abs(fit - coef).max() / np.abs(coef).max())}

    def coefficients(self, omegas) -> np.ndarray:
        x = (features(omegas) - self.f_mu) / self.f_sd
        y = self.mlp.forward(x).astype(np.float64)
        return y * self.c_sd + self.c_mu

    def predict(self, omega, out):
        """Fill operator ``out`` with the predicted Omega for one frequency."""
        self.basis.expand(self.coefficients([omega])[0], out)

    def close(self):
        for o in (self.mlp, self.basis):
            if o is not None:
                o.close()


def warm_flow(ctx, ws: WarmStart, omega: float, H0, HS, OMEGA, diis: int = 6, max_iter: int = 15):
    """Magnus flow for one frequency from the network's predicted Omega (with
    the Pulay-accelerated continuation); the cold start with the same
    acceleration is the fallback when the warm flow fails or has not
    converged within max_iter steps.  From a good prediction the warm flow
    takes 6-11 steps; a poor one (a relative error near 1 in Omega) crawls
    through the plain map's slow tail before the acceleration starts (40
    steps seen), so the budget cuts it short.
    Returns (result, fallback, iterations spent in the warm attempt)."""
    from . import QDError, magnus_flow
    ws.predict(omega, OMEGA)
    spent = 0
    try:
        r, _ = magnus_flow(ctx, H0, HS, OMEGA, warm_start=True, diis=diis, max_iter=max_iter)
        spent = r.iterations
        if r.converged:
            return r, False, spent
    except QDError:
        spent = max_iter
    r, _ = magnus_flow(ctx, H0, HS, OMEGA, diis=diis)  # cold: Omega is reset to zero
    return r, True, spent
