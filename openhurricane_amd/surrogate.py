"""Chemistry surrogate: an MLP advancing a cell over one flow time step,
trained on the plain integrator's own state pairs, with a gate that sends
every cell it cannot vouch for back to the integrator.

All arrays are torch tensors on the context's device (the surrogate entry
points of libohc.so take device memory only).  The device kernels do the
work: feature extraction, BF16 MFMA forward/backward GEMMs, Adam, and the
fused inference + gate + state update; this module only sequences them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Stats, check, lib


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Surrogate:
    def __init__(self, cs, width=256, depth=3, seed=0):
        if not cs.device_pointers:
            raise ValueError("the surrogate needs a ChemistrySource with device_pointers=1")
        self.cs = cs
        self.width, self.depth = width, depth
        self.ns = cs.mech.tables()["ns"]
        self.nin, self.nout = self.ns + 2, self.ns + 1
        h = C.c_void_p()
        check(lib().ohc_surrogate_create(cs._h, width, depth, seed, C.byref(h)), cs._h)
        self._h = h
        n = C.c_int64()
        check(lib().ohc_surrogate_num_params(self._h, C.byref(n)))
        self.num_params = n.value
        self.dt = None
        self.delta = None
        self._peer = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ohc_surrogate_free(h)
            self._h = None

    # ---- parameters ----
    def get_params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        check(lib().ohc_surrogate_get_params(self._h, p.ctypes.data_as(C.c_void_p)), self.cs._h)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        check(lib().ohc_surrogate_set_params(self._h, p.ctypes.data_as(C.c_void_p)), self.cs._h)

    def set_norm(self, in_mean, in_std, out_mean, out_std, dt, in_lo=None, in_hi=None,
                 out_weight=None):
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
        arrs = [f(a) for a in (in_mean, in_std, in_lo, in_hi, out_mean, out_std, out_weight)]
        self._norm_keep = arrs
        ptrs = [a.ctypes.data_as(C.c_void_p) if a is not None else None for a in arrs]
        check(lib().ohc_surrogate_set_norm(self._h, *ptrs, float(dt)), self.cs._h)
        self.norm = tuple(arrs)
        self.dt = dt

    def set_target_scale(self, delta):
        """Train / predict output k as sign(d) ln(1 + |d| / delta_k) of the
        increment d (None: plain increments).  Set before features()."""
        if delta is None:
            check(lib().ohc_surrogate_set_target_scale(self._h, None), self.cs._h)
            self.delta = None
            return
        d = np.ascontiguousarray(delta, dtype=np.float32)
        assert d.size == self.nout
        check(lib().ohc_surrogate_set_target_scale(self._h, d.ctypes.data_as(C.c_void_p)), self.cs._h)
        self.delta = d

    def set_precision(self, precision):
        """Inference precision: 0 BF16, 1 FP8 hidden-to-hidden layers."""
        check(lib().ohc_surrogate_set_precision(self._h, int(precision)), self.cs._h)

    def set_peer(self, peer, tol):
        """Gate on agreement with a second network (None: no peer)."""
        check(lib().ohc_surrogate_set_peer(self._h, peer._h if peer is not None else None,
                                           float(tol)), self.cs._h)
        self._peer = peer

    def save(self, path):
        check(lib().ohc_surrogate_save(self._h, str(path).encode()), self.cs._h)

    def load(self, path):
        check(lib().ohc_surrogate_load(self._h, str(path).encode()), self.cs._h)

    # ---- data ----
    def features(self, T0, p, Y0, T1=None, Y1=None):
        import torch

        n = T0.shape[0]
        X = torch.empty((n, self.nin), dtype=torch.float32, device=T0.device)
        Yt = torch.empty((n, self.nout), dtype=torch.float32, device=T0.device) if T1 is not None else None
        check(lib().ohc_surrogate_features(self._h, n, _ptr(T0), _ptr(p), _ptr(Y0), Y0.stride(0),
                                           _ptr(T1), _ptr(Y1), _ptr(X), _ptr(Yt)), self.cs._h)
        return X, Yt

    # ---- training ----
    def train_step(self, X, Yt, lr, want_loss=False):
        loss = C.c_float()
        check(lib().ohc_surrogate_train_step(self._h, X.shape[0], _ptr(X), _ptr(Yt), float(lr),
                                             C.byref(loss) if want_loss else None), self.cs._h)
        return loss.value if want_loss else None

    def gradient(self, X, Yt):
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = C.c_float()
        check(lib().ohc_surrogate_gradient(self._h, X.shape[0], _ptr(X), _ptr(Yt),
                                           g.ctypes.data_as(C.c_void_p), C.byref(loss)), self.cs._h)
        return g, loss.value

    def fit_norm(self, X, Yt, dt, margin=0.05, out_weight=None):
        """Feature / target statistics of a training set; the accepted feature
        box is the observed range widened by `margin` of its span."""
        mean, std = X.mean(0), X.std(0).clamp_min(1e-6)
        lo, hi = X.min(0).values, X.max(0).values
        span = (hi - lo).clamp_min(1e-6)
        om, osd = Yt.mean(0), Yt.std(0).clamp_min(1e-8)
        c = lambda t: t.double().cpu().numpy()
        self.set_norm(c(mean), c(std), c(om), c(osd), dt, in_lo=c(lo - margin * span),
                      in_hi=c(hi + margin * span), out_weight=out_weight)

    def fit(self, X, Yt, epochs=10, batch=16384, lr=2e-3, lr_end=2e-5, seed=0, log=None):
        """Adam over shuffled mini-batches with a cosine learning-rate decay."""
        import torch

        n = X.shape[0]
        g = torch.Generator(device=X.device)
        g.manual_seed(seed)
        steps = epochs * (n // batch)
        k, hist = 0, []
        for ep in range(epochs):
            perm = torch.randperm(n, device=X.device, generator=g)
            last = None
            for b in range(n // batch):
                idx = perm[b * batch:(b + 1) * batch]
                xb, yb = X[idx].contiguous(), Yt[idx].contiguous()
                a = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * k / max(steps, 1)))
                want = b == n // batch - 1
                last = self.train_step(xb, yb, a, want_loss=want)
                k += 1
            hist.append(last)
            if log:
                log(f"epoch {ep}: loss {last:.4g}")
        return hist

    # ---- inference ----
    def predict(self, T, p, Y):
        import torch

        n = T.shape[0]
        out = torch.empty((n, self.nout), dtype=torch.float32, device=T.device)
        check(lib().ohc_surrogate_predict(self._h, n, _ptr(T), _ptr(p), _ptr(Y), Y.stride(0),
                                          _ptr(out)), self.cs._h)
        return out

    def solve_update(self, T, p, rho, Y, dt, dtFactor=1.0, subDt=None, gate_tol=1e-3):
        """chemistrySource::solveUpdate through the surrogate with fallback;
        returns (cells that fell back, integrator counters of those cells)."""
        import torch

        n = T.shape[0]
        if subDt is None:
            subDt = torch.zeros(n, dtype=torch.float64, device=T.device)
        nf = C.c_int64()
        st = Stats()
        check(lib().ohc_chem_solve_update_surrogate(self.cs._h, self._h, n, _ptr(T), _ptr(p),
                                                    _ptr(rho), _ptr(Y), Y.stride(0), _ptr(dt),
                                                    float(dtFactor), _ptr(subDt), float(gate_tol),
                                                    C.byref(nf), C.byref(st)), self.cs._h)
        return nf.value, st.as_dict()
