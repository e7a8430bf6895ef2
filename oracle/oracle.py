"""ctypes wrapper of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the product, never in
its place.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBPATH = HERE / "_build" / "liboracle.so"

_V = C.c_void_p
_dp = C.POINTER(C.c_double)
_L = None


def build():
    subprocess.run(["make", "-C", str(HERE)], check=True)


def lib():
    global _L
    if _L is None:
        if not LIBPATH.exists():
            build()
        L = C.CDLL(str(LIBPATH))
        L.orc_mech_new.restype = _V
        L.orc_mech_new.argtypes = [_V]
        L.orc_mech_free.argtypes = [_V]
        L.orc_species_thermo.argtypes = [_V, C.c_double] + [_V] * 6
        L.orc_rates.argtypes = [_V, C.c_double, C.c_double, _V, _V, _V, _V, _V]
        L.orc_equilibrium.argtypes = [_V, C.c_double, _V, _V]
        L.orc_omega.argtypes = [_V, C.c_double, C.c_double, _V, _V]
        L.orc_dydt.argtypes = [_V, C.c_int, C.c_double, _V, _V]
        L.orc_jacobian.argtypes = [_V, C.c_int, C.c_double, _V, _V, _V, C.c_int]
        L.orc_lu_solve.restype = C.c_int
        L.orc_lu_solve.argtypes = [C.c_int, _V, _V]
        L.orc_chem_solve.restype = C.c_int
        L.orc_chem_solve.argtypes = [_V, _V, C.c_long, _V, _V, _V, _V, C.c_long, _V, C.c_double,
                                     _V, _V, C.c_long, C.c_int, _V, _V]
        L.orc_chem_solve_update.restype = C.c_int
        L.orc_chem_solve_update.argtypes = [_V, _V, C.c_long, _V, _V, _V, _V, C.c_long, _V,
                                            C.c_double, _V, C.c_int, _V, _V]
        L.orc_chem_source.argtypes = [_V, _V, C.c_long, _V, _V, _V, _V, C.c_long, _V, C.c_long,
                                      _V, C.c_long, C.c_int]
        L.orc_T_from_h.argtypes = [_V, _V, C.c_long, _V, _V, _V, C.c_long, _V, _V]
        L.orc_T_from_e.argtypes = [_V, _V, C.c_long, _V, _V, _V, C.c_long, _V, _V]
        L.orc_mix_thermo.argtypes = [_V, C.c_long, _V, _V, _V, C.c_long, _V, _V, _V, _V]
        L.orc_mix_props.argtypes = [_V, C.c_long, _V, _V, _V, C.c_long, _V, _V, _V, C.c_long,
                                    _V, _V, _V]
        L.orc_mts_timescales.argtypes = [_V, C.c_double, _V, _V]
        L.orc_chem_source_kim.argtypes = [_V, C.c_long, _V, _V, _V, _V, C.c_long, _V, C.c_long,
                                          _V, C.c_long, C.c_int]
        L.orc_mts_cell.restype = C.c_int
        L.orc_mts_cell.argtypes = [_V, _V, C.c_double, C.c_double, _V, _V, _V]
        _L = L
    return _L


def _p(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags.c_contiguous, "contiguous numpy array expected"
    return C.c_void_p(a.ctypes.data)


class Oracle:
    """Oracle bound to one mechanism (tables taken from a Mechanism)."""

    def __init__(self, mech):
        self.mech = mech
        t = mech.raw_tables()
        self._t = t
        self.ns = t.ns
        self.nr = t.nr
        self._h = lib().orc_mech_new(C.byref(t))
        if not self._h:
            raise ValueError("the oracle restates Arrhenius, third-body, fall-off and "
                             "chemically-activated rates only (no PLOG/CHEB)")

    def __del__(self):
        try:
            lib().orc_mech_free(self._h)
        except Exception:
            pass

    def species_thermo(self, T):
        out = [np.zeros(self.ns) for _ in range(6)]
        lib().orc_species_thermo(self._h, T, *[_p(a) for a in out])
        return dict(zip(["Cp", "dCp", "H", "G", "dG", "S"], out))

    def rates(self, p, T, c):
        out = [np.zeros(self.nr) for _ in range(4)]
        lib().orc_rates(self._h, p, T, _p(np.ascontiguousarray(c, dtype=float)), *[_p(a) for a in out])
        return dict(zip(["kf", "kr", "qf", "qr"], out))

    def equilibrium(self, T):
        kc, dkc = np.zeros(self.nr), np.zeros(self.nr)
        lib().orc_equilibrium(self._h, T, _p(kc), _p(dkc))
        return kc, dkc

    def omega(self, p, T, c):
        w = np.zeros(self.ns)
        lib().orc_omega(self._h, p, T, _p(np.ascontiguousarray(c, dtype=float)), _p(w))
        return w

    def dydt(self, y, pr, flow=0):
        f = np.zeros(self.ns + 1)
        lib().orc_dydt(self._h, flow, pr, _p(np.ascontiguousarray(y, dtype=float)), _p(f))
        return f

    def jacobian(self, y, pr, flow=0, exact_dkp=False):
        n = self.ns + 1
        f, J = np.zeros(n), np.zeros((n, n))
        lib().orc_jacobian(self._h, flow, pr, _p(np.ascontiguousarray(y, dtype=float)), _p(f),
                           _p(J), int(exact_dkp))
        return f, J

    def mts_timescales(self, T, c):
        tci = np.zeros(self.ns)
        lib().orc_mts_timescales(self._h, T, _p(np.ascontiguousarray(c, dtype=float)), _p(tci))
        return tci

    def mts_cell(self, opts, y, pr, tn):
        """One cell of the MTS explicit integrator: returns (y_new, groups, steps)."""
        y = np.ascontiguousarray(y, dtype=float).copy()
        nm, st = C.c_int(), C.c_int64()
        rc = lib().orc_mts_cell(self._h, C.byref(opts), pr, tn, _p(y), C.byref(nm), C.byref(st))
        if rc:
            raise RuntimeError(f"oracle MTS cell failed ({rc})")
        return y, nm.value, st.value

    @staticmethod
    def lu_solve(A, b):
        A = np.ascontiguousarray(A, dtype=float).copy()
        b = np.ascontiguousarray(b, dtype=float).copy()
        lib().orc_lu_solve(A.shape[0], _p(A), _p(b))
        return b

    def solve(self, opts, T, p, rho, Y, dt, dtFactor=1.0, subDt=None, nthreads=1,
              cell_steps=False):
        from openhurricane_amd._lib import Stats

        n = T.shape[0]
        subDt = np.zeros(n) if subDt is None else subDt.copy()
        R = np.zeros((self.ns, n))
        st = Stats()
        cs = np.zeros(n, dtype=np.int32) if cell_steps else None
        rc = lib().orc_chem_solve(self._h, C.byref(opts), n, _p(T), _p(p), _p(rho), _p(Y), n,
                                  _p(dt), dtFactor, _p(subDt), _p(R), n, nthreads, C.byref(st),
                                  _p(cs))
        if rc:
            raise RuntimeError(f"oracle integration failed ({rc})")
        return R, subDt, st.as_dict(), cs

    def solve_update(self, opts, T, p, rho, Y, dt, dtFactor=1.0, subDt=None, nthreads=1):
        from openhurricane_amd._lib import Stats

        n = T.shape[0]
        T = T.copy()
        Y = np.ascontiguousarray(Y).copy()
        subDt = np.zeros(n) if subDt is None else subDt.copy()
        st = Stats()
        rc = lib().orc_chem_solve_update(self._h, C.byref(opts), n, _p(T), _p(p), _p(rho), _p(Y),
                                         n, _p(dt), dtFactor, _p(subDt), nthreads, C.byref(st),
                                         None)
        if rc:
            raise RuntimeError(f"oracle integration failed ({rc})")
        return T, Y, subDt, st.as_dict()

    def chem_source(self, opts, T, p, rho, Y, diag=False, without_last=False):
        n = T.shape[0]
        R = np.zeros((self.ns, n))
        D = np.zeros((self.ns, n)) if diag else None
        lib().orc_chem_source(self._h, C.byref(opts), n, _p(T), _p(p), _p(rho), _p(Y), n, _p(R),
                              n, _p(D), n, int(without_last))
        return (R, D) if diag else R

    def chem_source_kim(self, T, p, rho, Y, fp32=False):
        """The GPU path's source terms R [ns, n] and Kim diagonal [ns-1, n]
        (float32 for the Hybrid variant)."""
        n = T.shape[0]
        R = np.zeros((self.ns, n))
        D = np.zeros((self.ns - 1, n), dtype=np.float32 if fp32 else np.float64)
        lib().orc_chem_source_kim(self._h, n, _p(T), _p(p), _p(rho), _p(Y), n, _p(R), n, _p(D), n,
                                  int(fp32))
        return R, D

    def T_from_h(self, opts, h, p, Y, T0):
        n = h.shape[0]
        T = T0.copy()
        fl = np.zeros(n, dtype=np.int32)
        lib().orc_T_from_h(self._h, C.byref(opts), n, _p(h), _p(p), _p(Y), n, _p(T), _p(fl))
        return T, fl

    def T_from_e(self, opts, e, p, Y, T0):
        n = e.shape[0]
        T = T0.copy()
        fl = np.zeros(n, dtype=np.int32)
        lib().orc_T_from_e(self._h, C.byref(opts), n, _p(e), _p(p), _p(Y), n, _p(T), _p(fl))
        return T, fl

    def mix_thermo(self, T, p, Y):
        n = T.shape[0]
        out = {k: np.zeros(n) for k in ["cp", "cv", "ha", "ea"]}
        lib().orc_mix_thermo(self._h, n, _p(T), _p(p), _p(Y), n, *[_p(out[k]) for k in
                                                                  ["cp", "cv", "ha", "ea"]])
        return out

    def mix_props(self, T, p, Y, hk=False):
        n = T.shape[0]
        out = {k: np.zeros(n) for k in ["mu", "kappa", "cp", "ha"]}
        out["Dk"] = np.zeros((self.ns, n))
        H = np.zeros((self.ns, n)) if hk else None
        lib().orc_mix_props(self._h, n, _p(T), _p(p), _p(Y), n, _p(out["mu"]), _p(out["kappa"]),
                            _p(out["Dk"]), n, _p(out["cp"]), _p(out["ha"]), _p(H))
        if hk:
            out["hk"] = H
        return out
