"""Python mirror of the reference chemistry call surface over libohc.so.

``ChemistrySource`` mirrors OpenHurricane's ``chemistrySource`` entry points
(src/speciesTransport/chemistrySource/include/chemistrySource.hpp:576-662):
``solve(dt, dtFactor)`` returns the source terms R and the minimum suggested
sub-step, ``solve_update`` advances Y and T in place, and the thermo /
transport calls mirror ``thermoList::THa_p`` and ``transportList::muKappa`` /
``Df``.  Arrays are numpy arrays (host pointers) or, for a context created
with ``device_pointers=True``, torch CUDA tensors (device pointers, calls are
asynchronous on the context stream).  Species fields are (ns, n) blocks,
species k of cell i at Y[k, i].
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import MechTables, Stats, check, lib

DATA = Path(__file__).resolve().parent / "data"


def _as_np(ptr, n, dtype=np.float64):
    if not ptr or n == 0:
        return np.zeros(n, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Mechanism:
    """A loaded reaction mechanism (species, NASA-7 thermo, reactions, transport)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if isinstance(handle, int) else handle

    # -- constructors -------------------------------------------------
    @classmethod
    def load(cls, chemkin, thermo=None, transport=None):
        h = C.c_void_p()
        enc = lambda s: None if s is None else str(s).encode()
        check(lib().ohc_mech_load(enc(chemkin), enc(thermo), enc(transport), C.byref(h)))
        return cls(h)

    @classmethod
    def h2air(cls, transport=True):
        """The 9-species / 19-reaction H2-air mechanism shipped with the package."""
        d = DATA / "h2air"
        return cls.load(d / "mech.che", d / "thermo.dat", d / "transport.dat" if transport else None)

    @classmethod
    def synthetic(cls, ns, nr, seed, frac_rev=0.2, frac_tb=0.1):
        h = C.c_void_p()
        check(lib().ohc_mech_synthetic(ns, nr, seed, frac_rev, frac_tb, C.byref(h)))
        return cls(h)

    @classmethod
    def from_tables(cls, tab: dict):
        keep = []

        def dptr(name):
            a = tab.get(name)
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_double))

        def iptr(name):
            a = np.ascontiguousarray(tab[name], dtype=np.int32)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_int32))

        t = MechTables()
        t.ns, t.nr, t.ne = int(tab["ns"]), int(tab["nr"]), int(tab.get("ne", 0))
        for f in ["W", "Tlow", "Thigh", "Tcommon", "nasa_lo", "nasa_hi", "elem", "ekb", "sigma",
                  "f_nu", "b_nu", "f_ord", "b_ord", "arr", "arr_lo", "arr_rev", "fo", "eff"]:
            setattr(t, f, dptr(f))
        for f in ["nf", "nb", "f_idx", "b_idx", "rev", "kind", "fo_type"]:
            setattr(t, f, iptr(f))
        if tab.get("aux_off") is not None:  # PLOG / CHEB data
            t.aux_off, t.aux = iptr("aux_off"), dptr("aux")
        h = C.c_void_p()
        check(lib().ohc_mech_from_tables(C.byref(t), C.byref(h)))
        return cls(h)

    def __del__(self):
        try:
            if self._h:
                lib().ohc_mech_free(self._h)
                self._h = None
        except Exception:
            pass

    # -- accessors ------------------------------------------------------
    def raw_tables(self) -> MechTables:
        t = MechTables()
        check(lib().ohc_mech_get_tables(self._h, C.byref(t)))
        return t

    def tables(self) -> dict:
        t = self.raw_tables()
        ns, nr, ne, P = t.ns, t.nr, t.ne, _lib.OHC_MAXP
        d = {"ns": ns, "nr": nr, "ne": ne}
        for f, n in [("W", ns), ("Tlow", ns), ("Thigh", ns), ("Tcommon", ns), ("ekb", ns),
                     ("sigma", ns)]:
            d[f] = _as_np(getattr(t, f), n) if getattr(t, f) else None
        d["nasa_lo"] = _as_np(t.nasa_lo, 7 * ns).reshape(ns, 7)
        d["nasa_hi"] = _as_np(t.nasa_hi, 7 * ns).reshape(ns, 7)
        d["elem"] = _as_np(t.elem, ns * ne).reshape(ns, ne)
        for f in ["nf", "nb", "rev", "kind", "fo_type"]:
            d[f] = _as_np(getattr(t, f), nr, np.int32)
        for f in ["f_idx", "b_idx"]:
            d[f] = _as_np(getattr(t, f), nr * P, np.int32).reshape(nr, P)
        for f in ["f_nu", "b_nu", "f_ord", "b_ord"]:
            d[f] = _as_np(getattr(t, f), nr * P).reshape(nr, P)
        for f in ["arr", "arr_lo", "arr_rev"]:
            d[f] = _as_np(getattr(t, f), nr * 3).reshape(nr, 3)
        d["fo"] = _as_np(t.fo, nr * 5).reshape(nr, 5)
        d["eff"] = _as_np(t.eff, nr * ns).reshape(nr, ns)
        d["aux_off"] = _as_np(t.aux_off, nr + 1, np.int32)
        d["aux"] = _as_np(t.aux, int(d["aux_off"][-1])) if d["aux_off"][-1] else np.zeros(0)
        return d

    def rate_data(self, r: int) -> dict:
        """The PLOG pressure points or the CHEB grid of reaction r, in SI units
        (p in Pa, A in kmol-m^3-s, Ta in K); empty for the other rate laws."""
        d = self.tables()
        a = d["aux"][d["aux_off"][r]:d["aux_off"][r + 1]]
        if d["kind"][r] == 4:
            return {"law": "PLOG", "points": a.reshape(-1, 4)}
        if d["kind"][r] == 5:
            n, m = int(a[0]), int(a[1])
            return {"law": "CHEB", "N": n, "M": m, "T": (a[2], a[3]), "P": (a[4], a[5]),
                    "coef": a[6:].reshape(n, m)}
        return {}

    @property
    def ns(self):
        return self.raw_tables().ns

    @property
    def nr(self):
        return self.raw_tables().nr

    @property
    def species(self):
        out = []
        for k in range(self.ns):
            s = C.c_char_p()
            check(lib().ohc_mech_species_name(self._h, k, C.byref(s)))
            out.append(s.value.decode())
        return out

    @property
    def hash(self):
        h = C.c_uint64()
        check(lib().ohc_mech_hash(self._h, C.byref(h)))
        return h.value


def _ptr(a, device, dtype="float64"):
    """Raw pointer of a numpy array (host) or torch tensor (device)."""
    if a is None:
        return None
    if device:
        import torch

        if not isinstance(a, torch.Tensor) or not a.is_cuda:
            raise TypeError("device-pointer context expects torch CUDA tensors")
        want = torch.float64 if dtype == "float64" else torch.int32
        if a.dtype != want or not a.is_contiguous():
            raise TypeError(f"expected contiguous {want} tensor")
        return C.c_void_p(a.data_ptr())
    if not isinstance(a, np.ndarray) or a.dtype != np.dtype(dtype) or not a.flags.c_contiguous:
        raise TypeError(f"expected C-contiguous numpy {dtype} array")
    return C.c_void_p(a.ctypes.data)


def _ld(a, n):
    """Leading dimension (row stride in elements) of an (ns, n) block."""
    if a is None:
        return n
    if a.ndim == 1:
        return n
    st = a.stride(0) if hasattr(a, "stride") and callable(a.stride) else a.strides[0] // a.itemsize
    return int(st)


class ChemistrySource:
    """chemistrySource / thermoList / transportList calls on one GPU context."""

    def __init__(self, mech: Mechanism, device=0, opts=None, **optkw):
        self.mech = mech
        self.opts = opts if opts is not None else _lib.default_opts(**optkw)
        self.device_pointers = bool(self.opts.device_pointers)
        self.ns = mech.ns
        self.device = device
        h = C.c_void_p()
        check(lib().ohc_ctx_create(mech._h, C.byref(self.opts), device, C.byref(h)))
        self._h = h
        if self.device_pointers:
            import torch

            # torch produces and consumes the tensors on its current stream;
            # the library's kernels run on that same stream so that both
            # stay ordered without extra synchronisation
            self.set_stream(torch.cuda.current_stream(device).cuda_stream)

    def __del__(self):
        try:
            if self._h:
                lib().ohc_ctx_free(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def kernel_name(self):
        return lib().ohc_ctx_kernel_name(self._h).decode()

    def set_stream(self, stream_ptr: int):
        check(lib().ohc_ctx_set_stream(self._h, C.c_void_p(stream_ptr)), self._h)

    def autotune(self, T, p, rho, Y, dt, dtFactor=1.0):
        """Pick the persistent-grid size (wavefronts per CU) by timing two
        flow steps on these cells; returns (waves_per_cu, timed step in ms).
        The arrays are not modified."""
        n = T.shape[0]
        w = C.c_int32()
        ms = C.c_double()
        check(lib().ohc_ctx_autotune(self._h, n, self._p(T), self._p(p), self._p(rho), self._p(Y),
                                     _ld(Y, n), self._p(dt), float(dtFactor), C.byref(w),
                                     C.byref(ms)), self._h)
        return w.value, ms.value

    def sync(self):
        check(lib().ohc_sync(self._h), self._h)

    def _p(self, a, dtype="float64"):
        return _ptr(a, self.device_pointers, dtype)

    # chemistrySource::solve(dt, dtFactor)
    def solve(self, T, p, rho, Y, dt, dtFactor=1.0, subDt=None, R=None, stats=True):
        n = T.shape[0]
        if subDt is None:
            subDt = self._zeros(n)
        if R is None:
            R = self._zeros((self.ns, n))
        st = Stats() if stats else None
        rc = lib().ohc_chem_solve(self._h, n, self._p(T), self._p(p), self._p(rho), self._p(Y),
                                  _ld(Y, n), self._p(dt), float(dtFactor), self._p(subDt),
                                  self._p(R), _ld(R, n), C.byref(st) if st else None)
        check(rc, self._h)
        return R, subDt, (st.as_dict() if st else None)

    # chemistrySource::solveUpdate(dt, dtFactor): T and Y advanced in place
    def solve_update(self, T, p, rho, Y, dt, dtFactor=1.0, subDt=None, stats=True):
        n = T.shape[0]
        if subDt is None:
            subDt = self._zeros(n)
        st = Stats() if stats else None
        rc = lib().ohc_chem_solve_update(self._h, n, self._p(T), self._p(p), self._p(rho),
                                         self._p(Y), _ld(Y, n), self._p(dt), float(dtFactor),
                                         self._p(subDt), C.byref(st) if st else None)
        check(rc, self._h)
        return subDt, (st.as_dict() if st else None)

    # solveUpdate without any device allocation (ohc_chem_solve_update_noalloc):
    # device-pointer contexts only, n <= OHC_NOALLOC_MAX_CELLS
    def solve_update_noalloc(self, T, p, rho, Y, dt, dtFactor=1.0, subDt=None, stats=True):
        n = T.shape[0]
        if subDt is None:
            subDt = self._zeros(n)
        st = Stats() if stats else None
        rc = lib().ohc_chem_solve_update_noalloc(self._h, n, self._p(T), self._p(p),
                                                 self._p(rho), self._p(Y), _ld(Y, n),
                                                 self._p(dt), float(dtFactor), self._p(subDt),
                                                 C.byref(st) if st else None)
        check(rc, self._h)
        return subDt, (st.as_dict() if st else None)

    # chemistrySource::DyDt / jacobian at states y (N, n)
    def eval_rhs(self, y, p):
        n = y.shape[1]
        f = self._zeros(y.shape)
        check(lib().ohc_eval_rhs(self._h, n, self._p(y), _ld(y, n), self._p(p), self._p(f),
                                 _ld(f, n)), self._h)
        return f

    def eval_jacobian(self, y, p):
        N, n = y.shape
        f = self._zeros(y.shape)
        J = self._zeros((N * N, n))
        check(lib().ohc_eval_jacobian(self._h, n, self._p(y), _ld(y, n), self._p(p), self._p(f),
                                      _ld(f, n), self._p(J), _ld(J, n)), self._h)
        return f, J

    # SquareMatrix::LUDecomposeDPivoting + LUBacksubstituteDPivoting
    def lu_solve(self, A, b):
        """x with A_m x_m = b_m for a batch of n x n systems: A is
        (n * n, nmat) with entry (i, j) of system m at A[i * n + j, m] (the
        eval_jacobian layout), b is (n, nmat)."""
        n, nmat = b.shape
        if A.shape != (n * n, nmat):
            raise ValueError(f"A must be ({n * n}, {nmat})")
        x = self._zeros(b.shape)
        check(lib().ohc_lu_solve(self._h, nmat, n, self._p(A), _ld(A, nmat), self._p(b),
                                 _ld(b, nmat), self._p(x), _ld(x, nmat)), self._h)
        return x

    # thermoList::THa_p
    def T_from_h(self, h, p, Y, T0, want_flag=False):
        n = h.shape[0]
        T = T0.copy() if not self.device_pointers else T0.clone()
        flag = self._zeros(n, np.int32) if want_flag else None
        check(lib().ohc_thermo_T_from_h(self._h, n, self._p(h), self._p(p), self._p(Y),
                                        _ld(Y, n), self._p(T),
                                        self._p(flag, "int32") if want_flag else None), self._h)
        return (T, flag) if want_flag else T

    # thermoList::TEa_p / TEa_rho
    def T_from_e(self, e, p, Y, T0, want_flag=False):
        n = e.shape[0]
        T = T0.copy() if not self.device_pointers else T0.clone()
        flag = self._zeros(n, np.int32) if want_flag else None
        check(lib().ohc_thermo_T_from_e(self._h, n, self._p(e), self._p(p), self._p(Y),
                                        _ld(Y, n), self._p(T),
                                        self._p(flag, "int32") if want_flag else None), self._h)
        return (T, flag) if want_flag else T

    # thermoList::cp_p / cv_p / ha_p / ea_p
    def mix_thermo(self, T, p, Y):
        n = T.shape[0]
        out = {k: self._zeros(n) for k in ("cp", "cv", "ha", "ea")}
        check(lib().ohc_mix_thermo(self._h, n, self._p(T), self._p(p), self._p(Y), _ld(Y, n),
                                   *[self._p(out[k]) for k in ("cp", "cv", "ha", "ea")]), self._h)
        return out

    # transportList::muKappa + Df, thermoList cp / ha
    # chemistrySource::calculateSourceTerms(Imp): R [ns, n] and, with diag, the
    # (coupled) diagonal d(omega_k)/d(c_k) [ns or ns-1, n]
    def source_terms(self, T, p, rho, Y, diag=False, without_last=False):
        n = T.shape[0]
        R = self._zeros((self.ns, n))
        D = self._zeros((self.ns, n)) if diag else None
        rc = lib().ohc_chem_source(self._h, n, self._p(T), self._p(p), self._p(rho), self._p(Y),
                                   _ld(Y, n), self._p(R), _ld(R, n),
                                   self._p(D) if diag else None, _ld(D, n) if diag else n,
                                   int(without_last))
        check(rc, self._h)
        return (R, D) if diag else R

    # the GPU path's calculateSourceTermsImp[Async][Hybrid]: R [ns, n] and the
    # substituted mass-action diagonal [ns-1, n] (float32 with fp32=True)
    def source_terms_kim(self, T, p, rho, Y, fp32=False):
        n = T.shape[0]
        R = self._zeros((self.ns, n))
        if self.device_pointers:
            import torch

            D = torch.zeros((self.ns - 1, n), dtype=torch.float32 if fp32 else torch.float64,
                            device=f"cuda:{self.device}")
            dptr = C.c_void_p(D.data_ptr())
        else:
            D = np.zeros((self.ns - 1, n), dtype=np.float32 if fp32 else np.float64)
            dptr = C.c_void_p(D.ctypes.data)
        rc = lib().ohc_chem_source_kim(self._h, n, self._p(T), self._p(p), self._p(rho),
                                       self._p(Y), _ld(Y, n), self._p(R), _ld(R, n), dptr, n,
                                       int(fp32))
        check(rc, self._h)
        return R, D

    def mix_props(self, T, p, Y, mu=True, kappa=True, Dk=True, cp=True, ha=True, hk=False):
        n = T.shape[0]
        out = {}
        for key, want, shape in [("mu", mu, n), ("kappa", kappa, n), ("Dk", Dk, (self.ns, n)),
                                 ("cp", cp, n), ("ha", ha, n), ("hk", hk, (self.ns, n))]:
            out[key] = self._zeros(shape) if want else None
        check(lib().ohc_mix_props(self._h, n, self._p(T), self._p(p), self._p(Y), _ld(Y, n),
                                  self._p(out["mu"]), self._p(out["kappa"]), self._p(out["Dk"]),
                                  n, self._p(out["cp"]), self._p(out["ha"]), self._p(out["hk"])),
              self._h)
        return {k: v for k, v in out.items() if v is not None}

    def _zeros(self, shape, dtype=np.float64):
        if self.device_pointers:
            import torch

            tdt = torch.float64 if dtype == np.float64 else torch.int32
            return torch.zeros(shape, dtype=tdt, device=f"cuda:{self.device}")
        return np.zeros(shape, dtype=dtype)
