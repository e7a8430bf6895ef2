"""The C-ABI library loads and exports every entry point include/ohc.h declares."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "ohc.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ohc_\w+)\s*\(", text, re.M)))


def test_exports_match_header():
    from openhurricane_amd import _lib

    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(_lib.EXPORTS) == syms


def test_opts_defaults_are_reference_defaults():
    from openhurricane_amd import _lib

    o = _lib.default_opts()
    assert (o.max_steps, o.atol, o.rtol) == (10000, 1e-15, 1e-4)
    assert (o.atol_inc, o.rtol_inc) == (1e-7, 1e-3)
    assert (o.safe_scale, o.min_scale, o.max_scale, o.alpha_inc) == (0.8, 0.25, 2.0, 0.25)
    assert (o.T_threshold, o.T_alpha, o.T_beta) == (200.0, 0.2, 0.25)


def test_errors_are_reported_not_raised_in_c():
    from openhurricane_amd import _lib

    L = _lib.lib()
    h = C.c_void_p()
    rc = L.ohc_mech_load(b"/nonexistent.che", None, None, C.byref(h))
    assert rc == _lib.OHC_ERR_ARG
    assert b"cannot read" in L.ohc_last_error(None)
    rc = L.ohc_ctx_create(None, None, 0, C.byref(h))
    assert rc == _lib.OHC_ERR_ARG


def test_tables_round_trip(h2mech):
    from openhurricane_amd.chemistry import Mechanism

    t = h2mech.tables()
    m2 = Mechanism.from_tables(t)
    t2 = m2.tables()
    for k in ["W", "nasa_lo", "nasa_hi", "arr", "eff", "f_idx", "b_ord", "rev", "kind"]:
        np.testing.assert_array_equal(t[k], t2[k])
    assert m2.hash == h2mech.hash
