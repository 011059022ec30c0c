"""ctypes loader for the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; it is the checker, never the thing measured or shipped.

* ``liboracle.so`` (built from oracle/gpp_oracle.c) -- the C restatement of
  the reference's band loop, gppComplex.cpp:16-150 and :335-383.
* ``_ref/libgppref.so`` (built from the reference's own gppComplex.cpp plus
  oracle/ref_harness.cpp, only where the reference sources exist) -- used to
  pin the restatement; optional.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libgppref.so")

_oracle: Optional[ctypes.CDLL] = None
_ref: Optional[ctypes.CDLL] = None


def _sigma_argtypes(extra):
    vp = ctypes.c_void_p
    return [ctypes.c_int] * 5 + [ctypes.c_double] + [vp] * 7 + extra


def load_oracle() -> ctypes.CDLL:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run make -C oracle")
        lib = ctypes.CDLL(ORACLE_LIB)
        vp = ctypes.c_void_p
        lib.oracle_sigma.argtypes = _sigma_argtypes([vp, vp, vp, vp, ctypes.c_int])
        lib.oracle_sigma.restype = ctypes.c_int
        lib.oracle_sigma_wxb.argtypes = ([ctypes.c_int] * 5 + [ctypes.c_double, vp, vp, ctypes.c_int]
                                         + [vp] * 5 + [vp, vp, vp, vp, ctypes.c_int])
        lib.oracle_sigma_wxb.restype = ctypes.c_int
        lib.oracle_sigma_dd.argtypes = _sigma_argtypes([vp, vp, ctypes.c_int])
        lib.oracle_sigma_dd.restype = ctypes.c_int
        lib.oracle_sigma_dd2.argtypes = _sigma_argtypes([vp, vp, vp, vp, ctypes.c_int])
        lib.oracle_sigma_dd2.restype = ctypes.c_int
        lib.oracle_pole_terms.argtypes = [ctypes.c_double, vp, vp, vp, vp]
        lib.oracle_pole_terms.restype = ctypes.c_int
        lib.oracle_max_threads.restype = ctypes.c_int
        lib.oracle_dd_sum.argtypes = [vp, ctypes.c_long, vp]
        lib.oracle_dd_sum.restype = None
        lib.oracle_dd_sum_dd.argtypes = [vp, ctypes.c_long, vp]
        lib.oracle_dd_sum_dd.restype = None
        lib.oracle_dd_mul_d.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, vp]
        lib.oracle_dd_mul_d.restype = None
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def load_ref() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_LIB)
        vp = ctypes.c_void_p
        lib.ref_sigma.argtypes = _sigma_argtypes([vp, vp, vp, vp])
        lib.ref_sigma.restype = ctypes.c_int
        lib.ref_pole_terms.argtypes = [ctypes.c_double, vp, vp, vp, vp]
        lib.ref_pole_terms.restype = None
        _ref = lib
    return _ref


def _prep(arrs):
    c = lambda k, t: np.ascontiguousarray(arrs[k], dtype=t)
    return [c("inv_igp_index", np.int32), c("wx_array", np.float64), c("vcoul", np.float64),
            c("aqsmtemp", np.complex128), c("aqsntemp", np.complex128),
            c("I_eps_array", np.complex128), c("wtilde_array", np.complex128)]


def _sizes(arrs, nvband, occ):
    nb, ncouls = arrs["aqsmtemp"].shape
    ngpown = arrs["I_eps_array"].shape[0]
    nfreq = arrs["wx_array"].shape[-1]
    return [nb, nvband, ncouls, ngpown, nfreq, ctypes.c_double(occ)]


def sigma(arrs: Dict[str, np.ndarray], nvband: int, occ: float = 1.0, threads: int = 1,
          use_ref: bool = False) -> Dict[str, np.ndarray]:
    """Runs the oracle (or the compiled reference harness) on the inputs.
    A 2-D wx_array [number_bands][nfreq] gives each band its own frequencies
    (oracle only: the reference shares one wx_array between all bands)."""
    ins = _prep(arrs)
    sizes = _sizes(arrs, nvband, occ)
    per_band = np.ndim(arrs["wx_array"]) == 2
    nb, nfreq = sizes[0], sizes[4]
    out = {"achtemp": np.zeros(nfreq, np.complex128), "asxtemp": np.zeros(nfreq, np.complex128),
           "acht_n1_loc": np.zeros(nb, np.complex128), "achstemp": np.zeros(1, np.complex128)}
    optr = [out[k].ctypes.data for k in ("achtemp", "asxtemp", "acht_n1_loc", "achstemp")]
    iptr = [a.ctypes.data for a in ins]
    if per_band:
        if use_ref:
            raise ValueError("the reference harness takes one wx_array for all bands")
        if np.shape(arrs["wx_array"])[0] != nb:
            raise ValueError("per-band wx_array must be [number_bands][nfreq]")
        rc = load_oracle().oracle_sigma_wxb(*sizes, iptr[0], iptr[1], nfreq, *iptr[2:], *optr, threads)
    elif use_ref:
        rc = load_ref().ref_sigma(*sizes, *iptr, *optr)
    else:
        rc = load_oracle().oracle_sigma(*sizes, *iptr, *optr, threads)
    if rc != 0:
        raise ValueError(f"oracle rejected the inputs (code {rc})")
    out["achstemp"] = out["achstemp"][0]
    return out


def sigma_dd(arrs: Dict[str, np.ndarray], nvband: int, occ: float = 1.0, threads: int = 1,
             all_outputs: bool = False):
    """Compensated (double-double accumulated) achtemp, asxtemp and, with
    all_outputs, acht_n1_loc and achstemp: the sums of the reference's own
    rounded terms, without the error of its left-to-right sums."""
    if np.ndim(arrs["wx_array"]) != 1:
        raise ValueError("sigma_dd takes one wx_array for all bands")
    ins = _prep(arrs)
    sizes = _sizes(arrs, nvband, occ)
    nb, nfreq = sizes[0], sizes[4]
    ach = np.zeros(nfreq, np.complex128)
    asx = np.zeros(nfreq, np.complex128)
    if not all_outputs:
        rc = load_oracle().oracle_sigma_dd(*sizes, *[a.ctypes.data for a in ins], ach.ctypes.data,
                                           asx.ctypes.data, threads)
        if rc != 0:
            raise ValueError(f"oracle rejected the inputs (code {rc})")
        return {"achtemp": ach, "asxtemp": asx}
    n1 = np.zeros(nb, np.complex128)
    achs = np.zeros(1, np.complex128)
    rc = load_oracle().oracle_sigma_dd2(*sizes, *[a.ctypes.data for a in ins], ach.ctypes.data, asx.ctypes.data,
                                        n1.ctypes.data, achs.ctypes.data, threads)
    if rc != 0:
        raise ValueError(f"oracle rejected the inputs (code {rc})")
    return {"achtemp": ach, "asxtemp": asx, "acht_n1_loc": n1, "achstemp": achs[0]}


def pole_terms(wxt: float, wtilde: complex, ieps: complex, use_ref: bool = False):
    """(branch, sch, ssx) of the occupied-band solver for one element;
    branch is None for the reference harness (it does not report it)."""
    wt = np.array([wtilde.real, wtilde.imag])
    ie = np.array([ieps.real, ieps.imag])
    sch = np.zeros(2)
    ssx = np.zeros(2)
    if use_ref:
        load_ref().ref_pole_terms(wxt, wt.ctypes.data, ie.ctypes.data, sch.ctypes.data, ssx.ctypes.data)
        br = None
    else:
        br = load_oracle().oracle_pole_terms(wxt, wt.ctypes.data, ie.ctypes.data, sch.ctypes.data,
                                             ssx.ctypes.data)
    return br, complex(sch[0], sch[1]), complex(ssx[0], ssx[1])


def dd_sum(x) -> tuple:
    """(hi, lo) of the oracle's double-double running sum of x, in order."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(2)
    load_oracle().oracle_dd_sum(x.ctypes.data, x.size, out.ctypes.data)
    return float(out[0]), float(out[1])


def dd_sum_dd(pairs) -> tuple:
    """(hi, lo) of the oracle's double-double sum of (hi, lo) pairs, in order."""
    p = np.ascontiguousarray(pairs, dtype=np.float64).reshape(-1, 2)
    out = np.zeros(2)
    load_oracle().oracle_dd_sum_dd(p.ctypes.data, p.shape[0], out.ctypes.data)
    return float(out[0]), float(out[1])


def dd_mul_d(hi: float, lo: float, b: float) -> tuple:
    """(hi, lo) of the oracle's double-double (hi, lo) times the double b."""
    out = np.zeros(2)
    load_oracle().oracle_dd_mul_d(hi, lo, b, out.ctypes.data)
    return float(out[0]), float(out[1])


def max_threads() -> int:
    return load_oracle().oracle_max_threads()
