"""ctypes binding of libgpp_hip.so, the C ABI declared in include/gpp.h.

Used by the tests, bench.py and the smoke entry point to drive the HIP
library exactly as a C caller would.  There is no fallback: if the built
library is missing, :func:`load` raises.

Arrays follow the reference mini-app's layout (gppComplex.cpp:264-271):
complex arrays are interleaved (re, im) float64 -- numpy ``complex128`` --
row-major, aqsmtemp/aqsntemp ``[number_bands, ncouls]``, I_eps_array /
wtilde_array ``[ngpown, ncouls]``, vcoul ``[ncouls]`` float64,
inv_igp_index ``[ngpown]`` int32, wx_array ``[nfreq]`` float64.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libgpp_hip.so")

# gpp_status, include/gpp.h
GPP_OK = 0
GPP_ERR_INVALID_ARGUMENT = 1
GPP_ERR_BAD_INDEX = 2
GPP_ERR_NO_DEVICE = 3
GPP_ERR_OUT_OF_MEMORY = 4
GPP_ERR_LAUNCH = 5
GPP_ERR_UNSUPPORTED = 6
GPP_ERR_NO_INPUTS = 7

# gpp_option
GPP_OPT_OZAKI_SLICES = 1
GPP_OPT_OZAKI_LAYOUT = 2
GPP_OPT_GRAM_TILE = 3
GPP_OPT_REDUCE_CHUNK = 4
GPP_OPT_REDUCE_FREQS = 5
GPP_OPT_MAX_BUCKETS = 6
GPP_OPT_REDUCE_FREQS_PER_BLOCK = 7
GPP_OPT_OZAKI_GROUP = 8
GPP_OPT_FUSED_REDUCE = 9

# gpp_form
FORMS = {
    "direct": 0,
    "direct-strict": 1,
    "factored-f64": 2,
    "factored-ozaki": 3,
    "factored-ozaki-i8": 4,
}
FORM_NAMES = {v: k for k, v in FORMS.items()}

# Every symbol include/gpp.h declares.
EXPORTED_SYMBOLS = (
    "gpp_last_error",
    "gpp_abi_version",
    "gpp_ctx_create",
    "gpp_ctx_destroy",
    "gpp_upload",
    "gpp_bind_device",
    "gpp_fill_deterministic",
    "gpp_fill_seeded",
    "gpp_generate_seeded",
    "gpp_sigma",
    "gpp_sigma_async",
    "gpp_ctx_wait",
    "gpp_ctx_set_stream",
    "gpp_ctx_set_columns",
    "gpp_ctx_timing_summary",
    "gpp_set_band_frequencies",
    "gpp_file_write",
    "gpp_file_info",
    "gpp_file_read",
    "gpp_upload_file",
    "gpp_ctx_band_buckets",
    "gpp_ctx_last_form",
    "gpp_ctx_set_option",
    "gpp_ctx_get_option",
    "gpp_ctx_last_timing",
    "gpp_ctx_stage_name",
    "gpp_ctx_autotune",
    "gpp_batch_create",
    "gpp_batch_create_seeded",
    "gpp_batch_sigma",
    "gpp_batch_autotune",
    "gpp_batch_use_graph",
    "gpp_batch_last_timing",
    "gpp_batch_last_stages",
    "gpp_batch_destroy",
    "d_till_nvband",
    "d_reduce_achstemp",
    "d_noFlagOCCSolver",
    "d_noFlagOCCSolver_noalloc",
)


class GppProblem(ctypes.Structure):
    _fields_ = [
        ("number_bands", ctypes.c_int32),
        ("nvband", ctypes.c_int32),
        ("ncouls", ctypes.c_int32),
        ("ngpown", ctypes.c_int32),
        ("nfreq", ctypes.c_int32),
        ("occ", ctypes.c_double),
    ]


class GppOutputs(ctypes.Structure):
    _fields_ = [
        ("achtemp", ctypes.c_void_p),
        ("asxtemp", ctypes.c_void_p),
        ("acht_n1_loc", ctypes.c_void_p),
        ("achstemp", ctypes.c_void_p),
    ]


class GppFileHeader(ctypes.Structure):
    _fields_ = [
        ("problem", GppProblem),
        ("flags", ctypes.c_uint32),
        ("payload_bytes", ctypes.c_uint64),
        ("checksum", ctypes.c_uint64),
    ]


GPP_FILE_WX_PER_BAND = 1
GPP_FILE_HAS_EXPECTED = 2
GPP_ERR_IO = 9


class GppDeviceInputs(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("inv_igp_index", "wx_array", "vcoul", "aqsmtemp", "aqsntemp",
                                              "I_eps_array", "wtilde_array")]


class GppError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"gpp status {status}: {message}")
        self.status = status


_lib: Optional[ctypes.CDLL] = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libgpp_hip.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    vp, i32, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    lib.gpp_last_error.restype = ctypes.c_char_p
    lib.gpp_last_error.argtypes = []
    lib.gpp_abi_version.restype = i32
    lib.gpp_ctx_create.argtypes = [P(vp), i32]
    lib.gpp_ctx_destroy.argtypes = [vp]
    lib.gpp_upload.argtypes = [vp, P(GppProblem)] + [vp] * 7
    lib.gpp_bind_device.argtypes = [vp, P(GppProblem)] + [vp] * 7
    lib.gpp_fill_deterministic.argtypes = [P(GppProblem), i32] + [vp] * 7
    lib.gpp_fill_seeded.argtypes = [P(GppProblem), ctypes.c_uint64, dbl, dbl] + [vp] * 7
    lib.gpp_generate_seeded.argtypes = [vp, P(GppProblem), ctypes.c_uint64, dbl, dbl]
    lib.gpp_sigma.argtypes = [vp, i32, P(GppOutputs)]
    lib.gpp_sigma_async.argtypes = [vp, i32, i32]
    lib.gpp_ctx_wait.argtypes = [vp, P(GppOutputs)]
    lib.gpp_ctx_set_stream.argtypes = [vp, vp]
    lib.gpp_ctx_set_columns.argtypes = [vp, i32, i32]
    lib.gpp_ctx_timing_summary.argtypes = [vp, i32, P(i32), P(dbl), P(dbl), i32, P(i32)]
    lib.gpp_set_band_frequencies.argtypes = [vp, vp]
    lib.gpp_file_write.argtypes = [ctypes.c_char_p, P(GppProblem), vp, vp, i32] + [vp] * 5 + [P(GppOutputs)]
    lib.gpp_file_info.argtypes = [ctypes.c_char_p, P(GppFileHeader)]
    lib.gpp_file_read.argtypes = [ctypes.c_char_p] + [vp] * 7 + [P(GppOutputs)]
    lib.gpp_upload_file.argtypes = [vp, ctypes.c_char_p, P(GppOutputs)]
    lib.gpp_ctx_band_buckets.argtypes = [vp, P(i32)]
    lib.gpp_ctx_last_form.argtypes = [vp, P(i32)]
    lib.gpp_ctx_last_timing.argtypes = [vp, P(dbl), P(dbl), i32, P(i32)]
    lib.gpp_ctx_set_option.argtypes = [vp, i32, i32]
    lib.gpp_ctx_get_option.argtypes = [vp, i32, P(i32)]
    lib.gpp_ctx_stage_name.restype = ctypes.c_char_p
    lib.gpp_ctx_stage_name.argtypes = [vp, i32]
    lib.gpp_ctx_autotune.argtypes = [vp, i32]
    lib.gpp_batch_create.argtypes = [vp, i32, P(GppProblem), P(GppDeviceInputs), P(vp)]
    lib.gpp_batch_create_seeded.argtypes = [vp, i32, P(GppProblem), P(ctypes.c_uint64), P(dbl), P(dbl), P(vp)]
    lib.gpp_batch_sigma.argtypes = [vp, i32, P(GppOutputs)]
    lib.gpp_batch_autotune.argtypes = [vp, i32]
    lib.gpp_batch_use_graph.argtypes = [vp, i32]
    lib.gpp_batch_last_timing.argtypes = [vp, P(dbl), P(i32), P(i32), P(i32)]
    lib.gpp_batch_last_stages.argtypes = [vp, P(dbl)]
    lib.gpp_batch_destroy.argtypes = [vp]
    lib.d_till_nvband.argtypes = [i32, i32, i32] + [vp] * 11 + [dbl]
    lib.d_reduce_achstemp.argtypes = [i32, vp, i32, vp, vp, vp, vp, vp, i32, vp]
    lib.d_noFlagOCCSolver.argtypes = [i32, i32, i32, i32] + [vp] * 10
    lib.d_noFlagOCCSolver_noalloc.argtypes = [i32, i32, i32, i32] + [vp] * 10
    for name in ("gpp_ctx_create", "gpp_ctx_destroy", "gpp_upload", "gpp_bind_device", "gpp_ctx_set_stream",
                 "gpp_ctx_set_columns", "gpp_sigma_async", "gpp_ctx_wait",
                 "gpp_fill_deterministic", "gpp_fill_seeded", "gpp_generate_seeded", "gpp_sigma",
                 "gpp_ctx_last_timing", "gpp_ctx_set_option", "gpp_ctx_get_option", "gpp_ctx_autotune",
                 "gpp_batch_create", "gpp_batch_create_seeded", "gpp_batch_sigma", "gpp_batch_autotune",
                 "gpp_batch_use_graph", "gpp_batch_last_timing", "gpp_batch_last_stages", "gpp_batch_destroy",
                 "d_till_nvband", "d_reduce_achstemp", "d_noFlagOCCSolver", "d_noFlagOCCSolver_noalloc"):
        getattr(lib, name).restype = i32
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != GPP_OK:
        raise GppError(status, load().gpp_last_error().decode())


def _ptr(a: Optional[np.ndarray]) -> Optional[int]:
    return None if a is None else a.ctypes.data


@dataclass
class Problem:
    number_bands: int
    nvband: int
    ncouls: int
    ngpown: int
    nfreq: int = 3
    occ: float = 1.0

    def c(self) -> GppProblem:
        return GppProblem(self.number_bands, self.nvband, self.ncouls, self.ngpown, self.nfreq, self.occ)

    @property
    def updates(self) -> int:
        """GPP sigma updates, one per (n1, igp, ig, iw)."""
        return self.number_bands * self.ngpown * self.ncouls * self.nfreq


def empty_inputs(p: Problem) -> Dict[str, np.ndarray]:
    return {
        "inv_igp_index": np.zeros(p.ngpown, np.int32),
        "wx_array": np.zeros(p.nfreq, np.float64),
        "vcoul": np.zeros(p.ncouls, np.float64),
        "aqsmtemp": np.zeros((p.number_bands, p.ncouls), np.complex128),
        "aqsntemp": np.zeros((p.number_bands, p.ncouls), np.complex128),
        "I_eps_array": np.zeros((p.ngpown, p.ncouls), np.complex128),
        "wtilde_array": np.zeros((p.ngpown, p.ncouls), np.complex128),
    }


_INPUT_ORDER = ("inv_igp_index", "wx_array", "vcoul", "aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array")


def fill_deterministic(p: Problem, igp_zero_based: bool = False) -> Dict[str, np.ndarray]:
    """The reference's deterministic inputs (gppComplex.cpp:296-329)."""
    arrs = empty_inputs(p)
    prob = p.c()
    _check(load().gpp_fill_deterministic(ctypes.byref(prob), int(igp_zero_based),
                                         *[_ptr(arrs[k]) for k in _INPUT_ORDER]))
    return arrs


def fill_seeded(p: Problem, seed: int, wx_center: float = 4.0, dw: float = 1.0) -> Dict[str, np.ndarray]:
    """Seeded inputs with the benchmark distributions (see gpp_fill_seeded)."""
    arrs = empty_inputs(p)
    prob = p.c()
    _check(load().gpp_fill_seeded(ctypes.byref(prob), seed, wx_center, dw,
                                  *[_ptr(arrs[k]) for k in _INPUT_ORDER]))
    return arrs


_OUT_KEYS = ("achtemp", "asxtemp", "acht_n1_loc", "achstemp")


def _expected_arrays(p: Problem) -> Dict[str, np.ndarray]:
    return {"achtemp": np.zeros(p.nfreq, np.complex128), "asxtemp": np.zeros(p.nfreq, np.complex128),
            "acht_n1_loc": np.zeros(p.number_bands, np.complex128), "achstemp": np.zeros(1, np.complex128)}


def write_problem(path: str, p: Problem, arrs: Dict[str, np.ndarray],
                  expected: Optional[Dict[str, np.ndarray]] = None) -> None:
    """Writes a problem file (.gpp).  A 2-D wx_array is stored as
    band-dependent frequencies; expected outputs are optional."""
    wx = np.ascontiguousarray(arrs["wx_array"], np.float64)
    a = {k: _contig(arrs[k], np.complex128) for k in ("aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array")}
    a["inv_igp_index"] = _contig(arrs["inv_igp_index"], np.int32)
    a["vcoul"] = _contig(arrs["vcoul"], np.float64)
    a["wx_array"] = wx
    o = None
    if expected is not None:
        e = {k: (None if expected.get(k) is None else np.ascontiguousarray(np.atleast_1d(expected[k]), np.complex128))
             for k in _OUT_KEYS}
        o = GppOutputs(*[_ptr(e[k]) for k in _OUT_KEYS])
    prob = p.c()
    _check(load().gpp_file_write(os.fsencode(path), ctypes.byref(prob), _ptr(a["inv_igp_index"]), _ptr(wx),
                                 1 if wx.ndim == 2 else 0, _ptr(a["vcoul"]),
                                 *[_ptr(a[k]) for k in ("aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array")],
                                 ctypes.byref(o) if o is not None else None))


def file_info(path: str) -> Dict[str, object]:
    h = GppFileHeader()
    _check(load().gpp_file_info(os.fsencode(path), ctypes.byref(h)))
    q = h.problem
    return {"problem": Problem(q.number_bands, q.nvband, q.ncouls, q.ngpown, q.nfreq, q.occ),
            "flags": h.flags, "payload_bytes": h.payload_bytes, "checksum": h.checksum}


def read_problem(path: str):
    """(Problem, inputs, expected outputs or None) from a problem file."""
    info = file_info(path)
    p = info["problem"]
    arrs = empty_inputs(p)
    if info["flags"] & GPP_FILE_WX_PER_BAND:
        arrs["wx_array"] = np.zeros((p.number_bands, p.nfreq), np.float64)
    exp = _expected_arrays(p) if info["flags"] & GPP_FILE_HAS_EXPECTED else None
    o = GppOutputs(*[_ptr(exp[k]) for k in _OUT_KEYS]) if exp is not None else None
    _check(load().gpp_file_read(os.fsencode(path), *[_ptr(arrs[k]) for k in _INPUT_ORDER],
                                ctypes.byref(o) if o is not None else None))
    if exp is not None:
        exp["achstemp"] = exp["achstemp"][0]
    return p, arrs, exp


def _contig(a: np.ndarray, dtype) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=dtype)


class Context:
    """A gpp_ctx on one HIP device."""

    def __init__(self, device: int = 0):
        lib = load()
        h = ctypes.c_void_p()
        _check(lib.gpp_ctx_create(ctypes.byref(h), device))
        self._h = h
        self.problem: Optional[Problem] = None

    def close(self) -> None:
        if self._h:
            load().gpp_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, p: Problem, arrs: Dict[str, np.ndarray]) -> None:
        """Copies the inputs to the device.  A 2-D wx_array [number_bands][nfreq]
        uploads its first row as the shared wx_array and then sets the rows as
        band-dependent frequencies (set_band_frequencies)."""
        wx = np.asarray(arrs["wx_array"], np.float64)
        a = {
            "inv_igp_index": _contig(arrs["inv_igp_index"], np.int32),
            "wx_array": _contig(wx[0] if wx.ndim == 2 else wx, np.float64),
            "vcoul": _contig(arrs["vcoul"], np.float64),
        }
        for k in ("aqsmtemp", "aqsntemp", "I_eps_array", "wtilde_array"):
            a[k] = _contig(arrs[k], np.complex128)
        prob = p.c()
        _check(load().gpp_upload(self._h, ctypes.byref(prob), *[_ptr(a[k]) for k in _INPUT_ORDER]))
        self.problem = p
        if wx.ndim == 2:
            self.set_band_frequencies(wx)

    def upload_file(self, path: str):
        """Streams a problem file to the device; returns the file's expected
        outputs, or None when it holds none."""
        info = file_info(path)
        p = info["problem"]
        exp = _expected_arrays(p)
        o = GppOutputs(*[_ptr(exp[k]) for k in _OUT_KEYS])
        _check(load().gpp_upload_file(self._h, os.fsencode(path), ctypes.byref(o)))
        self.problem = p
        if not info["flags"] & GPP_FILE_HAS_EXPECTED:
            return None
        exp["achstemp"] = exp["achstemp"][0]
        return exp

    def set_band_frequencies(self, wx_bands: Optional[np.ndarray]) -> None:
        """Band n1 uses wx_bands[n1] ([number_bands][nfreq]); None restores
        the shared wx_array."""
        if wx_bands is None:
            _check(load().gpp_set_band_frequencies(self._h, None))
            return
        w = _contig(wx_bands, np.float64)
        p = self.problem
        if p is None or w.shape != (p.number_bands, p.nfreq):
            raise ValueError("wx_bands must be [number_bands][nfreq] of the uploaded problem")
        _check(load().gpp_set_band_frequencies(self._h, _ptr(w)))

    def band_buckets(self) -> int:
        n = ctypes.c_int()
        _check(load().gpp_ctx_band_buckets(self._h, ctypes.byref(n)))
        return n.value

    def last_form(self) -> str:
        f = ctypes.c_int()
        _check(load().gpp_ctx_last_form(self._h, ctypes.byref(f)))
        return FORM_NAMES[f.value]

    def bind_device(self, p: Problem, ptrs: Dict[str, int]) -> None:
        """Binds caller-owned device buffers (integer device addresses)."""
        prob = p.c()
        _check(load().gpp_bind_device(self._h, ctypes.byref(prob), *[ptrs[k] for k in _INPUT_ORDER]))
        self.problem = p

    def generate_seeded(self, p: Problem, seed: int, wx_center: float = 4.0, dw: float = 1.0) -> None:
        prob = p.c()
        _check(load().gpp_generate_seeded(self._h, ctypes.byref(prob), seed, wx_center, dw))
        self.problem = p

    def sigma(self, form: str = "direct", acht_n1_loc: Optional[bool] = None) -> Dict[str, np.ndarray]:
        """Evaluates the band loop.  acht_n1_loc (per-band sums) is requested
        by default from the direct forms only: the factored forms compute it
        with a second product of the Gram product's size, so it is opt-in."""
        if self.problem is None:
            raise GppError(GPP_ERR_NO_INPUTS, "no inputs")
        if acht_n1_loc is None:
            acht_n1_loc = form.startswith("direct")
        p = self.problem
        out = {
            "achtemp": np.zeros(p.nfreq, np.complex128),
            "asxtemp": np.zeros(p.nfreq, np.complex128),
            "acht_n1_loc": np.zeros(p.number_bands, np.complex128) if acht_n1_loc else None,
            "achstemp": np.zeros(1, np.complex128),
        }
        o = GppOutputs(*[_ptr(out[k]) for k in ("achtemp", "asxtemp", "acht_n1_loc", "achstemp")])
        _check(load().gpp_sigma(self._h, FORMS[form], ctypes.byref(o)))
        out["achstemp"] = out["achstemp"][0]
        return out

    def sigma_async(self, form: str = "direct", acht_n1_loc: bool = False) -> None:
        """Enqueues one evaluation without waiting (see wait)."""
        _check(load().gpp_sigma_async(self._h, FORMS[form], 1 if acht_n1_loc else 0))

    def wait(self, acht_n1_loc: bool = False) -> Dict[str, np.ndarray]:
        """Waits for the enqueued evaluations; returns the last one's outputs."""
        p = self.problem
        out = {
            "achtemp": np.zeros(p.nfreq, np.complex128),
            "asxtemp": np.zeros(p.nfreq, np.complex128),
            "acht_n1_loc": np.zeros(p.number_bands, np.complex128) if acht_n1_loc else None,
            "achstemp": np.zeros(1, np.complex128),
        }
        o = GppOutputs(*[_ptr(out[k]) for k in ("achtemp", "asxtemp", "acht_n1_loc", "achstemp")])
        _check(load().gpp_ctx_wait(self._h, ctypes.byref(o)))
        out["achstemp"] = out["achstemp"][0]
        return out

    def set_stream(self, stream: Optional[int]) -> None:
        """Enqueues the context's work on a caller-owned HIP stream (its
        integer handle); None returns to the context's own stream."""
        _check(load().gpp_ctx_set_stream(self._h, stream))

    def set_columns(self, j0: int = 0, j1: int = 0) -> None:
        """Evaluates only the G' columns [j0, j1) of the current inputs (one
        rank's slab); (0, 0) restores every column."""
        _check(load().gpp_ctx_set_columns(self._h, j0, j1))

    def timing_summary(self, last_n: int) -> Dict[str, float]:
        """Average GPU milliseconds of the last evaluations of the same kind:
        'evaluations', 'total' and one entry per stage."""
        lib = load()
        nev, total, n = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        stages = (ctypes.c_double * 16)()
        _check(lib.gpp_ctx_timing_summary(self._h, last_n, ctypes.byref(nev), ctypes.byref(total), stages, 16,
                                          ctypes.byref(n)))
        res = {"evaluations": nev.value, "total": total.value}
        for i in range(n.value):
            res[lib.gpp_ctx_stage_name(self._h, i).decode()] = stages[i]
        return res

    def autotune(self, form: str = "factored-f64") -> Dict[str, int]:
        """Times candidate Gram tiles / reduce schedules and keeps the fastest."""
        _check(load().gpp_ctx_autotune(self._h, FORMS[form]))
        return {"gram_tile": self.get_option(GPP_OPT_GRAM_TILE),
                "reduce_chunk": self.get_option(GPP_OPT_REDUCE_CHUNK),
                "reduce_freqs": self.get_option(GPP_OPT_REDUCE_FREQS)}

    def set_option(self, option: int, value: int) -> None:
        _check(load().gpp_ctx_set_option(self._h, option, value))

    def get_option(self, option: int) -> int:
        v = ctypes.c_int()
        _check(load().gpp_ctx_get_option(self._h, option, ctypes.byref(v)))
        return v.value

    def last_timing(self) -> Dict[str, float]:
        """GPU milliseconds of the last sigma: 'total' and one entry per stage."""
        total = ctypes.c_double()
        stages = (ctypes.c_double * 16)()
        n = ctypes.c_int()
        lib = load()
        _check(lib.gpp_ctx_last_timing(self._h, ctypes.byref(total), stages, 16, ctypes.byref(n)))
        res = {"total": total.value}
        for i in range(n.value):
            res[lib.gpp_ctx_stage_name(self._h, i).decode()] = stages[i]
        return res


class Batch:
    """A gpp_batch: many problems through the persistent factored pipeline.

    Either ``seeds`` (library-owned inputs generated on the device; optional
    per-instance ``wx_center`` and ``dw``) or ``device_inputs`` (a list of
    dicts of device addresses, keys as in gpp_device_inputs) must be given.
    """

    def __init__(self, ctx: Context, problems, seeds=None, wx_center=None, dw=None, device_inputs=None):
        lib = load()
        n = len(problems)
        self.problems = list(problems)
        self._ctx = ctx  # keeps the context alive
        probs = (GppProblem * n)(*[p.c() for p in problems])
        h = ctypes.c_void_p()
        if device_inputs is not None:
            ins = (GppDeviceInputs * n)(*[GppDeviceInputs(*[d[k] for k, _ in GppDeviceInputs._fields_])
                                          for d in device_inputs])
            _check(lib.gpp_batch_create(ctx._h, n, probs, ins, ctypes.byref(h)))
        else:
            sd = (ctypes.c_uint64 * n)(*seeds)
            wc = (ctypes.c_double * n)(*wx_center) if wx_center is not None else None
            dd = (ctypes.c_double * n)(*dw) if dw is not None else None
            _check(lib.gpp_batch_create_seeded(ctx._h, n, probs, sd, wc, dd, ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if self._h:
            load().gpp_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sigma(self, form: str = "factored-f64"):
        outs_np = []
        outs = (GppOutputs * len(self.problems))()
        for i, p in enumerate(self.problems):
            o = {"achtemp": np.zeros(p.nfreq, np.complex128), "asxtemp": np.zeros(p.nfreq, np.complex128),
                 "achstemp": np.zeros(1, np.complex128)}
            outs[i] = GppOutputs(_ptr(o["achtemp"]), _ptr(o["asxtemp"]), None, _ptr(o["achstemp"]))
            outs_np.append(o)
        _check(load().gpp_batch_sigma(self._h, FORMS[form], outs))
        for o in outs_np:
            o["achstemp"] = o["achstemp"][0]
        return outs_np

    def autotune(self, form: str = "factored-f64") -> None:
        _check(load().gpp_batch_autotune(self._h, FORMS[form]))

    def use_graph(self, flag: bool) -> None:
        _check(load().gpp_batch_use_graph(self._h, int(bool(flag))))

    def last_timing(self) -> Dict[str, float]:
        ms = ctypes.c_double()
        t, c, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(load().gpp_batch_last_timing(self._h, ctypes.byref(ms), ctypes.byref(t), ctypes.byref(c),
                                            ctypes.byref(f)))
        st = (ctypes.c_double * 4)()
        _check(load().gpp_batch_last_stages(self._h, st))
        return {"ms": ms.value, "gram_tile": t.value, "reduce_chunk": c.value, "reduce_freqs": f.value,
                "stages_ms": {"gather_achstemp": st[0], "gram": st[1], "pole_reduce": st[2], "final": st[3]}}
