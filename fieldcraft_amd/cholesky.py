"""The solver crate's factor/solve API (crates/solver/src/cholesky.rs) over the C ABI of
``libfc_chol.so`` (include/fc_chol.h): the symbolic phase on the host, the numeric phase and the
solves on one MI355X.

Names, argument meaning and error behaviour follow the reference:

* ``UpperColumns`` (cholesky.rs:99-145) — the upper triangle by columns, ``from_symmetric``.
* ``FactorError`` and its kinds (cholesky.rs:346-361) — raised, never a silent fallback.
* ``Symbolic`` — ``Symbolic::new`` (cholesky.rs:1213): analyse once, factor many
  (cholesky.rs:1390-1395); exports the reference's symbolic outputs for parity checks.
* ``Factor`` — ``Factor::new`` / ``from_symbolic`` / ``entries`` / ``solve`` / ``solve_many``
  (cholesky.rs:1532-1570), plus ``solve_refined`` (FP64 refinement, ldlt.rs:524-563 made
  general) and ``values`` (the reference's trapezoid layout).

Nothing here computes on the CPU: every numeric call goes to the HIP library, and a missing
library raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _build

FP64, MIXED, FP32 = 0, 1, 2
PRECISIONS = {"fp64": FP64, "mixed": MIXED, "fp32": FP32}

_OK, _NPD, _SINGULAR, _TOO_LARGE, _OOM, _INTERNAL, _CANCELLED, _INVALID, _DEVICE, _DEGENERATE, _NOT_CONVERGED = range(11)


class FactorError(Exception):
    """Why a factorization or solve stopped (FactorError, cholesky.rs:346-361)."""

    kind = "internal"

    def __init__(self, message: str = "", column: int | None = None, entries: int | None = None,
                 limit: int | None = None):
        self.column, self.entries, self.limit = column, entries, limit
        super().__init__(message)


class NotPositiveDefinite(FactorError):
    kind = "not_positive_definite"


class Singular(FactorError):
    kind = "singular"


class TooLarge(FactorError):
    kind = "too_large"


class OutOfMemory(FactorError):
    kind = "out_of_memory"


class Internal(FactorError):
    kind = "internal"


class Cancelled(FactorError):
    kind = "cancelled"


class InvalidArgument(FactorError):
    kind = "invalid_argument"


class DeviceError(FactorError):
    kind = "device"


class DegenerateElement(FactorError):
    """An inverted or degenerate element met by the assembly (Error::Degenerate, fem
    element.rs:326-328); ``column`` holds its index in element order."""

    kind = "degenerate_element"


class NotConverged(FactorError):
    """The Lanczos iteration stopped without the wanted modes (Error::NotConverged,
    crates/solver/src/eigen.rs:407-411, 463, 467)."""

    kind = "not_converged"


_KINDS = {_NPD: NotPositiveDefinite, _SINGULAR: Singular, _TOO_LARGE: TooLarge, _OOM: OutOfMemory,
          _INTERNAL: Internal, _CANCELLED: Cancelled, _INVALID: InvalidArgument, _DEVICE: DeviceError,
          _DEGENERATE: DegenerateElement, _NOT_CONVERGED: NotConverged}


class _Error(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("column", C.c_uint64), ("entries", C.c_uint64),
                ("limit", C.c_uint64), ("message", C.c_char * 256)]


class _Opts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("device", C.c_int32), ("stream", C.c_void_p),
                ("cancel", C.POINTER(C.c_int32)), ("bf16_terms", C.c_int32), ("use_graph", C.c_int32),
                ("fallback", C.c_int32), ("pad0", C.c_int32), ("profile_mask", C.c_int32), ("pad", C.c_int32)]


class _Info(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ["n", "supernodes", "entries", "rows", "updates", "tasks", "top", "fronts",
                                           "levels", "stored"]] + \
               [("flops", C.c_double), ("front_flops", C.c_double), ("l_bytes_fp64", C.c_uint64),
                ("u_bytes_fp64", C.c_uint64), ("by_blocks", C.c_int32), ("pad", C.c_int32),
                ("launches", C.c_uint64), ("kind_flops", C.c_double * 8), ("kind_entries", C.c_double * 8),
                ("solve_bytes", C.c_uint64)]


class _LanczosOpts(C.Structure):
    _fields_ = [("count", C.c_uint64), ("min", C.c_double), ("max", C.c_double), ("block", C.c_uint32),
                ("pad", C.c_uint32), ("max_basis", C.c_uint64), ("tolerance", C.c_double)]


class _LanczosReport(C.Structure):
    _fields_ = [("found", C.c_uint64), ("basis", C.c_uint64), ("negative_count", C.c_uint64),
                ("negative", C.c_double * 16), ("seconds", C.c_double)]


class _RefineOpts(C.Structure):
    _fields_ = [("method", C.c_int32), ("max_iter", C.c_int32), ("target_berr", C.c_double)]


class _RefineReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("berr", C.c_double), ("seconds", C.c_double),
                ("berr_history", C.c_double * 64)]


class _Schedule(C.Structure):
    _fields_ = [("slab", C.c_int32), ("split_target", C.c_int32), ("split_side_fp64", C.c_int32),
                ("split_side_fp32", C.c_int32)]


class _TuneOpts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("device", C.c_int32), ("stream", C.c_void_p), ("repeats", C.c_int32),
                ("pad", C.c_int32)]


TUNE_MAX = 32


class _TuneReport(C.Structure):
    _fields_ = [("chosen", _Schedule), ("seconds_start", C.c_double), ("seconds_chosen", C.c_double),
                ("candidates", C.c_int32), ("pad", C.c_int32), ("tried", _Schedule * TUNE_MAX),
                ("tried_seconds", C.c_double * TUNE_MAX)]


def _sched_dict(c) -> dict:
    return {k: int(getattr(c, k)) for k, _ in _Schedule._fields_}


_lib = None
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

#: every symbol include/fc_chol.h declares
EXPORTS = ["fc_analyze", "fc_symbolic_info_get", "fc_symbolic_export", "fc_symbolic_free", "fc_factor_numeric",
           "fc_factor_numeric_device", "fc_factor_new", "fc_refactor", "fc_refactor_device", "fc_entries", "fc_solve",
           "fc_solve_many", "fc_solve_device", "fc_solve_device_noalloc", "fc_factor_export", "fc_factor_seconds", "fc_factor_precision", "fc_factor_profile",
           "fc_factor_free", "fc_matrix_new", "fc_matrix_free", "fc_matrix_set_segments", "fc_solve_refined",
           "fc_solve_refined_device", "fc_solve_refined_many_device", "fc_assembly_new", "fc_assemble_device",
           "fc_assembly_info", "fc_assembly_free", "fc_lanczos_device", "fc_ldlt_numeric", "fc_ldlt_solve",
           "fc_ldlt_inertia", "fc_ldlt_export", "fc_ldlt_free", "fc_zldlt_numeric", "fc_zldlt_solve",
           "fc_lanczos_device_ldlt", "fc_symbolic_schedule_get", "fc_symbolic_schedule_set", "fc_tune",
           "fc_version"]

#: kernel kinds of the numeric phase (FC_KIND_* in include/fc_chol.h)
KINDS = ["extend_l", "extend_u", "update", "potrf", "trsm", "syrk", "reduce", "small"]


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(_build.chol_library())
        ep = C.POINTER(_Error)
        vp = C.c_void_p
        L.fc_analyze.argtypes = [C.c_uint64, _u64p, _u64p, C.c_uint64, C.POINTER(vp), ep]
        L.fc_symbolic_info_get.argtypes = [vp, C.POINTER(_Info)]
        L.fc_symbolic_export.argtypes = [vp] + [vp] * 8
        L.fc_symbolic_free.argtypes = [vp]
        L.fc_symbolic_free.restype = None
        L.fc_symbolic_schedule_get.argtypes = [vp, C.POINTER(_Schedule)]
        L.fc_symbolic_schedule_set.argtypes = [vp, C.POINTER(_Schedule), ep]
        L.fc_tune.argtypes = [vp, vp, C.c_uint64, C.POINTER(_TuneOpts), C.POINTER(_TuneReport), ep]
        L.fc_factor_numeric.argtypes = [vp, _f64p, C.c_uint64, C.POINTER(_Opts), C.POINTER(vp), ep]
        L.fc_factor_numeric_device.argtypes = [vp, vp, C.c_uint64, C.POINTER(_Opts), C.POINTER(vp), ep]
        L.fc_refactor.argtypes = [vp, _f64p, C.c_uint64, ep]
        L.fc_refactor_device.argtypes = [vp, vp, C.c_uint64, ep]
        L.fc_factor_profile.argtypes = [vp, C.c_int32] + [C.POINTER(C.c_double)] * 4
        L.fc_entries.argtypes = [vp]
        L.fc_entries.restype = C.c_uint64
        L.fc_solve.argtypes = [vp, _f64p, C.c_uint64, ep]
        L.fc_solve_many.argtypes = [vp, C.POINTER(C.c_void_p), C.c_uint64, C.c_uint64, ep]
        L.fc_solve_device.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, ep]
        L.fc_solve_device_noalloc.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.c_uint64, ep]
        L.fc_factor_export.argtypes = [vp, _f64p, C.c_uint64, ep]
        L.fc_factor_seconds.argtypes = [vp]
        L.fc_factor_seconds.restype = C.c_double
        L.fc_factor_precision.argtypes = [vp]
        L.fc_factor_precision.restype = C.c_int32
        L.fc_factor_free.argtypes = [vp]
        L.fc_factor_free.restype = None
        L.fc_matrix_new.argtypes = [C.c_uint64, _u64p, _u64p, _f64p, C.c_int32, C.POINTER(vp), ep]
        L.fc_matrix_free.argtypes = [vp]
        L.fc_matrix_free.restype = None
        L.fc_solve_refined.argtypes = [vp, vp, _f64p, _f64p, C.c_uint64, C.POINTER(_RefineOpts),
                                       C.POINTER(_RefineReport), ep]
        L.fc_solve_refined_device.argtypes = [vp, vp, vp, vp, C.c_uint64, C.POINTER(_RefineOpts),
                                              C.POINTER(_RefineReport), ep]
        L.fc_matrix_set_segments.argtypes = [vp, C.c_uint64, _u64p, ep]
        L.fc_solve_refined_many_device.argtypes = [vp, vp, vp, vp, C.c_uint64, C.c_uint64, C.POINTER(_RefineOpts),
                                                   C.POINTER(_RefineReport), _f64p, ep]
        L.fc_lanczos_device.argtypes = [vp, vp, C.c_double, C.POINTER(_LanczosOpts), _f64p, vp,
                                         C.POINTER(_LanczosReport), ep]
        L.fc_lanczos_device_ldlt.argtypes = [vp, vp, C.c_double, C.POINTER(_LanczosOpts), _f64p, vp,
                                         C.POINTER(_LanczosReport), ep]
        L.fc_ldlt_numeric.argtypes = [vp, _f64p, C.c_uint64, C.c_int32, C.POINTER(vp), ep]
        L.fc_ldlt_solve.argtypes = [vp, _f64p, C.c_uint64, ep]
        L.fc_ldlt_inertia.argtypes = [vp] + [C.POINTER(C.c_uint64)] * 3
        L.fc_ldlt_export.argtypes = [vp, _f64p, C.c_uint64, ep]
        L.fc_ldlt_free.argtypes = [vp]
        L.fc_zldlt_numeric.argtypes = [vp, _f64p, C.c_uint64, C.c_int32, C.POINTER(vp), ep]
        L.fc_zldlt_solve.argtypes = [vp, _f64p, C.c_uint64, ep]
        L.fc_ldlt_free.restype = None
        L.fc_version.restype = C.c_char_p
        _lib = L
    return _lib


def _raise(status: int, err: _Error):
    if status == _OK:
        return
    cls = _KINDS.get(status, Internal)
    msg = err.message.decode(errors="replace")
    raise cls(msg or cls.kind, column=int(err.column), entries=int(err.entries), limit=int(err.limit))


def _raise_status(status: int):
    if status != _OK:
        cls = _KINDS.get(status, Internal)
        raise cls(cls.kind)


def _precision(p) -> int:
    return PRECISIONS[p] if isinstance(p, str) else int(p)


def _u64(a):
    a = np.asarray(a)
    return np.ascontiguousarray(a, dtype=np.uint64) if a.size else np.zeros(1, np.uint64)


@dataclass
class UpperColumns:
    """The upper triangle by columns: column k holds A(i, k) for i <= k, rows sorted
    (cholesky.rs:99-108)."""

    n: int
    ptr: np.ndarray
    idx: np.ndarray
    val: np.ndarray

    @staticmethod
    def from_symmetric(n: int, ptr, idx, val) -> "UpperColumns":
        """From a symmetric matrix by full rows: row k's entries with column <= k are column k
        (cholesky.rs:112-144)."""
        ptr = np.asarray(ptr, np.int64)
        idx = np.asarray(idx, np.int64)
        val = np.asarray(val, np.float64)
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ptr))
        keep = idx <= rows
        counts = np.bincount(rows[keep], minlength=n)
        uptr = np.zeros(n + 1, np.int64)
        np.cumsum(counts, out=uptr[1:])
        return UpperColumns(n, uptr, np.ascontiguousarray(idx[keep]), np.ascontiguousarray(val[keep]))


class Symbolic:
    """Symbolic::new (cholesky.rs:1213-1378) on the host, through fc_analyze."""

    def __init__(self, a: UpperColumns, max_entries: int = 2 ** 63 - 1):
        L = lib()
        self.n = int(a.n)
        self._h = C.c_void_p()
        err = _Error()
        st = L.fc_analyze(self.n, _u64(a.ptr), _u64(a.idx), int(max_entries), C.byref(self._h), C.byref(err))
        if st != _OK:
            self._h = None
            _raise(st, err)
        self.nnz = int(a.ptr[-1]) if self.n else 0

    def __del__(self):
        # the library object itself, not the module's lib(): at interpreter exit module
        # globals may already be gone
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fc_symbolic_free(self._h)
            self._h = None

    def info(self) -> dict:
        i = _Info()
        lib().fc_symbolic_info_get(self._h, C.byref(i))
        out = {k: getattr(i, k) for k, _ in _Info._fields_ if k not in ("pad", "kind_flops", "kind_entries")}
        out["kind_flops"] = {KINDS[k]: i.kind_flops[k] for k in range(len(KINDS))}
        out["kind_entries"] = {KINDS[k]: i.kind_entries[k] for k in range(len(KINDS))}
        return out

    def schedule(self) -> dict:
        """The analysis's schedule: slab (panels a side-stream update covers), split targets."""
        c = _Schedule()
        _raise_status(lib().fc_symbolic_schedule_get(self._h, C.byref(c)))
        return _sched_dict(c)

    def set_schedule(self, **kw) -> None:
        """Rebuild the launch program for a new schedule (keys of schedule(); the others keep
        their values). Factors made afterwards use it."""
        cur = self.schedule()
        unknown = set(kw) - set(cur)
        if unknown:
            raise ValueError(f"unknown schedule keys {sorted(unknown)}")
        cur.update(kw)
        c = _Schedule(**cur)
        err = _Error()
        _raise(lib().fc_symbolic_schedule_set(self._h, C.byref(c), C.byref(err)), err)

    def tune(self, device_values: int, nnz: int | None = None, precision: str | int = "mixed", device: int = 0,
             repeats: int = 0, stream: int | None = None) -> dict:
        """Time candidate schedules on this matrix (upper values in device memory) and install
        the fastest; returns the report (chosen, seconds_start, seconds_chosen, tried)."""
        o = _TuneOpts(_precision(precision), int(device), C.c_void_p(stream) if stream else None, int(repeats), 0)
        r = _TuneReport()
        err = _Error()
        _raise(lib().fc_tune(self._h, C.c_void_p(device_values), self.nnz if nnz is None else int(nnz), C.byref(o),
                             C.byref(r), C.byref(err)), err)
        return dict(chosen=_sched_dict(r.chosen), seconds_start=r.seconds_start, seconds_chosen=r.seconds_chosen,
                    tried=[dict(_sched_dict(r.tried[q]), seconds=r.tried_seconds[q]) for q in range(r.candidates)])

    def export(self) -> dict:
        """The reference's symbolic outputs: perm, first, rptr, rows, uptr, updates (k, p1, p2),
        tasks ([start, end)) and top."""
        i = self.info()
        n, ns = i["n"], i["supernodes"]
        out = dict(perm=np.zeros(max(n, 1), np.uint64), first=np.zeros(ns + 1, np.uint64),
                   rptr=np.zeros(ns + 1, np.uint64), rows=np.zeros(max(i["rows"], 1), np.uint64),
                   uptr=np.zeros(ns + 1, np.uint64), updates=np.zeros(max(3 * i["updates"], 1), np.uint64),
                   tasks=np.zeros(max(2 * i["tasks"], 1), np.uint64), top=np.zeros(max(i["top"], 1), np.uint64))
        keys = ["perm", "first", "rptr", "rows", "uptr", "updates", "tasks", "top"]
        lib().fc_symbolic_export(self._h, *[out[k].ctypes.data for k in keys])
        res = {k: out[k].astype(np.int64) for k in keys}
        res["perm"] = res["perm"][:n]
        res["rows"] = res["rows"][: i["rows"]]
        res["updates"] = res["updates"][: 3 * i["updates"]].reshape(-1, 3)
        res["tasks"] = res["tasks"][: 2 * i["tasks"]].reshape(-1, 2)
        res["top"] = res["top"][: i["top"]]
        res["entries"] = i["entries"]
        return res


class Matrix:
    """The matrix by full rows on the device, for FP64 residuals (Csr, sparse.rs)."""

    def __init__(self, n: int, ptr, idx, val, device: int = 0):
        self.n = int(n)
        self._h = C.c_void_p()
        err = _Error()
        st = lib().fc_matrix_new(self.n, _u64(ptr), _u64(idx), np.ascontiguousarray(val, np.float64), int(device),
                                 C.byref(self._h), C.byref(err))
        if st != _OK:
            self._h = None
            _raise(st, err)

    def set_segments(self, offsets) -> None:
        """Rows [offsets[s], offsets[s+1]) are independent models of a batch: refinement
        measures each one's backward error with its own norms."""
        off = _u64(offsets)
        err = _Error()
        _raise(lib().fc_matrix_set_segments(self._h, len(off) - 1, off, C.byref(err)), err)

    def __del__(self):
        # the library object itself, not the module's lib(): at interpreter exit module
        # globals may already be gone
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fc_matrix_free(self._h)
            self._h = None


class Factor:
    """A Cholesky factor on the GPU (Factor, cholesky.rs:1532-1570)."""

    def __init__(self, symbolic: Symbolic, val=None, precision: str | int = "fp64", device: int = 0,
                 bf16_terms: int = 0, cancel: C.c_int32 | None = None, fallback: int = 0,
                 use_graph: int = 0, profile: tuple = (), device_values: int | None = None,
                 stream: int | None = None):
        """Factor the matrix with values `val` (host array) or `device_values` (a device
        pointer to the upper columns' values, nnz doubles). bf16_terms (mixed form): 1 or 3
        BF16 products per operand pair, 0 for the library's default (3). use_graph: bit 0
        replays the factorization from a hipGraph, bit 1 the solves. fallback (mixed, fp32):
        0 refactors in the next precision when a pivot fails in the reduced one (mixed -> fp32
        -> fp64) and raises NotPositiveDefinite only when FP64 fails too; 1 raises at once.
        ``precision_used()`` tells which precision the factor holds."""
        L = lib()
        self.symbolic = symbolic
        self.n = symbolic.n
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self._cancel = cancel
        mask = 0
        for k in profile:
            mask |= 1 << (KINDS.index(k) if isinstance(k, str) else int(k))
        opts = _Opts(self.precision, int(device), C.c_void_p(stream) if stream else None,
                     C.pointer(cancel) if cancel is not None else None, int(bf16_terms), int(use_graph),
                     int(fallback), 0, mask, 0)
        self._h = C.c_void_p()
        err = _Error()
        if device_values is not None:
            st = L.fc_factor_numeric_device(symbolic._h, C.c_void_p(device_values), symbolic.nnz, C.byref(opts),
                                            C.byref(self._h), C.byref(err))
        else:
            val = np.ascontiguousarray(val, np.float64)
            st = L.fc_factor_numeric(symbolic._h, val if val.size else np.zeros(1), val.size, C.byref(opts),
                                     C.byref(self._h), C.byref(err))
        if st != _OK:
            self._h = None
            _raise(st, err)

    @staticmethod
    def new(a: UpperColumns, max_entries: int = 2 ** 63 - 1, **kw) -> "Factor":
        """Factor::new (cholesky.rs:1545-1547)."""
        return Factor(Symbolic(a, max_entries), a.val, **kw)

    def __del__(self):
        # the library object itself, not the module's lib(): at interpreter exit module
        # globals may already be gone
        if getattr(self, "_h", None) and _lib is not None:
            _lib.fc_factor_free(self._h)
            self._h = None

    def refactor(self, val) -> None:
        err = _Error()
        val = np.ascontiguousarray(val, np.float64)
        st = lib().fc_refactor(self._h, val if val.size else np.zeros(1), val.size, C.byref(err))
        _raise(st, err)

    def refactor_device(self, device_values: int) -> None:
        err = _Error()
        st = lib().fc_refactor_device(self._h, C.c_void_p(device_values), self.symbolic.nnz, C.byref(err))
        _raise(st, err)

    def profile(self, kind) -> dict:
        """Kernel kind `kind` in the last factorization: launches timed, their device seconds, and
        the kind's algorithmic flops and extend-add entries."""
        k = KINDS.index(kind) if isinstance(kind, str) else int(kind)
        v = [C.c_double() for _ in range(4)]
        lib().fc_factor_profile(self._h, k, *[C.byref(x) for x in v])
        return dict(launches=v[0].value, seconds=v[1].value, flops=v[2].value, entries=v[3].value)

    def entries(self) -> int:
        """Factor::entries (cholesky.rs:1557-1559)."""
        return int(lib().fc_entries(self._h))

    def seconds(self) -> float:
        return float(lib().fc_factor_seconds(self._h))

    def precision_used(self) -> str:
        """The precision of the last successful factorization ("fp64", "mixed", "fp32"): the one
        asked for, or the one a fallback reached."""
        p = int(lib().fc_factor_precision(self._h))
        return {v: k for k, v in PRECISIONS.items()}[p]

    def solve(self, b) -> np.ndarray:
        """Factor::solve (cholesky.rs:1562-1564); returns x (b is not modified)."""
        x = np.array(b, dtype=np.float64, copy=True)
        err = _Error()
        st = lib().fc_solve(self._h, x if x.size else np.zeros(1), x.shape[0], C.byref(err))
        _raise(st, err)
        return x

    def solve_many(self, bs) -> list:
        """Factor::solve_many (cholesky.rs:1568-1570): each the same bit for bit as solve."""
        xs = [np.array(b, dtype=np.float64, copy=True) for b in bs]
        n = xs[0].shape[0] if xs else self.n
        ptrs = (C.c_void_p * max(len(xs), 1))(*[x.ctypes.data for x in xs])
        err = _Error()
        st = lib().fc_solve_many(self._h, ptrs, len(xs), n, C.byref(err))
        _raise(st, err)
        return xs

    def solve_device(self, b_ptr: int, nrhs: int, ldb: int | None = None) -> None:
        """Solve in place a column-major block of nrhs right-hand sides in device memory."""
        err = _Error()
        st = lib().fc_solve_device(self._h, C.c_void_p(b_ptr), self.n, nrhs, ldb or self.n, C.byref(err))
        _raise(st, err)

    def solve_device_noalloc(self, b_ptr: int, nrhs: int, ldb: int | None = None) -> None:
        """solve_device without any device allocation: the work vectors live in the library's
        fixed per-device array (fc_solve_device_noalloc); raises TooLarge when one right-hand
        side's work does not fit it."""
        err = _Error()
        st = lib().fc_solve_device_noalloc(self._h, C.c_void_p(b_ptr), self.n, nrhs, ldb or self.n, C.byref(err))
        _raise(st, err)

    def values(self) -> np.ndarray:
        """The factor's values in the reference layout (trapezoids, cholesky.rs:393-420)."""
        e = self.entries()
        out = np.zeros(max(e, 1))
        err = _Error()
        st = lib().fc_factor_export(self._h, out, e, C.byref(err))
        _raise(st, err)
        return out[:e]

    def solve_refined(self, matrix: Matrix, b, method: str = "auto", max_iter: int = 50,
                      target_berr: float = 1e-12):
        """x with A x = b to normwise backward error <= target_berr (FP64 residuals on the
        device). Returns (x, report)."""
        m = {"refine": 0, "cg": 1, "auto": 2}[method]
        opts = _RefineOpts(m, int(max_iter), float(target_berr))
        rep = _RefineReport()
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros_like(b)
        err = _Error()
        st = lib().fc_solve_refined(self._h, matrix._h, b, x, b.shape[0], C.byref(opts), C.byref(rep), C.byref(err))
        _raise(st, err)
        report = dict(iterations=rep.iterations, converged=bool(rep.converged), berr=rep.berr, seconds=rep.seconds,
                      berr_history=list(rep.berr_history[: min(rep.iterations + 1, 64)]))
        return x, report


    def solve_refined_device(self, matrix: Matrix, b_ptr: int, x_ptr: int, method: str = "auto",
                             max_iter: int = 50, target_berr: float = 1e-12) -> dict:
        """Refined solve with b and x in device memory; returns the report."""
        m = {"refine": 0, "cg": 1, "auto": 2}[method]
        opts = _RefineOpts(m, int(max_iter), float(target_berr))
        rep = _RefineReport()
        err = _Error()
        st = lib().fc_solve_refined_device(self._h, matrix._h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), self.n,
                                           C.byref(opts), C.byref(rep), C.byref(err))
        _raise(st, err)
        return _report(rep)

    def lanczos_device(self, mass: "Matrix", shift: float, count: int, lo: float = -np.inf, hi: float = np.inf,
                       block: int = 6, max_basis: int | None = None, tolerance: float = 1e-9, vectors: bool = True):
        """The lowest `count` modes of K phi = lambda M phi in [lo, hi] by the shifted block Lanczos
        method (lanczos, eigen.rs:281-466) on the device, this factor being that of K - shift M and
        `mass` M by rows. Returns dict(values, vectors (a torch tensor n x found, M-orthonormal, or
        None), negative, basis, seconds)."""
        import torch

        if max_basis is None:
            # the reference's cap (Limits::MAX_LANCZOS_VECTORS = 3000, lib.rs:114), and the basis
            # and its M products in at most a quarter of the free device memory
            free, _ = torch.cuda.mem_get_info()
            max_basis = max(min(3000, self.n, int(free / 4 / (16 * max(self.n, 1)))), min(block, self.n))
        opts = _LanczosOpts(int(count), float(lo), float(hi), int(block), 0, int(max_basis), float(tolerance))
        vals = np.zeros(max(int(count), 1))
        vec = torch.empty((max(int(count), 1), max(self.n, 1)), dtype=torch.float64, device="cuda") if vectors else None
        rep = _LanczosReport()
        err = _Error()
        st = lib().fc_lanczos_device(self._h, mass._h, float(shift), C.byref(opts), vals,
                                     C.c_void_p(vec.data_ptr()) if vec is not None else None, C.byref(rep),
                                     C.byref(err))
        _raise(st, err)
        found = int(rep.found)
        return dict(values=vals[:found].copy(), vectors=vec[:found].T if vec is not None else None,
                    negative=np.array(rep.negative[:min(int(rep.negative_count), 16)]), basis=int(rep.basis),
                    seconds=float(rep.seconds))

    def solve_refined_many_device(self, matrix: Matrix, b_ptr: int, x_ptr: int, nrhs: int, method: str = "auto",
                                  max_iter: int = 50, target_berr: float = 1e-12):
        """Refined solves of nrhs columns (n x nrhs, column-major, device memory); returns the
        report (the worst column) and each column's backward error."""
        m = {"refine": 0, "cg": 1, "auto": 2}[method]
        opts = _RefineOpts(m, int(max_iter), float(target_berr))
        rep = _RefineReport()
        err = _Error()
        berr = np.zeros(max(int(nrhs), 1))
        st = lib().fc_solve_refined_many_device(self._h, matrix._h, C.c_void_p(b_ptr), C.c_void_p(x_ptr), self.n,
                                                int(nrhs), C.byref(opts), C.byref(rep), berr, C.byref(err))
        _raise(st, err)
        return _report(rep), berr[: int(nrhs)]


def _report(rep) -> dict:
    return dict(iterations=rep.iterations, converged=bool(rep.converged), berr=rep.berr, seconds=rep.seconds,
                berr_history=list(rep.berr_history[: min(rep.iterations + 1, 64)]))


def version() -> str:
    return lib().fc_version().decode()


class LdltFactor:
    """LdltFactor (crates/solver/src/ldlt.rs:362-445) on the GPU: P A P^T = L D L^T with bounded
    Bunch-Kaufman pivoting inside supernodes, on a Cholesky analysis (``Symbolic``): ``solve``,
    ``inertia``, ``two_by_two``, ``values`` (the reference's trapezoids). The factor and the
    solutions are the reference's bit for bit. Raises ``Singular`` naming the original column."""

    def __init__(self, symbolic: Symbolic, val, device: int = 0):
        L = lib()
        self.symbolic = symbolic
        self.n = symbolic.n
        self._L = L
        self._h = C.c_void_p()
        err = _Error()
        val = np.ascontiguousarray(val, np.float64)
        st = L.fc_ldlt_numeric(symbolic._h, val if val.size else np.zeros(1), val.size, int(device), C.byref(self._h),
                               C.byref(err))
        if st != _OK:
            self._h = None
            _raise(st, err)

    @staticmethod
    def new(a: UpperColumns, max_entries: int = 2 ** 63 - 1, **kw) -> "LdltFactor":
        """LdltFactor::new (ldlt.rs:363-365)."""
        return LdltFactor(Symbolic(a, max_entries), a.val, **kw)

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.fc_ldlt_free(self._h)
            self._h = None

    def solve(self, b) -> np.ndarray:
        x = np.array(b, dtype=np.float64, copy=True)
        err = _Error()
        st = lib().fc_ldlt_solve(self._h, x if x.size else np.zeros(1), x.shape[0], C.byref(err))
        _raise(st, err)
        return x

    def solve_many(self, bs) -> list:
        """LdltFactor::solve_many (ldlt.rs:410-412): each the same bit for bit as solve; a wrong
        length anywhere fails before any is solved."""
        xs = [np.array(b, dtype=np.float64, copy=True) for b in bs]
        if any(x.shape[0] != self.n for x in xs):
            raise Internal("right-hand side length")
        return [self.solve(x) for x in xs]

    def _counts(self):
        v = [C.c_uint64() for _ in range(3)]
        lib().fc_ldlt_inertia(self._h, *[C.byref(x) for x in v])
        return [int(x.value) for x in v]

    def lanczos_device(self, mass: "Matrix", shift: float, count: int, lo: float = -np.inf, hi: float = np.inf,
                       block: int = 6, max_basis: int | None = None, tolerance: float = 1e-9):
        """Factor.lanczos_device with this LDL^T factor of K - shift M: a shift between modes
        (indefinite K - shift M; the modes below it come back in ``negative``). Returns
        dict(values, negative, basis)."""
        import torch

        if max_basis is None:
            free, _ = torch.cuda.mem_get_info()
            max_basis = max(min(3000, self.n, int(free / 4 / (16 * max(self.n, 1)))), min(block, self.n))
        opts = _LanczosOpts(int(count), float(lo), float(hi), int(block), 0, int(max_basis), float(tolerance))
        vals = np.zeros(max(int(count), 1))
        rep = _LanczosReport()
        err = _Error()
        st = lib().fc_lanczos_device_ldlt(self._h, mass._h, float(shift), C.byref(opts), vals, None, C.byref(rep),
                                          C.byref(err))
        _raise(st, err)
        found = int(rep.found)
        return dict(values=vals[:found].copy(), negative=np.array(rep.negative[:min(int(rep.negative_count), 16)]),
                    basis=int(rep.basis))

    def inertia(self) -> tuple:
        """(positive, negative) eigenvalue counts (Sylvester's law on D, ldlt.rs:412-445)."""
        p, q, _ = self._counts()
        return p, q

    def two_by_two(self) -> int:
        return self._counts()[2]

    def values(self) -> np.ndarray:
        e = self.symbolic.info()["entries"]
        out = np.zeros(max(e, 1))
        err = _Error()
        st = lib().fc_ldlt_export(self._h, out, e, C.byref(err))
        _raise(st, err)
        return out[:e]


class ZLdltFactor:
    """LdltFactor::<Complex64> (crates/solver/src/ldlt.rs over scalar.rs's Complex64) on the GPU:
    the complex-symmetric A = A^T factored as P A P^T = L D L^T with the real path's pivoting,
    the modulus for its tests and num-complex's arithmetic (transposes, never conjugates), on a
    Cholesky analysis of the pattern: ``solve``, ``solve_many``, ``two_by_two``, ``values``.
    Raises ``Singular`` naming the original column."""

    def __init__(self, symbolic: Symbolic, val, device: int = 0):
        L = lib()
        self.symbolic = symbolic
        self.n = symbolic.n
        self._L = L
        self._h = C.c_void_p()
        err = _Error()
        pairs = np.ascontiguousarray(np.asarray(val, np.complex128)).view(np.float64)
        st = L.fc_zldlt_numeric(symbolic._h, pairs if pairs.size else np.zeros(2), pairs.size // 2, int(device),
                                C.byref(self._h), C.byref(err))
        if st != _OK:
            self._h = None
            _raise(st, err)

    @staticmethod
    def new(n: int, ptr, idx, val, max_entries: int = 2 ** 63 - 1, **kw) -> "ZLdltFactor":
        """LdltFactor::<Complex64>::new (ldlt.rs:374-376) from the upper columns (ptr, idx) and
        their complex values."""
        return ZLdltFactor(Symbolic(UpperColumns(n, np.asarray(ptr, np.int64), np.asarray(idx, np.int64),
                                                np.zeros(len(idx))), max_entries), val, **kw)

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.fc_ldlt_free(self._h)
            self._h = None

    def solve(self, b) -> np.ndarray:
        x = np.array(b, dtype=np.complex128, copy=True)
        err = _Error()
        st = lib().fc_zldlt_solve(self._h, x.view(np.float64) if x.size else np.zeros(2), x.shape[0], C.byref(err))
        _raise(st, err)
        return x

    def solve_many(self, bs) -> list:
        xs = [np.array(b, dtype=np.complex128, copy=True) for b in bs]
        if any(x.shape[0] != self.n for x in xs):
            raise Internal("right-hand side length")
        return [self.solve(x) for x in xs]

    def two_by_two(self) -> int:
        v = [C.c_uint64() for _ in range(3)]
        lib().fc_ldlt_inertia(self._h, *[C.byref(x) for x in v])
        return int(v[2].value)

    def values(self) -> np.ndarray:
        e = self.symbolic.info()["entries"]
        out = np.zeros(2 * max(e, 1))
        err = _Error()
        st = lib().fc_ldlt_export(self._h, out, 2 * e, C.byref(err))
        _raise(st, err)
        return out[:2 * e].view(np.complex128)


class Symmetric:
    """Symmetric (crates/solver/src/ldlt.rs:453-560): a factored symmetric matrix, Cholesky or
    LDL^T. ``kind`` is "definite" (Cholesky, NotPositiveDefinite otherwise), "indefinite" (LDL^T)
    or "auto" (one analysis; Cholesky, and LDL^T on the same analysis when the matrix is not
    positive definite). Both factors run on the device."""

    def __init__(self, a: UpperColumns, max_entries: int = 2 ** 63 - 1, kind: str = "auto", device: int = 0):
        if kind not in ("definite", "indefinite", "auto"):
            raise InvalidArgument(f"unknown kind {kind!r}")
        self.n = a.n
        if kind == "indefinite":
            self.factor = LdltFactor.new(a, max_entries, device=device)
        elif kind == "definite":
            self.factor = Factor.new(a, max_entries, device=device)
        else:
            symbolic = Symbolic(a, max_entries)
            try:
                self.factor = Factor(symbolic, a.val, device=device)
            except NotPositiveDefinite:
                self.factor = LdltFactor(symbolic, a.val, device=device)

    @property
    def indefinite(self) -> bool:
        return isinstance(self.factor, LdltFactor)

    def inertia(self) -> tuple:
        """A Cholesky factor's matrix is positive definite (ldlt.rs:502-508)."""
        return self.factor.inertia() if self.indefinite else (self.n, 0)

    def solve(self, b) -> np.ndarray:
        return self.factor.solve(b)

    def solve_many(self, bs) -> list:
        return self.factor.solve_many(bs)

    def solve_refined(self, n: int, ptr, idx, val, b, steps: int) -> np.ndarray:
        """solve_refined (ldlt.rs:526-558): solve, then while the residual's largest magnitude
        shrinks, at most ``steps`` times, x += solve(b - A x) (iterative refinement in working
        precision, Higham 2002, sec. 12.2). A is the factored matrix by full rows; its product
        sums each row in column order, as the reference's Csr::mul. Cholesky solutions are
        returned as solved."""
        rhs = np.array(b, dtype=np.float64, copy=True)
        x = self.solve(rhs)
        if not self.indefinite:
            return x
        ptr = np.asarray(ptr, np.int64)
        idx = np.asarray(idx, np.int64)
        val = np.asarray(val, np.float64)

        def norm(v):  # the largest magnitude, NaN ignored (f64::max)
            return float(np.fmax.reduce(np.abs(v), initial=0.0))

        def residual(v):
            ax = np.zeros(n)
            prod = val * v[idx]
            # row sums in column order: one term per pass over the rows' k-th entries
            lens = np.diff(ptr)
            for k in range(int(lens.max()) if n else 0):
                sel = ptr[:-1] + k < ptr[1:]
                ax[sel] = ax[sel] + prod[ptr[:-1][sel] + k]
            return rhs - ax

        r = residual(x)
        size = norm(r)
        for _ in range(steps):
            if not size > 0.0:
                break
            trial = x + self.solve(r)
            nxt = residual(trial)
            nsize = norm(nxt)
            if not nsize < size:
                break
            x, r, size = trial, nxt, nsize
        return x
