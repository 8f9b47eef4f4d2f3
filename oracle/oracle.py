"""ORACLE — test infrastructure only (tests/, smoke() and bench.py's CPU-baseline leg).

ctypes wrapper of ``liboracle.so``, the plain-C restatement of FieldCraft's supernodal Cholesky
path (crates/solver/src/cholesky.rs, ordering.rs, sparse.rs). The shipped library never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

OK, NOT_POSITIVE_DEFINITE, SINGULAR, TOO_LARGE, OUT_OF_MEMORY, INTERNAL, CANCELLED = range(7)
KIND_NAMES = {
    OK: "ok", NOT_POSITIVE_DEFINITE: "not_positive_definite", SINGULAR: "singular",
    TOO_LARGE: "too_large", OUT_OF_MEMORY: "out_of_memory", INTERNAL: "internal", CANCELLED: "cancelled",
}

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so from chol_oracle.c (make in this directory)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        i64 = C.c_int64
        L.orc_analyze.argtypes = [i64, _i64p, _i64p, i64, C.POINTER(C.c_void_p), C.POINTER(i64)]
        L.orc_analyze.restype = C.c_int
        L.orc_factor_new.argtypes = [i64, _i64p, _i64p, _f64p, i64, C.c_void_p, C.POINTER(C.c_void_p),
                                     C.POINTER(i64), C.POINTER(i64)]
        L.orc_factor_new.restype = C.c_int
        L.orc_factor_numeric.argtypes = [C.c_void_p, _f64p, C.c_void_p, C.POINTER(i64)]
        L.orc_factor_numeric.restype = C.c_int
        L.orc_free.argtypes = [C.c_void_p]
        for name in ["orc_n", "orc_supernodes", "orc_entries", "orc_row_count", "orc_update_count",
                     "orc_task_count", "orc_top_count"]:
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = i64
        L.orc_by_blocks.argtypes = [C.c_void_p]
        L.orc_by_blocks.restype = C.c_int
        L.orc_export_symbolic.argtypes = [C.c_void_p] + [_i64p] * 10
        L.orc_export_values.argtypes = [C.c_void_p, _f64p]
        L.orc_export_values.restype = C.c_int
        L.orc_solve.argtypes = [C.c_void_p, _f64p, i64]
        L.orc_solve.restype = C.c_int
        L.orc_symbolic.argtypes = [i64, _i64p, _i64p, C.c_int, _i64p, _i64p]
        L.orc_postorder.argtypes = [i64, _i64p, _i64p]
        L.orc_from_symmetric.argtypes = [i64, _i64p, _i64p, _f64p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_from_symmetric.restype = i64
        L.orc_csr_mul.argtypes = [i64, _i64p, _i64p, _f64p, _f64p, _f64p]
        L.orc_nested_dissection.argtypes = [i64, _f64p, _i64p, _i64p, _i64p]
        L.orc_graph_dissection.argtypes = [i64, _i64p, _i64p, _i64p]
        L.orc_fill.argtypes = [i64, _i64p, _i64p, _i64p]
        L.orc_fill.restype = i64
        L.orc_fill_reducing_order.argtypes = [i64, _f64p, _i64p, _i64p, _i64p]
        L.orc_fill_reducing_order.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_threads.restype = C.c_int
        L.orc_set_parallel_inner.argtypes = [C.c_int]
        L.orc_ldlt_new.argtypes = [i64, _i64p, _i64p, _f64p, i64, C.POINTER(C.c_void_p), C.POINTER(i64),
                                   C.POINTER(i64)]
        L.orc_ldlt_new.restype = C.c_int
        L.orc_ldlt_solve.argtypes = [C.c_void_p, _f64p, i64]
        L.orc_ldlt_solve.restype = C.c_int
        L.orc_zldlt_new.argtypes = [i64, _i64p, _i64p, _f64p, i64, C.POINTER(C.c_void_p), C.POINTER(i64),
                                    C.POINTER(i64)]
        L.orc_zldlt_new.restype = C.c_int
        L.orc_zldlt_solve.argtypes = [C.c_void_p, _f64p, i64]
        L.orc_zldlt_solve.restype = C.c_int
        L.orc_zexport_values.argtypes = [C.c_void_p, _f64p]
        L.orc_zexport_values.restype = None
        L.orc_ldlt_inertia.argtypes = [C.c_void_p, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
        L.orc_solid_stiffness.argtypes = [C.c_int, _f64p, C.c_double, C.c_double, _f64p]
        L.orc_solid_stiffness.restype = C.c_int
        L.orc_tet10_mass.argtypes = [_f64p, C.c_double, _f64p]
        L.orc_tet10_mass.restype = C.c_int
        L.orc_assemble_upper.argtypes = [C.c_int, i64, _f64p, i64, _i64p, _f64p, i64, C.c_double, C.c_double,
                                         C.c_double, C.c_int, i64, _i64p, _i64p, _i64p, _f64p]
        L.orc_assemble_upper.restype = i64
        _lib = L
    return _lib


def _a(x, dt=np.int64):
    return np.ascontiguousarray(x, dtype=dt)


class OracleError(Exception):
    def __init__(self, kind: int, column: int = -1, entries: int = -1):
        self.kind, self.column, self.entries = kind, column, entries
        super().__init__(f"{KIND_NAMES.get(kind, kind)} column={column} entries={entries}")


def from_symmetric(n, rptr, ridx, rval):
    """UpperColumns::from_symmetric (cholesky.rs:112-144)."""
    L = lib()
    rptr, ridx, rval = _a(rptr), _a(ridx), _a(rval, np.float64)
    total = L.orc_from_symmetric(n, rptr, ridx, rval, None, None, None)
    ptr = np.zeros(n + 1, np.int64)
    idx = np.zeros(max(total, 1), np.int64)
    val = np.zeros(max(total, 1), np.float64)
    L.orc_from_symmetric(n, rptr, ridx, rval, ptr.ctypes.data, idx.ctypes.data, val.ctypes.data)
    return ptr, idx[:total].copy(), val[:total].copy()


ELEMENT_KINDS = {"hex8": 0, "tet10": 1}


def solid_stiffness(kind: str, xyz, young: float, poisson: float):
    """Solid::stiffness (fem element.rs:416-422): 3n x 3n, node by node (x, y, z); None when the
    element is degenerate."""
    k = ELEMENT_KINDS[kind]
    nn = 8 if k == 0 else 10
    ke = np.zeros((3 * nn, 3 * nn))
    ok = lib().orc_solid_stiffness(k, _a(xyz, np.float64).reshape(-1), float(young), float(poisson), ke.reshape(-1))
    return ke if ok else None


def tet10_mass(xyz, density: float):
    """Solid::mass of a tet10 (fem element.rs:427-441): 10 x 10 per displacement component."""
    me = np.zeros((10, 10))
    ok = lib().orc_tet10_mass(_a(xyz, np.float64).reshape(-1), float(density), me.reshape(-1))
    return me if ok else None


class DegenerateElement(Exception):
    def __init__(self, element: int):
        self.element = element
        super().__init__(f"element {element} is inverted or degenerate")


def assemble_upper(kind: str, xyz, conn, young, poisson: float, eq_dof, uptr, uidx, *, density: float = 0.0,
                   what: str = "K", shift: float = 0.0):
    """Upper columns of K_ff, M_ff or K_ff - shift*M_ff on the caller's pattern (uptr, uidx): the
    element matrices added in element order (analysis.rs:199-233, sparse.rs:199-243), reduced to
    the free equations (constraints.rs:820-906), K + (-shift) M (sparse.rs:81-87), and
    UpperColumns::from_symmetric (cholesky.rs:110-144)."""
    young = np.atleast_1d(np.asarray(young, np.float64))
    xyz, conn = _a(xyz, np.float64), _a(conn)
    uptr, uidx = _a(uptr), _a(uidx)
    out = np.zeros(max(int(uptr[-1]), 1))
    w = {"K": 0, "M": 1, "shifted": 2}[what]
    r = lib().orc_assemble_upper(ELEMENT_KINDS[kind], len(xyz), xyz.reshape(-1), len(conn), conn.reshape(-1), young,
                                 len(young), float(poisson), float(density), float(shift), w, len(eq_dof), _a(eq_dof),
                                 uptr, uidx, out)
    if r >= 0:
        raise DegenerateElement(int(r))
    if r == -2:
        raise ValueError("the assembled pattern differs from the given upper columns")
    if r == -3:
        raise ValueError("bad element kind, mass request or material")
    return out[:int(uptr[-1])]


def csr_mul(n, ptr, idx, val, x):
    """Csr::mul (sparse.rs:42-50)."""
    y = np.zeros(n, np.float64)
    lib().orc_csr_mul(n, _a(ptr), _a(idx), _a(val, np.float64), _a(x, np.float64), y)
    return y


def symbolic(n, ptr, idx, by_columns=False):
    parent = np.zeros(max(n, 1), np.int64)
    counts = np.zeros(max(n, 1), np.int64)
    lib().orc_symbolic(n, _a(ptr), _a(idx), int(by_columns), parent, counts)
    return parent[:n], counts[:n]


def postorder(parent):
    parent = _a(parent)
    n = len(parent)
    post = np.zeros(max(n, 1), np.int64)
    lib().orc_postorder(n, parent if n else np.zeros(1, np.int64), post)
    return post[:n]


def _graph(adj_ptr, adj_idx):
    return _a(adj_ptr), _a(adj_idx) if len(adj_idx) else np.zeros(1, np.int64)


def nested_dissection(xyz, adj_ptr, adj_idx):
    n = len(adj_ptr) - 1
    order = np.zeros(max(n, 1), np.int64)
    p, i = _graph(adj_ptr, adj_idx)
    lib().orc_nested_dissection(n, _a(np.reshape(xyz, -1), np.float64) if n else np.zeros(3), p, i, order)
    return order[:n]


def graph_dissection(adj_ptr, adj_idx):
    n = len(adj_ptr) - 1
    order = np.zeros(max(n, 1), np.int64)
    p, i = _graph(adj_ptr, adj_idx)
    lib().orc_graph_dissection(n, p, i, order)
    return order[:n]


def fill(adj_ptr, adj_idx, order):
    n = len(adj_ptr) - 1
    p, i = _graph(adj_ptr, adj_idx)
    return int(lib().orc_fill(n, p, i, _a(order) if n else np.zeros(1, np.int64)))


def fill_reducing_order(xyz, adj_ptr, adj_idx):
    n = len(adj_ptr) - 1
    order = np.zeros(max(n, 1), np.int64)
    p, i = _graph(adj_ptr, adj_idx)
    graph = lib().orc_fill_reducing_order(n, _a(np.reshape(xyz, -1), np.float64) if n else np.zeros(3), p, i, order)
    return order[:n], bool(graph)


def set_threads(t: int) -> None:
    lib().orc_set_threads(int(t))


def threads() -> int:
    return int(lib().orc_threads())


class Factor:
    """The oracle's Factor (cholesky.rs:1532-1570) over upper columns (ptr, idx, val)."""

    def __init__(self, n, ptr, idx, val=None, max_entries=2**62, analyze_only=False, cancel=False):
        L = lib()
        self.n = int(n)
        self._h = C.c_void_p()
        self.ptr = _a(ptr)
        self.idx = _a(idx) if len(idx) else np.zeros(1, np.int64)
        col, ent = C.c_int64(-1), C.c_int64(-1)
        if analyze_only:
            e = L.orc_analyze(self.n, self.ptr, self.idx, int(max_entries), C.byref(self._h), C.byref(ent))
        else:
            v = _a(val, np.float64) if len(val) else np.zeros(1)
            flag = C.c_int(1 if cancel else 0)
            e = L.orc_factor_new(self.n, self.ptr, self.idx, v, int(max_entries), C.byref(flag),
                                 C.byref(self._h), C.byref(col), C.byref(ent))
        if e != OK:
            self._h = None
            raise OracleError(e, col.value, ent.value)

    def refactor(self, val):
        col = C.c_int64(-1)
        e = lib().orc_factor_numeric(self._h, _a(val, np.float64), None, C.byref(col))
        if e != OK:
            raise OracleError(e, col.value)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def entries(self) -> int:
        return int(lib().orc_entries(self._h))

    def supernodes(self) -> int:
        return int(lib().orc_supernodes(self._h))

    def by_blocks(self) -> bool:
        return bool(lib().orc_by_blocks(self._h))

    def symbolic(self) -> dict:
        L = lib()
        h = self._h
        n, ns = self.n, L.orc_supernodes(h)
        nr, nu = L.orc_row_count(h), L.orc_update_count(h)
        nt, ntop = L.orc_task_count(h), L.orc_top_count(h)
        out = dict(perm=np.zeros(max(n, 1), np.int64), parent0=np.zeros(max(n, 1), np.int64),
                   counts0=np.zeros(max(n, 1), np.int64), first=np.zeros(ns + 1, np.int64),
                   rptr=np.zeros(ns + 1, np.int64), rows=np.zeros(max(nr, 1), np.int64),
                   uptr=np.zeros(ns + 1, np.int64), updates=np.zeros(max(3 * nu, 1), np.int64),
                   tasks=np.zeros(max(2 * nt, 1), np.int64), top=np.zeros(max(ntop, 1), np.int64))
        L.orc_export_symbolic(h, out["perm"], out["parent0"], out["counts0"], out["first"], out["rptr"],
                              out["rows"], out["uptr"], out["updates"], out["tasks"], out["top"])
        out["perm"], out["parent0"], out["counts0"] = out["perm"][:n], out["parent0"][:n], out["counts0"][:n]
        out["rows"] = out["rows"][:nr]
        out["updates"] = out["updates"][: 3 * nu].reshape(-1, 3)
        out["tasks"] = out["tasks"][: 2 * nt].reshape(-1, 2)
        out["top"] = out["top"][:ntop]
        out["entries"] = self.entries()
        return out

    def values(self) -> np.ndarray:
        s = self.symbolic()
        m = np.diff(s["rptr"])
        w = np.diff(s["first"])
        total = int(np.sum(w * m - w * (w - 1) // 2))
        out = np.zeros(max(total, 1))
        if lib().orc_export_values(self._h, out) != OK:
            raise OracleError(INTERNAL)
        return out[:total]

    def solve(self, b):
        x = np.array(b, dtype=np.float64, copy=True)
        if x.shape[0] != self.n:
            raise OracleError(INTERNAL)
        if self.n == 0:
            return x
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        xt = np.ascontiguousarray(x.reshape(self.n, nrhs).T)
        if lib().orc_solve(self._h, xt, nrhs) != OK:
            raise OracleError(INTERNAL)
        return xt.T.reshape(x.shape).copy()


class LdltFactor:
    """LdltFactor::new (ldlt.rs:362-391): P A P^T = L D L^T with bounded Bunch-Kaufman pivoting
    inside supernodes, on the Cholesky symbolic analysis; solve (cholesky.rs:904-1037 with
    pivots), inertia and two_by_two (ldlt.rs:393-445). Raises OracleError(SINGULAR, column)."""

    def __init__(self, n, ptr, idx, val, max_entries=2**62):
        L = lib()
        self.n = n
        self._h = C.c_void_p()
        col, ent = C.c_int64(-1), C.c_int64(-1)
        val = _a(val, np.float64)
        st = L.orc_ldlt_new(n, _a(ptr), _a(idx), val if val.size else np.zeros(1), int(max_entries),
                            C.byref(self._h), C.byref(col), C.byref(ent))
        if st != OK:
            self._h = None
            raise OracleError(st, col.value, ent.value)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def solve(self, b):
        x = np.array(b, dtype=np.float64, copy=True)
        if self.n:
            lib().orc_ldlt_solve(self._h, x.reshape(-1, order="F") if x.ndim > 1 else x, 1 if x.ndim == 1 else x.shape[1])
        return x

    def inertia(self):
        p, q, t = C.c_int64(), C.c_int64(), C.c_int64()
        lib().orc_ldlt_inertia(self._h, C.byref(p), C.byref(q), C.byref(t))
        return int(p.value), int(q.value)

    def two_by_two(self) -> int:
        p, q, t = C.c_int64(), C.c_int64(), C.c_int64()
        lib().orc_ldlt_inertia(self._h, C.byref(p), C.byref(q), C.byref(t))
        return int(t.value)

    def values(self) -> np.ndarray:
        """Every supernode's trapezoid (L below the pivots, D on them), in supernode order."""
        L = lib()
        out = np.zeros(max(L.orc_entries(self._h), 1))
        L.orc_export_values(self._h, out)
        return out[:L.orc_entries(self._h)]


class ZLdltFactor:
    """LdltFactor::<Complex64>::new (ldlt.rs:362-412 over scalar.rs's Complex64): the
    complex-symmetric P A P^T = L D L^T (transposes, never conjugates) with the modulus for the
    pivot tests; solve and two_by_two. ``val`` is complex128. Raises OracleError(SINGULAR, column)."""

    def __init__(self, n, ptr, idx, val, max_entries=2**62):
        L = lib()
        self.n = n
        self._h = C.c_void_p()
        col, ent = C.c_int64(-1), C.c_int64(-1)
        pairs = np.ascontiguousarray(np.asarray(val, np.complex128)).view(np.float64)
        st = L.orc_zldlt_new(n, _a(ptr), _a(idx), pairs if pairs.size else np.zeros(2), int(max_entries),
                             C.byref(self._h), C.byref(col), C.byref(ent))
        if st != OK:
            self._h = None
            raise OracleError(st, col.value, ent.value)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_free(self._h)
            self._h = None

    def solve(self, b):
        x = np.array(b, dtype=np.complex128, copy=True)
        if self.n:
            lib().orc_zldlt_solve(self._h, x.view(np.float64), 1)
        return x

    def two_by_two(self) -> int:
        p, q, t = C.c_int64(), C.c_int64(), C.c_int64()
        lib().orc_ldlt_inertia(self._h, C.byref(p), C.byref(q), C.byref(t))
        return int(t.value)

    def values(self) -> np.ndarray:
        L = lib()
        e = L.orc_entries(self._h)
        out = np.zeros(2 * max(e, 1))
        L.orc_zexport_values(self._h, out)
        return out[:2 * e].view(np.complex128)
