"""Synthetic FEM inputs for the supernodal Cholesky, in the reference's own order of work.

A structured hex8 or tet10 block is meshed, its node graph built (sparse.rs:104-120), its nodes
ordered by the fill-reducing order (ordering.rs:102-110), the element matrices assembled in
label order (analysis.rs:199-233) and the free equations numbered node by node (constraints.rs:
755-766); ``Problem`` holds K_ff (and M_ff) by full rows in that numbering, the matrix the
reference's static analysis hands to ``Factor::new`` after ``UpperColumns::from_symmetric``
(analysis.rs:285-287). The heavy lifting is C++ (``csrc/problems.cpp``, built into
``libfc_problems.so``).

The named workloads follow the hot path's configurations:
  * ``cantilever_hex8``: 40x8x8 hex8 on 0.4 x 0.08 x 0.08 m, x=0 clamped, 1 kN tip load in -z
    split equally over the x=L face nodes (9 720 equations).
  * ``cube_hex8``: k^3 hex8 on the unit cube, z=0 clamped; uniform pressure on z=1 plus a
    seeded N(0,1) body-force field scaled to the same 2-norm (k=80: 1 574 640 equations).
  * ``cube_hex8_contrast``: the same pattern, element E in {210 GPa, 21 MPa} by a seeded
    Bernoulli(0.3), nu = 0.49.
  * ``random_model``: one of the batched models (hex8 or tet10, seeded size and material).
  * ``block_tet10``: a tet10 Kuhn block with consistent mass (shift-invert workload).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import _build

_lib = None
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        path = _build.problems_library()
        L = C.CDLL(path)
        i64 = C.c_int64
        L.fcp_block.argtypes = [C.c_int, i64, i64, i64, C.c_double, C.c_double, C.c_double, _f64p, i64, C.c_double,
                                C.c_int, C.c_double, C.c_uint64, C.c_double]
        L.fcp_block.restype = C.c_void_p
        L.fcp_free.argtypes = [C.c_void_p]
        for name in ["fcp_nodes", "fcp_elements", "fcp_neq", "fcp_nnz", "fcp_adjacency_nnz"]:
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = i64
        for name in ["fcp_nodes_per_element", "fcp_graph_chosen", "fcp_has_mass"]:
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = C.c_int
        L.fcp_get_mesh.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        L.fcp_get_csr.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.fcp_fill_reducing_order.argtypes = [i64, _f64p, _i64p, _i64p, _i64p]
        L.fcp_fill_reducing_order.restype = C.c_int
        L.fcp_nested_dissection.argtypes = [i64, _f64p, _i64p, _i64p, _i64p]
        L.fcp_graph_dissection.argtypes = [i64, _i64p, _i64p, _i64p]
        L.fcp_fill.argtypes = [i64, _i64p, _i64p, _i64p]
        L.fcp_fill.restype = i64
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


@dataclass
class Problem:
    """K_ff (and M_ff) of a structured block, by full rows in equation order."""

    kind: str
    cells: tuple
    n: int
    ptr: np.ndarray
    idx: np.ndarray
    kval: np.ndarray
    mval: np.ndarray | None
    xyz: np.ndarray
    conn: np.ndarray
    adj_ptr: np.ndarray
    adj_idx: np.ndarray
    order: np.ndarray
    eq_dof: np.ndarray
    graph_chosen: bool
    rhs: np.ndarray | None = None
    meta: dict = field(default_factory=dict)
    young: np.ndarray | None = None   # per element, or one value for all
    poisson: float = 0.3
    rho: float = 0.0

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    def upper(self, values: np.ndarray | None = None):
        """UpperColumns::from_symmetric (cholesky.rs:112-144): row k's entries with column <= k
        are column k of the upper triangle (rows stay sorted)."""
        vals = self.kval if values is None else values
        keep = self.idx <= np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.ptr))
        counts = np.add.reduceat(keep.astype(np.int64), self.ptr[:-1]) if self.n else np.zeros(0, np.int64)
        # reduceat misreads empty rows; every row here holds its diagonal
        uptr = np.zeros(self.n + 1, np.int64)
        np.cumsum(counts, out=uptr[1:])
        return uptr, np.ascontiguousarray(self.idx[keep]), np.ascontiguousarray(vals[keep])

    def mul(self, x: np.ndarray, values: np.ndarray | None = None) -> np.ndarray:
        """y = K x by rows (Csr::mul, sparse.rs:42-50)."""
        vals = self.kval if values is None else values
        prod = vals * x[self.idx] if x.ndim == 1 else vals[:, None] * x[self.idx]
        out = np.add.reduceat(prod, self.ptr[:-1], axis=0)
        return out

    def norm_inf(self, values: np.ndarray | None = None) -> float:
        vals = self.kval if values is None else values
        return float(np.max(np.add.reduceat(np.abs(vals), self.ptr[:-1])))


def block(kind: str, cells, size, young, poisson: float, clamp_face: int, jitter: float = 0.0, seed: int = 0,
          rho: float = 0.0) -> Problem:
    """Mesh, assemble, order and reduce a structured block (see module doc)."""
    L = lib()
    nx, ny, nz = (int(c) for c in cells)
    young = np.atleast_1d(np.asarray(young, dtype=np.float64))
    k = {"hex8": 0, "tet10": 1}[kind]
    h = L.fcp_block(k, nx, ny, nz, float(size[0]), float(size[1]), float(size[2]), np.ascontiguousarray(young),
                    len(young), float(poisson), int(clamp_face), float(jitter), int(seed) & (2**64 - 1), float(rho))
    if not h:
        raise ValueError("a degenerate element in the generated mesh")
    try:
        nn, ne, npe = L.fcp_nodes(h), L.fcp_elements(h), L.fcp_nodes_per_element(h)
        n, nnz, na = L.fcp_neq(h), L.fcp_nnz(h), L.fcp_adjacency_nnz(h)
        xyz = np.zeros((nn, 3))
        conn = np.zeros((ne, npe), np.int64)
        adj_ptr = np.zeros(nn + 1, np.int64)
        adj_idx = np.zeros(na, np.int64)
        order = np.zeros(nn, np.int64)
        eq_dof = np.zeros(n, np.int64)
        L.fcp_get_mesh(h, _p(xyz), _p(conn), _p(adj_ptr), _p(adj_idx), _p(order), _p(eq_dof))
        ptr = np.zeros(n + 1, np.int64)
        idx = np.zeros(nnz, np.int64)
        kval = np.zeros(nnz)
        mval = np.zeros(nnz) if L.fcp_has_mass(h) else None
        L.fcp_get_csr(h, _p(ptr), _p(idx), _p(kval), _p(mval))
        return Problem(kind, (nx, ny, nz), int(n), ptr, idx, kval, mval, xyz, conn, adj_ptr, adj_idx, order, eq_dof,
                       bool(L.fcp_graph_chosen(h)), young=young.copy(), poisson=float(poisson),
                       rho=float(rho) if mval is not None else 0.0)
    finally:
        L.fcp_free(h)


def fill_reducing_order(xyz, adj_ptr, adj_idx):
    n = len(adj_ptr) - 1
    order = np.zeros(max(n, 1), np.int64)
    g = lib().fcp_fill_reducing_order(n, np.ascontiguousarray(np.reshape(xyz, -1), np.float64) if n else np.zeros(3),
                                      np.ascontiguousarray(adj_ptr, np.int64),
                                      np.ascontiguousarray(adj_idx, np.int64) if len(adj_idx) else np.zeros(1, np.int64),
                                      order)
    return order[:n], bool(g)


def nested_dissection(xyz, adj_ptr, adj_idx):
    n = len(adj_ptr) - 1
    order = np.zeros(max(n, 1), np.int64)
    lib().fcp_nested_dissection(n, np.ascontiguousarray(np.reshape(xyz, -1), np.float64) if n else np.zeros(3),
                                np.ascontiguousarray(adj_ptr, np.int64),
                                np.ascontiguousarray(adj_idx, np.int64) if len(adj_idx) else np.zeros(1, np.int64), order)
    return order[:n]


def graph_dissection(adj_ptr, adj_idx):
    n = len(adj_ptr) - 1
    order = np.zeros(max(n, 1), np.int64)
    lib().fcp_graph_dissection(n, np.ascontiguousarray(adj_ptr, np.int64),
                               np.ascontiguousarray(adj_idx, np.int64) if len(adj_idx) else np.zeros(1, np.int64), order)
    return order[:n]


def fill(adj_ptr, adj_idx, order):
    n = len(adj_ptr) - 1
    return int(lib().fcp_fill(n, np.ascontiguousarray(adj_ptr, np.int64),
                              np.ascontiguousarray(adj_idx, np.int64) if len(adj_idx) else np.zeros(1, np.int64),
                              np.ascontiguousarray(order, np.int64) if n else np.zeros(1, np.int64)))


# ------------------------------------------------------------------------------------------
# Named workloads

E_STEEL = 210e9


def _dof_eq(p: Problem) -> np.ndarray:
    """Mesh dof -> equation (-1 where fixed)."""
    eq = -np.ones(3 * len(p.xyz), np.int64)
    eq[p.eq_dof] = np.arange(p.n, dtype=np.int64)
    return eq


def cantilever_hex8(nx: int = 40, ny: int = 8, nz: int = 8) -> Problem:
    """Hex8 cantilever 0.4 x 0.08 x 0.08 m (nodes 10 mm apart at 40x8x8), E = 210 GPa,
    nu = 0.3, x = 0 face clamped, (0, 0, -1000) N split equally over the x = L face nodes
    (model loads.rs:45-55, Distribution::Total)."""
    p = block("hex8", (nx, ny, nz), (0.4, 0.08, 0.08), E_STEEL, 0.3, clamp_face=0)
    eq = _dof_eq(p)
    tip = np.nonzero(p.xyz[:, 0] == p.xyz[:, 0].max())[0]
    f = np.zeros(p.n)
    f[eq[3 * tip + 2]] = -1000.0 / len(tip)
    p.rhs = f
    p.meta = dict(workload="cantilever_hex8", cells=[nx, ny, nz], young=E_STEEL, poisson=0.3)
    return p


def pressure_load(p: Problem, pressure: float = 1e6) -> np.ndarray:
    """Consistent nodal loads of a uniform pressure on the z = max face of a hex8 block:
    p h_x h_y x {1, 1/2, 1/4} at interior, edge and corner nodes, in -z."""
    eq = _dof_eq(p)
    z = p.xyz[:, 2]
    top = np.nonzero(z == z.max())[0]
    nx, ny, _ = p.cells
    hx = (p.xyz[:, 0].max() - p.xyz[:, 0].min()) / nx
    hy = (p.xyz[:, 1].max() - p.xyz[:, 1].min()) / ny
    x, y = p.xyz[top, 0], p.xyz[top, 1]
    wx = np.where((x == x.min()) | (x == x.max()), 0.5, 1.0)
    wy = np.where((y == y.min()) | (y == y.max()), 0.5, 1.0)
    f = np.zeros(p.n)
    f[eq[3 * top + 2]] = -pressure * hx * hy * wx * wy
    return f


def cube_rhs(p: Problem, seed: int = 20241) -> np.ndarray:
    """Uniform pressure on z = 1 plus a seeded N(0,1) body-force field scaled to its 2-norm."""
    fp = pressure_load(p)
    g = np.random.default_rng(seed).standard_normal(p.n)
    g *= np.linalg.norm(fp) / np.linalg.norm(g)
    return fp + g


def cube_hex8(k: int = 80, seed: int = 20241) -> Problem:
    """Hex8 k^3 on the unit cube, E = 210 GPa, nu = 0.3, z = 0 clamped."""
    p = block("hex8", (k, k, k), (1.0, 1.0, 1.0), E_STEEL, 0.3, clamp_face=4)
    p.rhs = cube_rhs(p, seed)
    p.meta = dict(workload=f"cube_hex8_{k}", cells=[k, k, k], young=E_STEEL, poisson=0.3)
    return p


def cube_hex8_contrast(k: int = 80, seed: int = 20242) -> Problem:
    """The cube with E in {210 GPa, 21 MPa} per element (Bernoulli p = 0.3) and nu = 0.49."""
    rng = np.random.default_rng(seed)
    young = np.where(rng.random(k ** 3) < 0.3, 21e6, E_STEEL)
    p = block("hex8", (k, k, k), (1.0, 1.0, 1.0), young, 0.49, clamp_face=4)
    p.rhs = cube_rhs(p, seed + 1)
    p.meta = dict(workload=f"cube_hex8_contrast_{k}", cells=[k, k, k], young="210e9|21e6 (p=0.3)", poisson=0.49)
    return p


def random_model(seed: int, dofs_lo: float = 2e3, dofs_hi: float = 6e4) -> Problem:
    """One model of the batched workload: hex8 or tet10 (even odds), box cell counts drawn so
    the equation count is log-uniform in [dofs_lo, dofs_hi], unit spacing, E ~ U[70, 210] GPa,
    nu ~ U[0.25, 0.35], tet10 corners jittered by +-10 % of the cell, x = 0 clamped."""
    rng = np.random.default_rng(seed)
    kind = "hex8" if rng.random() < 0.5 else "tet10"
    target = float(np.exp(rng.uniform(np.log(dofs_lo), np.log(dofs_hi))))
    aspect = rng.uniform(0.6, 1.6, size=3)
    per = 3.0 if kind == "hex8" else 24.0  # equations per cell, roughly
    cells = np.maximum(1, np.round(aspect * (target / per / np.prod(aspect)) ** (1.0 / 3.0))).astype(int)
    young = rng.uniform(70e9, 210e9)
    nu = rng.uniform(0.25, 0.35)
    p = block(kind, tuple(cells), tuple(float(c) for c in cells), young, nu, clamp_face=0,
              jitter=0.1 if kind == "tet10" else 0.0, seed=seed)
    p.meta = dict(workload="random_model", seed=seed, kind=kind, cells=cells.tolist(), young=young, poisson=nu)
    return p


@dataclass
class Batch:
    """Independent models as one block-diagonal matrix: model s owns equations
    [segments[s], segments[s + 1]). Its symbolic analysis is the union of the models' own
    (the elimination forest of a block-diagonal matrix is the models' trees side by side), so
    one factorization runs every model's fronts of a tree level in the same launches."""

    n: int
    ptr: np.ndarray
    idx: np.ndarray
    kval: np.ndarray
    segments: np.ndarray
    models: list
    rhs: np.ndarray | None = None  # n x nrhs, column-major

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    upper = Problem.upper
    mul = Problem.mul


def batch(models: list, nrhs: int = 64, seed: int = 20243) -> Batch:
    """Block-diagonal concatenation of `models` with nrhs seeded N(0, 1) right-hand sides per
    model (one generator per model, seeded by seed and the model's index)."""
    ns = np.array([m.n for m in models], np.int64)
    seg = np.zeros(len(models) + 1, np.int64)
    np.cumsum(ns, out=seg[1:])
    nz = np.array([m.nnz for m in models], np.int64)
    zoff = np.zeros(len(models) + 1, np.int64)
    np.cumsum(nz, out=zoff[1:])
    n = int(seg[-1])
    ptr = np.zeros(n + 1, np.int64)
    idx = np.zeros(int(zoff[-1]), np.int64)
    kval = np.zeros(int(zoff[-1]))
    rhs = np.zeros((n, nrhs), order="F")
    for s, m in enumerate(models):
        ptr[seg[s] + 1:seg[s + 1] + 1] = m.ptr[1:] + zoff[s]
        idx[zoff[s]:zoff[s + 1]] = m.idx + seg[s]
        kval[zoff[s]:zoff[s + 1]] = m.kval
        rng = np.random.default_rng([seed, s])
        rhs[seg[s]:seg[s + 1]] = rng.standard_normal((m.n, nrhs))
    meta = [dict(getattr(m, "meta", {}), n=m.n) for m in models]
    return Batch(n, ptr, idx, kval, seg, meta, rhs)


def block_tet10(cells=(50, 50, 50), size=(0.5, 0.5, 0.5), rho: float = 7850.0) -> Problem:
    """Tet10 Kuhn block of steel with its consistent mass, z = 0 clamped."""
    p = block("tet10", cells, size, E_STEEL, 0.3, clamp_face=4, rho=rho)
    p.meta = dict(workload="block_tet10", cells=list(cells), size=list(size), rho=rho)
    return p
