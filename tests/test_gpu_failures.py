"""Failure semantics of the device factorization through the C ABI, against the reference's
scheduler (cholesky.rs:1440-1530): a failing pivot is named by the first failing column a pass in
postorder would meet, as an original column, wherever it sits (deep inside a wide front, or in two
subtrees that fail at once); a cancelled factorization stops between levels (cholesky.rs:1721-1728)
and leaves the factor reusable; a device allocation that fails is reported as the factor's own lack
of memory with its entry count (cholesky.rs:1961-1975); a reduced-precision factor that loses
definiteness on an SPD matrix falls back to the next precision instead of failing.
"""
import ctypes as C

import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def supernodal_tree(e):
    """Supernode widths and parents from an exported analysis (the parent of s is the supernode
    of its first row below its columns, cholesky.rs:1265-1270)."""
    first, rptr, rows = e["first"], e["rptr"], e["rows"]
    w = np.diff(first)
    ns = len(w)
    snode = np.repeat(np.arange(ns), w)
    par = np.array([snode[rows[rptr[k] + w[k]]] if rptr[k] + w[k] < rptr[k + 1] else -1 for k in range(ns)])
    return w, par


def zero_diagonals(up, cols):
    uptr, uidx, uval = up
    v = uval.copy()
    for c in cols:
        assert uidx[uptr[c + 1] - 1] == c  # the diagonal closes its column
        v[uptr[c + 1] - 1] = 0.0
    return uptr, uidx, v


@pytest.fixture(scope="module")
def cube16():
    p = P.cube_hex8(16)
    up = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    return p, up, sym, sym.export()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
def test_failure_deep_inside_the_root_front(gpu, cube16, precision):
    """A zero pivot 570 columns into the 870-column root supernode (its fifth 128-column panel):
    every precision names that column, as the oracle does (the reduced precisions with their
    fallback to FP64, which fails at the same column)."""
    p, up, sym, e = cube16
    w, _ = supernodal_tree(e)
    assert w[-1] >= 600
    pos = p.n - 300  # postorder position inside the root supernode
    col = int(e["perm"][pos])
    bad = zero_diagonals(up, [col])
    with pytest.raises(O.OracleError) as o:
        O.Factor(p.n, *bad)
    assert o.value.column == col
    print(f"zero pivot at postorder {pos} (column {col}), {precision}", flush=True)
    with pytest.raises(CH.NotPositiveDefinite) as g:
        CH.Factor(sym, bad[2], precision=precision)
    assert g.value.column == col and g.value.entries == e["entries"]
    if precision != "fp64":  # without the fallback the reduced precision reports it itself
        with pytest.raises(CH.NotPositiveDefinite) as g1:
            CH.Factor(sym, bad[2], precision=precision, fallback=1)
        assert g1.value.column == col


def test_two_subtrees_fail_the_earlier_in_postorder_is_named(gpu, cube16):
    """Zero pivots in the two subtrees under the root: at the top of the first subtree (factored
    late, at a high level of the front tree) and in a leaf of the second (factored first). The
    device meets the leaf's failure first in time; the name is the column a pass in postorder
    meets first (cholesky.rs:1510-1528), the oracle's."""
    p, up, sym, e = cube16
    w, par = supernodal_tree(e)
    ns = len(w)
    kids = np.where(par == ns - 1)[0]
    assert len(kids) == 2
    a = int(kids[0])  # the first subtree's top supernode: postorder ends at it
    first = e["first"]
    leaves = np.setdiff1d(np.arange(ns), par[par >= 0])
    b = int(leaves[leaves > a][0])  # a leaf of the second subtree
    pa, pb = int(first[a]), int(first[b])
    assert pa < pb
    ca, cb = int(e["perm"][pa]), int(e["perm"][pb])
    bad = zero_diagonals(up, [ca, cb])
    with pytest.raises(O.OracleError) as o:
        O.Factor(p.n, *bad)
    assert o.value.column == ca
    with pytest.raises(CH.NotPositiveDefinite) as g:
        CH.Factor(sym, bad[2])
    assert g.value.column == ca
    # the leaf alone fails at its own column
    with pytest.raises(CH.NotPositiveDefinite) as g2:
        CH.Factor(sym, zero_diagonals(up, [cb])[2])
    assert g2.value.column == cb


@pytest.mark.parametrize("graph", [0, 1])
def test_cancel_stops_the_factorization_and_leaves_the_factor_reusable(gpu, cube16, graph):
    """A set cancel flag stops a factorization with Cancelled (polled between levels, or before a
    graph replay); cleared, the same factor refactors and solves as the oracle."""
    p, up, sym, e = cube16
    flag = C.c_int32(1)
    with pytest.raises(CH.Cancelled):
        CH.Factor(sym, up[2], cancel=flag, use_graph=graph)
    flag.value = 0
    f = CH.Factor(sym, up[2], cancel=flag, use_graph=graph)
    flag.value = 1
    with pytest.raises(CH.Cancelled):
        f.refactor(up[2] * 2.0)
    with pytest.raises(CH.InvalidArgument):
        f.solve(p.rhs)  # a cancelled factorization leaves no valid factor
    flag.value = 0
    f.refactor(up[2])
    x, xr = f.solve(p.rhs), O.Factor(p.n, *up).solve(p.rhs)
    assert np.linalg.norm(x - xr) <= 1e-12 * np.linalg.norm(xr)


def test_allocation_failure_is_the_factors_lack_of_memory(gpu, cube16, monkeypatch):
    """With the library's largest device allocation capped (FC_MAX_DEVICE_ALLOC) below the
    factor's storage, the factorization reports OutOfMemory with the factor's entries; lifted,
    the same analysis factors."""
    p, up, sym, e = cube16
    info = sym.info()
    monkeypatch.setenv("FC_MAX_DEVICE_ALLOC", str(info["l_bytes_fp64"] - 8))
    with pytest.raises(CH.OutOfMemory) as g:
        CH.Factor(sym, up[2])
    assert g.value.entries == e["entries"]
    f32 = CH.Factor(sym, up[2], precision="fp32")  # half the storage fits under the cap
    monkeypatch.delenv("FC_MAX_DEVICE_ALLOC")
    f = CH.Factor(sym, up[2])
    x, xr = f.solve(p.rhs), O.Factor(p.n, *up).solve(p.rhs)
    assert np.linalg.norm(x - xr) <= 1e-12 * np.linalg.norm(xr)
    assert np.linalg.norm(f32.solve(p.rhs) - xr) <= 1e-3 * np.linalg.norm(xr)


def test_mixed_form_falls_back_instead_of_failing_an_spd_matrix(gpu):
    """The single-product BF16 form (relative error about 4e-3 a product) loses definiteness on
    a dense Gram matrix B B^T / 199 + 1e-5 I of 200 equations (one front of two panels, whose
    second panel's update is a BF16 product; smallest eigenvalue 1e-5). With the fallback (the
    default) the factorization moves to FP32 and the refined solve reaches berr <= 1e-12;
    without it, NotPositiveDefinite."""
    from matrices import dense_to_csr, upper
    rng = np.random.default_rng(3)
    b = rng.standard_normal((200, 199))
    a = b @ b.T / 199 + 1e-5 * np.eye(200)
    n, ptr, idx, val = dense_to_csr(a)
    up = upper(n, ptr, idx, val)
    sym = CH.Symbolic(CH.UpperColumns(n, *up))
    assert sym.info()["kind_flops"]["small"] == 0  # the tiled launches, BF16 products included
    with pytest.raises(CH.NotPositiveDefinite):
        CH.Factor(sym, up[2], precision="mixed", bf16_terms=1, fallback=1)
    f = CH.Factor(sym, up[2], precision="mixed", bf16_terms=1)
    assert f.precision_used() == "fp32"
    A = CH.Matrix(n, ptr, idx, val)
    rhs = np.cos(np.arange(n) * 0.1)
    x, rep = f.solve_refined(A, rhs, target_berr=1e-12)
    assert rep["converged"], rep
    r = a @ x - rhs
    assert np.abs(r).max() / (np.abs(a).sum(axis=1).max() * np.abs(x).max() + np.abs(rhs).max()) <= 1e-12


def nearly_singular_pairs(p, delta=1e-9, pairs=3):
    """p's stiffness with `pairs` appended 2 x 2 blocks s [[1, 1 - delta], [1 - delta, 1]]: SPD in
    FP64 (eigenvalues s delta, s (2 - delta); the second pivot 2 s delta passes the test), exactly
    singular once rounded to FP32 (1 - delta rounds to 1), where the second pivot is exactly 0."""
    s = float(2.0 ** np.round(np.log2(np.median(np.abs(p.kval)))))  # a power of two: exact in FP32
    n = p.n + 2 * pairs
    rows = [list(zip(p.idx[p.ptr[i]:p.ptr[i + 1]], p.kval[p.ptr[i]:p.ptr[i + 1]])) for i in range(p.n)]
    for q in range(pairs):
        a, b = p.n + 2 * q, p.n + 2 * q + 1
        rows.append([(a, s), (b, s * (1 - delta))])
        rows.append([(a, s * (1 - delta)), (b, s)])
    ptr = np.zeros(n + 1, np.int64)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    idx = np.array([c for r in rows for c, _ in r], np.int64)
    val = np.array([v for r in rows for _, v in r])
    return n, ptr, idx, val


def check_fp64(f, ref, up, b, xr):
    """The FP64 factor of the nearly singular matrix: its values within 1e-12 of the oracle's
    (normwise) and a backward-stable solve. The solutions themselves differ from the oracle's by
    up to cond(A) eps ~ 2e9 x 1.1e-16 along the near-null directions (two backward-stable solves
    with different rounding), so they are held to that bound, not to 1e-12."""
    assert np.linalg.norm(f.values() - ref.values()) <= 1e-12 * np.linalg.norm(ref.values())
    x = f.solve(b)
    n = len(b)
    rows = np.repeat(np.arange(n), np.diff(up.ptr))
    off = up.idx != rows  # the upper columns hold each off-diagonal entry once
    ax = np.bincount(up.idx, up.val * x[rows], minlength=n) + np.bincount(rows[off], up.val[off] * x[up.idx[off]],
                                                                          minlength=n)
    norm = np.max(np.bincount(up.idx, np.abs(up.val), minlength=n) + np.bincount(rows[off], np.abs(up.val[off]),
                                                                                 minlength=n))
    berr = np.abs(b - ax).max() / (norm * np.abs(x).max() + np.abs(b).max())
    assert berr <= 1e-15, berr
    assert np.linalg.norm(x - xr) <= 1e-6 * np.linalg.norm(xr)


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_reduced_precision_falls_back_to_fp64_storage(gpu, monkeypatch, precision):
    """A matrix SPD in FP64 but singular in FP32: the FP32 (and mixed) factor fails, the fallback
    refactors on FP64 storage and the solution is the oracle's to 1e-12. With the FP64 storage's
    allocation refused, OutOfMemory; the factor then keeps its FP32 storage and refactors once the
    memory is there."""
    p = P.cube_hex8(6)
    n, ptr, idx, val = nearly_singular_pairs(p)
    up = CH.UpperColumns.from_symmetric(n, ptr, idx, val)
    ref = O.Factor(n, up.ptr, up.idx, up.val)
    sym = CH.Symbolic(up)
    b = np.cos(np.arange(n) * 0.21)
    with pytest.raises(CH.NotPositiveDefinite):
        CH.Factor(sym, up.val, precision=precision, fallback=1)
    f = CH.Factor(sym, up.val, precision=precision)
    assert f.precision_used() == "fp64"
    xr = ref.solve(b)
    check_fp64(f, ref, up, b, xr)
    g = CH.Factor(sym, up.val * 0 + np.where(up.idx == np.repeat(np.arange(n), np.diff(up.ptr)), 1.0, 0.0),
                  precision=precision)  # the identity: no fallback, FP32 storage
    assert g.precision_used() == precision
    monkeypatch.setenv("FC_MAX_DEVICE_ALLOC", str(sym.info()["l_bytes_fp64"] - 8))
    with pytest.raises(CH.OutOfMemory):
        g.refactor(up.val)
    monkeypatch.delenv("FC_MAX_DEVICE_ALLOC")
    g.refactor(up.val)
    assert g.precision_used() == "fp64"
    check_fp64(g, ref, up, b, xr)
