"""Every kernel reads only the LDS it has written in the same launch: before each library call
the LDS of every CU is filled with NaNs (tests/hip/lds_dirty.hip), so a kernel that reads LDS it
did not write — even with a 0 weight, NaN x 0 = NaN — turns its results into NaNs, where whatever
finite values earlier kernels left there would hide it. The results must still be the oracle's:
the Cholesky factor and solves in each precision (one and many right-hand sides, refined), the
real and complex LDL^T factors bit for bit, and shift-invert Lanczos.
"""
import ctypes as C
import os

import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P
from oracle import oracle as O

from test_zldlt_oracle import complex_grid, upper as zupper

pytestmark = pytest.mark.gpu

_HELPER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hip", "_lds_dirty.so")


@pytest.fixture(scope="module")
def dirty(gpu):
    if not os.path.exists(_HELPER):
        pytest.fail(f"{_HELPER} is not built (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(_HELPER)
    lib.lds_dirty.restype = C.c_int

    def run():
        st = lib.lds_dirty()
        assert st == 0, f"lds_dirty: HIP error {st}"

    return run


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def bits(v):
    return np.ascontiguousarray(v).view(np.uint64)


@pytest.fixture(scope="module")
def cube12():
    p = P.cube_hex8(12)
    up = p.upper()
    ref = O.Factor(p.n, *up)
    return p, up, ref, ref.solve(p.rhs)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
def test_cholesky_with_dirty_lds(dirty, cube12, precision):
    p, up, ref, xr = cube12
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    dirty()
    f = CH.Factor(sym, up[2], precision=precision)
    assert f.precision_used() == precision
    dirty()
    x = f.solve(p.rhs)
    bs = [p.rhs * (c + 1) + np.sin(np.arange(p.n) * (0.1 + c)) for c in range(20)]
    for nb in (4, 20):  # the 8- and 16-wide solve kernels
        dirty()
        xs = f.solve_many(bs[:nb])
        for b, xc in zip(bs[:nb], xs):
            dirty()
            assert np.array_equal(bits(xc), bits(f.solve(b)))
    if precision == "fp64":
        assert rel(f.values(), ref.values()) <= 1e-12
        assert rel(x, xr) <= 1e-12
    else:
        assert rel(x, xr) <= 1e-3
        A = CH.Matrix(p.n, p.ptr, p.idx, p.kval)
        dirty()
        xi, rep = f.solve_refined(A, p.rhs, target_berr=1e-12)
        assert rep["converged"], rep
        assert rel(xi, xr) <= 1e-9


def test_ldlt_with_dirty_lds(dirty, cube12):
    p, up, ref, xr = cube12
    uptr, uidx, uval = up
    diag = uidx == np.repeat(np.arange(p.n), np.diff(uptr))
    upi = (uptr, uidx, uval - 0.6180339887 * float(np.median(uval[diag])) * diag)
    lo = O.LdltFactor(p.n, *upi)
    dirty()
    lf = CH.LdltFactor.new(CH.UpperColumns(p.n, *upi))
    assert np.array_equal(bits(lf.values()), bits(lo.values()))
    dirty()
    assert np.array_equal(bits(lf.solve(p.rhs)), bits(lo.solve(p.rhs)))


def test_complex_ldlt_with_dirty_lds(dirty):
    a = complex_grid([10, 9, 8], 2.7)
    n = a[0]
    o = O.ZLdltFactor(n, *zupper(a))
    dirty()
    f = CH.ZLdltFactor.new(n, *zupper(a))
    assert np.array_equal(bits(f.values()), bits(o.values()))
    rhs = np.cos(np.arange(n) * 0.3) + 1j * np.sin(np.arange(n) * 0.7)
    dirty()
    assert np.array_equal(bits(f.solve(rhs)), bits(o.solve(rhs)))


def test_lanczos_with_dirty_lds(dirty):
    n = 300
    lam = 1.0 + np.arange(n) * 0.37
    m = 1.0 + (np.arange(n) % 5) * 0.2
    ptr, idx = np.arange(n + 1), np.arange(n)
    shift = 0.5
    sym = CH.Symbolic(CH.UpperColumns(n, ptr, idx, np.zeros(n)))
    dirty()
    f = CH.Factor(sym, lam * m - shift * m)
    M = CH.Matrix(n, ptr, idx, m)
    dirty()
    got = f.lanczos_device(M, shift, 8, block=6, max_basis=n, tolerance=1e-9)
    assert np.abs(got["values"] - lam[:8]).max() <= 1e-10 * lam[7]
    x = got["vectors"].cpu().numpy()
    assert np.abs(x.T @ (m[:, None] * x) - np.eye(8)).max() < 1e-10
