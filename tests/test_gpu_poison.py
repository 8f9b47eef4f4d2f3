"""Every step reads only what the same call wrote: with FC_POISON_ALLOC=1 every device
allocation of the library starts as 0xFF bytes (a NaN in FP64 and FP32), so a kernel that reads
memory no step of the current factorization or solve has written would turn the results into
NaNs or a false pivot failure. Each precision's factor and solve must still be the oracle's
(cholesky.rs:1146-1185), and so must a factor made right after a failed one (whose NaNs the
allocator hands back otherwise).
"""
import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.fixture(scope="module")
def cube12():
    p = P.cube_hex8(12)
    up = p.upper()
    ref = O.Factor(p.n, *up)
    return p, up, ref, ref.solve(p.rhs)


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
def test_factor_and_solve_read_nothing_unwritten(gpu, cube12, monkeypatch, precision):
    p, up, ref, xr = cube12
    monkeypatch.setenv("FC_POISON_ALLOC", "1")
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    f = CH.Factor(sym, up[2], precision=precision)
    assert f.precision_used() == precision
    x = f.solve(p.rhs)
    assert np.isfinite(x).all()
    if precision == "fp64":
        assert rel(f.values(), ref.values()) <= 1e-12
        assert rel(x, xr) <= 1e-12
    else:
        assert rel(x, xr) <= 1e-3
        A = CH.Matrix(p.n, p.ptr, p.idx, p.kval)
        xi, rep = f.solve_refined(A, p.rhs, target_berr=1e-12)
        assert rep["converged"], rep
        assert rel(xi, xr) <= 1e-9
    # four right-hand sides at once: each column as its own solve
    bs = [p.rhs * (c + 1) + np.sin(np.arange(p.n) * (0.1 + c)) for c in range(4)]
    xs = f.solve_many(bs)
    for b, xc in zip(bs, xs):
        assert np.array_equal(xc, f.solve(b))


def test_a_factor_after_a_failed_one(gpu, cube12, monkeypatch):
    """A factorization that fails leaves NaNs in the memory the next factor is given: with and
    without poisoning, the next factor of an SPD matrix is the oracle's."""
    p, up, ref, xr = cube12
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    e = sym.export()
    col = int(e["perm"][p.n // 2])
    uptr, uidx, uval = up
    bad = uval.copy()
    assert uidx[uptr[col + 1] - 1] == col
    bad[uptr[col + 1] - 1] = -1.0
    for poison in ("0", "1"):
        monkeypatch.setenv("FC_POISON_ALLOC", poison)
        with pytest.raises(CH.NotPositiveDefinite):
            CH.Factor(sym, bad)
        f = CH.Factor(sym, uval)
        assert rel(f.values(), ref.values()) <= 1e-12
        assert rel(f.solve(p.rhs), xr) <= 1e-12

