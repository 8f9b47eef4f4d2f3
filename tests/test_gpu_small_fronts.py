"""GPU parity of the small-front launch (a front of at most 64 columns taken whole by one
workgroup) against the oracle, through the C ABI: a dense 64-column front alone (FP64 values and
solutions within 1e-12); fronts whose FP32 attempt meets a small pivot (below 1e-3 |A_jj|) and are
factored again in FP64 (every value of the FP32-storage forms within FP32 rounding of the oracle's
factor of the FP32-rounded matrix, where FP32 arithmetic would be off by 1e-3); a failing pivot inside a small front, named by its
column in every precision."""
import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from oracle import oracle as O

from matrices import dense_to_csr, random_spd, upper

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def factor_both(a, precision):
    n, ptr, idx, val = dense_to_csr(a)
    up = upper(n, ptr, idx, val)
    sym = CH.Symbolic(CH.UpperColumns(n, *up))
    return sym, O.Factor(n, *up), CH.Factor(sym, up[2], precision=precision), (n, ptr, idx, val)


def test_one_dense_small_front(gpu):
    a = random_spd(64, 1.0, 11)
    sym, ref, f, _ = factor_both(a, "fp64")
    assert sym.info()["kind_flops"]["small"] > 0
    assert rel(f.values(), ref.values()) <= 1e-12
    b = np.sin(np.arange(64) * 0.7)
    assert rel(f.solve(b), ref.solve(b)) <= 1e-12
    assert rel(f.solve(b), np.linalg.solve(a, b)) <= 1e-12


def near_singular_pairs(m=10, eps=1e-5, delta=1e-4):
    """m pairs [[3, s], [s, 1]], s = sqrt(3) (1 - eps) (second pivot about 2 eps: below
    1e-3 |A_jj|), each coupled by delta to a last equation of diagonal 1 (SPD: its Schur
    complement is about 1 - m delta^2 / (2 eps) = 0.995). In FP32 arithmetic the first column's
    s / sqrt(3) carries a rounding of 6e-8, which the second pivot's cancellation turns into a
    relative error of about 1e-3 in L(1, 1)."""
    n = 2 * m + 1
    a = np.zeros((n, n))
    for i in range(m):
        a[2 * i, 2 * i], a[2 * i + 1, 2 * i + 1] = 3.0, 1.0
        a[2 * i, 2 * i + 1] = a[2 * i + 1, 2 * i] = np.sqrt(3.0) * (1.0 - eps)
        a[2 * i + 1, n - 1] = a[n - 1, 2 * i + 1] = delta
    a[n - 1, n - 1] = 1.0
    return a


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_small_pivots_are_factored_again_in_fp64(gpu, precision):
    """The FP32-storage forms hold the matrix rounded to FP32: their factor is the oracle's
    factor of that rounded matrix, each value within FP32 rounding (FP32 arithmetic would be off
    by about 1e-3 in each pair's L(1, 1))."""
    a = near_singular_pairs()
    sym, _, f, (n, ptr, idx, val) = factor_both(a, precision)
    assert sym.info()["kind_flops"]["small"] > 0
    up32 = upper(n, ptr, idx, val.astype(np.float32).astype(np.float64))
    ref32 = O.Factor(n, *up32)
    v, r = np.asarray(f.values()), np.asarray(ref32.values())
    nz = r != 0
    assert np.array_equal(v != 0, nz)
    assert (np.abs(v[nz] - r[nz]) / np.abs(r[nz])).max() <= 1e-6
    A = CH.Matrix(n, ptr, idx, val)
    b = np.cos(np.arange(n) * 0.3)
    x, rep = f.solve_refined(A, b, target_berr=1e-12)
    assert rep["converged"], rep
    assert rel(x, np.linalg.solve(a, b)) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32", "mixed"])
def test_a_failing_pivot_in_a_small_front_is_named(gpu, precision):
    a = random_spd(48, 1.0, 5)
    a[17, 17] = 1e-3 * a[17, 17]  # indefinite: a negative pivot at column 17 or after it
    n, ptr, idx, val = dense_to_csr(a)
    up = upper(n, ptr, idx, val)
    with pytest.raises(O.OracleError) as o:
        O.Factor(n, *up)
    sym = CH.Symbolic(CH.UpperColumns(n, *up))
    assert sym.info()["kind_flops"]["small"] > 0
    with pytest.raises(CH.NotPositiveDefinite) as g:
        CH.Factor(sym, up[2], precision=precision)
    assert g.value.column == o.value.column


@pytest.mark.parametrize("n", [65, 96, 97, 129, 160, 168, 200, 250])
@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_fp32_diagonal_blocks_of_every_quarter_shape(gpu, n, precision):
    """A dense front of n columns (wider than a small front) factored by the tiled kernels: its
    128-column panels' FP32 diagonal blocks go in 32-column quarters, and these widths give a
    last panel of 1, 32, 33, 40, 72, 122 columns and halves of every shape (a lone quarter, a
    quarter and a part, two whole ones). Every value within 2e-6 of the oracle's factor of the
    FP32-rounded matrix (normwise; 1e-4 for the mixed form, whose panel updates are BF16 split
    products), and its inverse blocks right: the refined solve reaches berr <= 1e-12. A wrong
    quarter product would be off by O(1)."""
    a = random_spd(n, 1.0, 100 + n)
    sym, _, f, (nn, ptr, idx, val) = factor_both(a, precision)
    assert sym.info()["kind_flops"]["small"] == 0
    ref32 = O.Factor(nn, *upper(nn, ptr, idx, val.astype(np.float32).astype(np.float64)))
    assert rel(f.values(), ref32.values()) <= (2e-6 if precision == "fp32" else 1e-4)
    A = CH.Matrix(nn, ptr, idx, val)
    b = np.sin(np.arange(nn) * 0.9)
    x, rep = f.solve_refined(A, b, target_berr=1e-12)
    assert rep["converged"], rep
    assert rel(x, np.linalg.solve(a, b)) <= 1e-10
