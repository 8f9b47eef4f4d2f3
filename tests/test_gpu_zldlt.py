"""GPU parity of the complex-symmetric LDL^T (fc_zldlt_numeric / fc_zldlt_solve through the C ABI)
with the oracle's restatement (tests/test_zldlt_oracle.py pins it by the reference's complex
tests, ldlt.rs:1053-1225): the factor's values and the solutions bit for bit on complex grids
and a complex saddle point, the reference's closed forms (damped oscillators, the real 2 x 2
block form), a located singular error, several right-hand sides as each alone, and a real solve
of a complex factor refused."""
import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from oracle import oracle as O

from matrices import rows_to_csr, saddle, shifted_grid
from test_zldlt_oracle import cmul, complex_grid, residual, sample_solution, upper

pytestmark = pytest.mark.gpu


def device(a):
    n = a[0]
    uptr, uidx, uval = upper(a)
    return CH.ZLdltFactor.new(n, uptr, uidx, uval)


def bits(v):
    return np.ascontiguousarray(v).view(np.uint64)


@pytest.mark.parametrize("case", ["grid 5 x 4 x 4", "grid 10 x 9 x 8", "saddle", "grid 16 x 16 x 14"])
def test_complex_factor_and_solution_match_the_oracle_bit_for_bit(gpu, case):
    if case == "saddle":
        n, ptr, idx, val = saddle(8, 3)
        a = (n, ptr, idx, val + 0.3j * val)
    else:
        dims = {"grid 5 x 4 x 4": ([5, 4, 4], 2.1), "grid 10 x 9 x 8": ([10, 9, 8], 2.7),
                "grid 16 x 16 x 14": ([16, 16, 14], 4.1)}[case]
        a = complex_grid(*dims)
    n = a[0]
    f = device(a)
    o = O.ZLdltFactor(n, *upper(a))
    assert f.two_by_two() == o.two_by_two()
    assert np.array_equal(bits(f.values()), bits(o.values()))
    x = sample_solution(n)
    rhs = cmul(a, x)
    y = f.solve(rhs)
    assert np.array_equal(bits(y), bits(o.solve(rhs)))
    assert residual(a, y, rhs) < 1e-11


def test_damped_oscillators_and_the_block_form_on_the_device(gpu):
    k, c, w = 4.0, 0.5, 1.3
    d = complex(2.0 * k - w * w, w * c)
    two = (2, np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.array([d, -k, -k, d], np.complex128))
    det = d * d - k * k
    want = [(d * 1.0 + k * 0.5) / det, (k * 1.0 + d * 0.5) / det]
    x = device(two).solve(np.array([1.0 + 0j, 0.5 + 0j]))
    assert all(abs(g - h) < 1e-12 * abs(h) for g, h in zip(x, want))
    kk = shifted_grid([4, 3, 3], -0.5)[0]
    cc = shifted_grid([4, 3, 3], 0.0)[0]
    a = (kk[0], kk[1], kk[2], kk[3] + 1j * cc[3])
    b = np.sin((np.arange(a[0]) + 1) * 0.7) + 1j * np.cos(np.arange(a[0]) * 0.4)
    assert residual(a, device(a).solve(b), b) < 1e-12


def test_singular_complex_matrix_and_argument_errors(gpu):
    n = 10
    rows = []
    for i in range(n):
        r, dd = [], 0.0
        for j in (i - 1, i + 1):
            if 0 <= j < n:
                r.append((j, -1.0))
                dd += 1.0
        r.append((i, dd))
        rows.append(r)
    m, ptr, idx, val = rows_to_csr(rows)
    with pytest.raises(CH.Singular) as e:
        device((m, ptr, idx, val.astype(np.complex128)))
    assert 0 <= e.value.column < n
    a = complex_grid([4, 3, 3], 2.1)
    f = device(a)
    rhs = [cmul(a, sample_solution(a[0]) * s) for s in (1.0, 0.5j, 2.0)]
    alone = [f.solve(b) for b in rhs]
    assert all(np.array_equal(bits(x), bits(y)) for x, y in zip(f.solve_many(rhs), alone))
    with pytest.raises(CH.FactorError):
        f.solve(np.zeros(a[0] + 1, np.complex128))
    err = CH._Error()
    x = np.zeros(a[0])
    assert CH.lib().fc_ldlt_solve(f._h, x, a[0], CH.C.byref(err)) != 0  # a real solve of a complex factor
