"""The LDL^T oracle (oracle/chol_oracle.c: the restatement of crates/solver/src/ldlt.rs, bounded
Bunch-Kaufman pivoting inside supernodes) pinned by the reference's own tests:

* its frozen fingerprints (ldlt.rs:760-846): FNV-1a over the bits of the factor's values and of
  the solution for an exact right-hand side, with the negative-pivot and 2 x 2 counts, on a
  definite and an indefinite shifted grid, a saddle point and a hex-pattern block numbered by
  nested dissection: bit for bit;
* inertia against the closed-form eigenvalue count of the shifted Dirichlet grid Laplacian,
  residual bounds, the saddle point's 2 x 2 pivots, located singular and NaN failures, the cap,
  the empty and 1 x 1 matrices (ldlt.rs:634-756);
* the same bits on any number of threads."""
import numpy as np
import pytest

from oracle import oracle as O

from matrices import fingerprint, frozen_rhs, hex_pattern, rows_to_csr, saddle, shifted_grid

FROZEN = [  # ldlt.rs:807-844
    ("shifted grid, definite", lambda: shifted_grid([5, 5, 5], -0.5)[0], 0, 0, 12318096910519023990,
     10217835541271929398),
    ("shifted grid, indefinite", lambda: shifted_grid([7, 6, 5], 3.3)[0], 78, 9, 17838274908630861643,
     4015666017640459628),
    ("saddle point, 2 x 2 pivots", lambda: saddle(8, 3), 3, 3, 4028760691601933113, 11718422723752440401),
    ("hex stiffness, indefinite", lambda: hex_pattern(6, 22.0, O.nested_dissection), 45, 40, 2687260628477475718,
     8749587493275634787),
]


def factor(a, **kw):
    n, ptr, idx, val = a
    return O.LdltFactor(n, *O.from_symmetric(n, ptr, idx, val), **kw)


def residual(a, f, seed):
    """residual (ldlt.rs:585-593): x = sin(0.37 (i + seed)), b = A x, max |solve(b) - x|."""
    n, ptr, idx, val = a
    x = np.sin((np.arange(n) + seed) * 0.37)
    b = O.csr_mul(n, ptr, idx, val, x)
    return np.abs(f.solve(b) - x).max()


@pytest.mark.parametrize("name,make,negative,two,factor_bits,solution_bits", FROZEN, ids=[c[0] for c in FROZEN])
def test_frozen_fingerprints(name, make, negative, two, factor_bits, solution_bits):
    a = make()
    n, ptr, idx, val = a
    f = factor(a)
    x = f.solve(frozen_rhs(n, lambda v: O.csr_mul(n, ptr, idx, val, v)))
    assert f.inertia()[1] == negative
    assert f.two_by_two() == two
    assert fingerprint(f.values()) == factor_bits
    assert fingerprint(x) == solution_bits


@pytest.mark.parametrize("sigma", [0.0, 0.9, 3.3, 7.1])
def test_indefinite_systems_solve_with_their_inertia(sigma):
    """ldlt.rs:634-652"""
    a, below = shifted_grid([7, 6, 5], sigma)
    f = factor(a)
    assert f.inertia() == (a[0] - below, below)
    assert residual(a, f, sigma) < 1e-10


def test_definite_matrix_as_cholesky():
    a, _ = shifted_grid([5, 5, 5], -0.5)
    f = factor(a)
    assert f.inertia()[1] == 0
    n, ptr, idx, val = a
    up = O.from_symmetric(n, ptr, idx, val)
    assert len(f.values()) == O.Factor(n, *up).entries()  # the same fill
    assert residual(a, f, 1.0) < 1e-12


def test_two_by_two_pivots_for_zero_diagonals():
    """ldlt.rs:690-698"""
    a = saddle(8, 3)
    f = factor(a)
    assert f.inertia() == (8, 3)
    assert f.two_by_two() > 0
    assert residual(a, f, 0.5) < 1e-12


def test_singular_and_hostile_matrices_are_located_errors():
    """ldlt.rs:700-734"""
    n = 12
    rows = []
    for i in range(n):
        r, d = [], 0.0
        for j in (i - 1, i + 1):
            if 0 <= j < n:
                r.append((j, -1.0))
                d += 1.0
        r.append((i, d))
        rows.append(r)
    with pytest.raises(O.OracleError) as e:
        factor(rows_to_csr(rows))
    assert e.value.kind == O.SINGULAR and 0 <= e.value.column < n
    bad = [list(r) for r in rows]
    bad[5][0] = (bad[5][0][0], float("nan"))
    bad[4] = [(c, float("nan") if c == 5 else v) for c, v in bad[4]]
    with pytest.raises(O.OracleError) as e:
        factor(rows_to_csr(bad))
    assert e.value.kind == O.SINGULAR
    a, _ = shifted_grid([3, 3, 3], 2.0)
    with pytest.raises(O.OracleError) as e:
        factor(a, max_entries=10)
    assert e.value.kind == O.TOO_LARGE
    f = factor(rows_to_csr([[(0, -4.0)]]), max_entries=10)
    assert f.solve(np.array([2.0])).tolist() == [-0.5]
    assert f.inertia() == (0, 1)
    empty = O.LdltFactor(0, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0), max_entries=10)
    assert empty.solve(np.zeros(0)).shape == (0,)


def test_same_factor_on_any_number_of_threads():
    """ldlt.rs:736-756"""
    a, below = shifted_grid([14, 12, 10], 2.7)
    threads = O.threads()
    try:
        O.set_threads(1)
        one = factor(a)
        O.set_threads(max(threads, 2))
        many = factor(a)
    finally:
        O.set_threads(threads)
    assert one.inertia()[1] == below
    assert np.array_equal(one.values().view(np.uint64), many.values().view(np.uint64))
    assert residual(a, one, 0.1) < 1e-9
