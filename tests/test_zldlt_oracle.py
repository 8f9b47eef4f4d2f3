"""The complex-symmetric LDL^T oracle (orc_zldlt_* in oracle/chol_oracle.c: ldlt.rs over
Complex64) pinned by the reference's own complex tests (ldlt.rs:1053-1225): a complex grid whose
Hermitian part differs from its symmetric part solved to 1e-10 / residual 1e-12 (a conjugate
anywhere would be a wrong answer), the real 2 x 2 block form [[K, -C], [-C, -K]] solved by the
real LDL^T oracle, damped oscillators against Cramer's rule, a saddle point that needs 2 x 2
pivots, the same bits on any number of threads, and a singular complex Laplacian's located
error."""
import numpy as np
import pytest

from oracle import oracle as O

from matrices import rows_to_csr, saddle, shifted_grid


def complex_grid(n3, sigma):
    """complex_grid (ldlt.rs:1067-1075): K + 0.3 i K0 on a real grid's pattern."""
    n, ptr, idx, kv = shifted_grid(n3, sigma)[0]
    k0 = shifted_grid(n3, 0.0)[0][3]
    return n, ptr, idx, kv + 0.3j * k0


def sample_solution(n):
    i = np.arange(n, dtype=float)
    return np.sin(i * 0.37) + 1j * np.cos(i * 0.11)


def cmul(a, x):
    n, ptr, idx, val = a
    rows = np.repeat(np.arange(n), np.diff(ptr))
    out = np.zeros(n, np.complex128)
    np.add.at(out, rows, val * x[idx])
    return out


def upper(a):
    n, ptr, idx, val = a
    rows = np.repeat(np.arange(n), np.diff(ptr))
    keep = idx <= rows
    uptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows[keep], minlength=n), out=uptr[1:])
    return uptr, idx[keep].astype(np.int64), val[keep]


def zfactor(a):
    return O.ZLdltFactor(a[0], *upper(a))


def residual(a, x, rhs):
    return np.abs(rhs - cmul(a, x)).max() / np.abs(rhs).max()


def test_complex_symmetric_systems_solve_with_the_transpose():
    a = complex_grid([5, 4, 4], 2.1)
    n, ptr, idx, val = a
    at = {(i, int(idx[e])): val[e] for i in range(n) for e in range(ptr[i], ptr[i + 1])}
    assert at[(0, 3)] == at[(3, 0)] and abs(at[(0, 3)] - np.conj(at[(3, 0)])) > 0.1
    x = sample_solution(n)
    b = cmul(a, x)
    y = zfactor(a).solve(b)
    assert np.abs(y - x).max() < 1e-10 and residual(a, y, b) < 1e-12


def test_complex_solve_matches_the_real_two_by_two_block_form():
    k = shifted_grid([4, 3, 3], -0.5)[0]
    c = shifted_grid([4, 3, 3], 0.0)[0]
    n = k[0]
    a = (n, k[1], k[2], k[3] + 1j * c[3])
    rows = [[] for _ in range(2 * n)]
    for i in range(n):
        for e in range(k[1][i], k[1][i + 1]):
            rows[i].append((int(k[2][e]), k[3][e]))
            rows[n + i].append((n + int(k[2][e]), -k[3][e]))
        for e in range(c[1][i], c[1][i + 1]):
            rows[i].append((n + int(c[2][e]), -c[3][e]))
            rows[n + i].append((int(c[2][e]), -c[3][e]))
    block = rows_to_csr(rows)
    b_re = np.sin((np.arange(n) + 1) * 0.7)
    b_im = np.cos(np.arange(n) * 0.4)
    real = O.LdltFactor(2 * n, *O.from_symmetric(*block)).solve(np.concatenate([b_re, -b_im]))
    z = zfactor(a).solve(b_re + 1j * b_im)
    gap = np.abs(z - (real[:n] + 1j * real[n:])).max()
    assert gap < 1e-10 * np.abs(z).max()
    assert residual(a, z, b_re + 1j * b_im) < 1e-12


def test_damped_oscillators_match_the_closed_form():
    k, c, w = 4.0, 0.5, 1.3
    d = complex(2.0 * k - w * w, w * c)
    one = (1, np.array([0, 1]), np.array([0]), np.array([complex(k - w * w, w * c)]))
    x = zfactor(one).solve(np.array([1.0 + 0j]))
    want = 1.0 / complex(k - w * w, w * c)
    assert abs(x[0] - want) < 1e-12 * abs(want)
    two = (2, np.array([0, 2, 4]), np.array([0, 1, 0, 1]), np.array([d, -k, -k, d], np.complex128))
    f1, f2 = 1.0 + 0j, 0.5 + 0j
    det = d * d - k * k
    want = [(d * f1 + k * f2) / det, (k * f1 + d * f2) / det]
    x = zfactor(two).solve(np.array([f1, f2]))
    assert all(abs(g - h) < 1e-12 * abs(h) for g, h in zip(x, want))


def test_complex_saddle_point_forces_two_by_two_pivots():
    n, ptr, idx, val = saddle(8, 3)
    a = (n, ptr, idx, val + 0.3j * val)
    f = zfactor(a)
    assert f.two_by_two() > 0
    x = sample_solution(n)
    rhs = cmul(a, x)
    assert residual(a, f.solve(rhs), rhs) < 1e-12


def test_complex_factor_is_the_same_on_any_number_of_threads():
    a = complex_grid([10, 9, 8], 2.7)
    x = sample_solution(a[0])
    rhs = cmul(a, x)
    threads = O.threads()
    try:
        O.set_threads(1)
        one = zfactor(a)
        ref = one.solve(rhs)
        for t in (2, 7):
            O.set_threads(t)
            other = zfactor(a)
            assert np.array_equal(one.values().view(np.uint64), other.values().view(np.uint64))
            assert np.array_equal(other.solve(rhs).view(np.uint64), ref.view(np.uint64))
    finally:
        O.set_threads(threads)
    assert residual(a, ref, rhs) < 1e-11


def test_singular_complex_matrices_are_located_errors():
    n = 10
    rows = []
    for i in range(n):
        r, d = [], 0.0
        for j in (i - 1, i + 1):
            if 0 <= j < n:
                r.append((j, -1.0))
                d += 1.0
        r.append((i, d))
        rows.append(r)
    m, ptr, idx, val = rows_to_csr(rows)
    with pytest.raises(O.OracleError) as e:
        zfactor((m, ptr, idx, val.astype(np.complex128)))
    assert e.value.kind == O.SINGULAR and 0 <= e.value.column < n
