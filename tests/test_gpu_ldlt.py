"""GPU parity of the LDL^T factor (fc_ldlt_numeric / fc_ldlt_solve through the C ABI) with the
reference: its frozen fingerprints (ldlt.rs:760-846) bit for bit on the device (the factor's
values, the solution of an exact right-hand side, the negative-pivot and 2 x 2 counts), which the
oracle reproduces too (tests/test_ldlt_oracle.py); inertia against the closed-form eigenvalue
count of the shifted grid Laplacian, the residual bounds, the saddle point's 2 x 2 pivots,
located singular and NaN failures, the 1 x 1 and empty matrices (ldlt.rs:634-734)."""
import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from oracle import oracle as O

from matrices import fingerprint, frozen_rhs, hex_pattern, rows_to_csr, saddle, shifted_grid
from test_ldlt_oracle import FROZEN

pytestmark = pytest.mark.gpu


def device_factor(a):
    n, ptr, idx, val = a
    up = O.from_symmetric(n, ptr, idx, val)
    return CH.LdltFactor(CH.Symbolic(CH.UpperColumns(n, *up)), up[2])


def residual(a, f, seed):
    n, ptr, idx, val = a
    x = np.sin((np.arange(n) + seed) * 0.37)
    return np.abs(f.solve(O.csr_mul(n, ptr, idx, val, x)) - x).max()


@pytest.mark.parametrize("name,make,negative,two,factor_bits,solution_bits", FROZEN, ids=[c[0] for c in FROZEN])
def test_frozen_fingerprints_on_the_device(gpu, name, make, negative, two, factor_bits, solution_bits):
    a = make()
    n, ptr, idx, val = a
    print(f"LDL^T {name}, n = {n}", flush=True)
    f = device_factor(a)
    x = f.solve(frozen_rhs(n, lambda v: O.csr_mul(n, ptr, idx, val, v)))
    assert (f.inertia()[1], f.two_by_two()) == (negative, two)
    assert fingerprint(f.values()) == factor_bits
    assert fingerprint(x) == solution_bits


@pytest.mark.parametrize("sigma", [0.0, 0.9, 3.3, 7.1])
def test_inertia_and_residual(gpu, sigma):
    a, below = shifted_grid([7, 6, 5], sigma)
    f = device_factor(a)
    assert f.inertia() == (a[0] - below, below)
    assert residual(a, f, sigma) < 1e-10


def test_saddle_point(gpu):
    a = saddle(8, 3)
    f = device_factor(a)
    assert f.inertia() == (8, 3) and f.two_by_two() > 0
    assert residual(a, f, 0.5) < 1e-12


def test_located_failures_and_small_cases(gpu):
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
    a = rows_to_csr(rows)
    with pytest.raises(CH.Singular) as e:
        device_factor(a)
    with pytest.raises(O.OracleError) as o:
        O.LdltFactor(a[0], *O.from_symmetric(*a))
    assert e.value.column == o.value.column
    bad = [list(r) for r in rows]
    bad[5][0] = (bad[5][0][0], float("nan"))
    bad[4] = [(c, float("nan") if c == 5 else v) for c, v in bad[4]]
    with pytest.raises(CH.Singular):
        device_factor(rows_to_csr(bad))
    f = device_factor(rows_to_csr([[(0, -4.0)]]))
    assert f.solve(np.array([2.0])).tolist() == [-0.5] and f.inertia() == (0, 1)
    empty = CH.LdltFactor(CH.Symbolic(CH.UpperColumns(0, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0))),
                          np.zeros(0))
    assert empty.solve(np.zeros(0)).shape == (0,)


def test_bigger_grid_matches_the_oracle_bit_for_bit(gpu):
    """ldlt.rs:736-756's grid (wide supernodes spanning panels, many 2 x 2 pivots)."""
    a, below = shifted_grid([14, 12, 10], 2.7)
    n, ptr, idx, val = a
    f = device_factor(a)
    o = O.LdltFactor(n, *O.from_symmetric(n, ptr, idx, val))
    assert f.inertia()[1] == below
    assert np.array_equal(f.values().view(np.uint64), o.values().view(np.uint64))
    b = np.cos(np.arange(n) * 0.3)
    assert np.array_equal(f.solve(b).view(np.uint64), o.solve(b).view(np.uint64))


@pytest.mark.parametrize("case", ["grid 16 x 16 x 14", "hex block 12^3"])
def test_wide_supernodes_match_the_oracle_bit_for_bit(gpu, case):
    """Supernodes wider than 192 columns are factored one panel per launch with the trailing
    update spread over many workgroups: the same bits as the oracle's one-pass order."""
    if case.startswith("grid"):
        a, below = shifted_grid([16, 16, 14], 4.1)
    else:
        a, below = hex_pattern(12, 22.0, O.nested_dissection), None
    n, ptr, idx, val = a
    up = O.from_symmetric(n, ptr, idx, val)
    sym = CH.Symbolic(CH.UpperColumns(n, *up))
    assert np.diff(sym.export()["first"]).max() > 192
    f = CH.LdltFactor(sym, up[2])
    o = O.LdltFactor(n, *up)
    assert f.inertia() == o.inertia() and f.two_by_two() == o.two_by_two()
    if below is not None:
        assert f.inertia()[1] == below
    assert np.array_equal(f.values().view(np.uint64), o.values().view(np.uint64))
    b = np.cos(np.arange(n) * 0.3)
    x = f.solve(b)
    assert np.array_equal(x.view(np.uint64), o.solve(b).view(np.uint64))
    assert np.abs(O.csr_mul(n, ptr, idx, val, x) - b).max() < 1e-8 * np.abs(b).max() * max(1.0, np.abs(val).max())


def test_several_right_hand_sides_solve_as_each_alone(gpu):
    """ldlt.rs:852-884: solve_many gives each right-hand side the bits of solving it alone; a
    wrong length fails before anything is solved."""
    cases = [rows_to_csr([[(0, 4.0)]]), shifted_grid([4, 3, 3], -0.5)[0], shifted_grid([4, 3, 3], 2.7)[0],
             shifted_grid([5, 4, 3], 7.1)[0]]
    for a in cases:
        n = a[0]
        f = device_factor(a)
        for k in (1, 2, 6):
            rhs = [np.cos((np.arange(n) * 5 + s * 3) * 0.37) for s in range(k)]
            alone = [f.solve(b) for b in rhs]
            together = f.solve_many(rhs)
            assert all(np.array_equal(x.view(np.uint64), y.view(np.uint64)) for x, y in zip(together, alone))
        good, short = np.ones(n), np.full(max(n - 1, 0), 0.25)
        with pytest.raises(CH.Internal):
            f.solve_many([good, short])


def test_solve_across_pivoted_supernodes_and_rhs_errors(gpu):
    """ldlt.rs:886-958: two dense blocks sharing a separator, interleaved so that the postorder
    differs from the input order, the first block needing an interchange and a 2 x 2 pivot; the
    solves of several cases to 1e-10, repeated solves bit-identical, wrong lengths refused."""
    rows = [[] for _ in range(7)]
    for i, row in enumerate([[0.0, 0.1, 3.0], [0.1, 5.0, 0.2], [3.0, 0.2, 0.5]]):
        rows[i] += [(j, v) for j, v in enumerate(row)]
    for i in range(3, 6):
        rows[i] += [(j, 4.0 if i == j else 0.25) for j in range(3, 6)]
    for i in range(6):
        rows[i].append((6, 0.15))
        rows[6].append((i, 0.15))
    rows[6].append((6, 8.0))
    order = [3, 0, 4, 1, 5, 2, 6]
    inverse = [order.index(i) for i in range(7)]
    mixed = rows_to_csr([[(inverse[j], v) for j, v in rows[i]] for i in order])
    cases = [rows_to_csr([[(0, 4.0)]]), rows_to_csr([[(0, -4.0)]]), mixed, shifted_grid([4, 3, 3], -0.5)[0],
             shifted_grid([4, 3, 3], 2.7)[0], shifted_grid([4, 3, 3], 7.1)[0]]
    for a in cases:
        n, ptr, idx, val = a
        f = device_factor(a)
        o = O.LdltFactor(n, *O.from_symmetric(n, ptr, idx, val))
        assert f.two_by_two() == o.two_by_two()
        for seed in (0.0, 0.5, 3.25):
            expected = np.sin((np.arange(n) + seed) * 0.37)
            rhs = O.csr_mul(n, ptr, idx, val, expected)
            x = f.solve(rhs)
            assert np.abs(x - expected).max() < 1e-10
            assert np.array_equal(x.view(np.uint64), f.solve(rhs).view(np.uint64))
            assert np.array_equal(x.view(np.uint64), o.solve(rhs).view(np.uint64))
        for length in (n + 1, n - 1):
            before = np.full(length, 0.25)
            b = before.copy()
            with pytest.raises(CH.FactorError):
                f.solve(b)
            assert np.array_equal(b, before)
    assert device_factor(mixed).two_by_two() >= 1


def test_auto_selection_and_refinement(gpu):
    """ldlt.rs:960-982: Kind::Auto picks Cholesky for a definite grid and LDL^T (on the same
    analysis) for an indefinite one, whose inertia it reports; solve_refined to 1e-12; Definite
    refuses the indefinite matrix."""
    spd, _ = shifted_grid([4, 4, 4], -1.0)
    s = CH.Symmetric(CH.UpperColumns.from_symmetric(*spd), kind="auto")
    assert not s.indefinite and s.inertia() == (spd[0], 0)
    ind, below = shifted_grid([4, 4, 4], 4.4)
    n, ptr, idx, val = ind
    s = CH.Symmetric(CH.UpperColumns.from_symmetric(*ind), kind="auto")
    assert s.indefinite and s.inertia()[1] == below
    x = np.cos(np.arange(n, dtype=float))
    b = O.csr_mul(n, ptr, idx, val, x)
    assert np.abs(s.solve_refined(n, ptr, idx, val, b, 3) - x).max() < 1e-12
    with pytest.raises(CH.NotPositiveDefinite):
        CH.Symmetric(CH.UpperColumns.from_symmetric(*ind), kind="definite")


def test_strongly_indefinite_stiffness_matches_the_oracle_bit_for_bit(gpu):
    """The 9 720-equation hex8 cantilever stiffness less 0.618 of its median diagonal: about half
    the pivots negative and hundreds of 2 x 2 pivots across panels and supernodes."""
    from fieldcraft_amd import problems as P

    p = P.cantilever_hex8()
    diag = p.idx == np.repeat(np.arange(p.n), np.diff(p.ptr))
    vals = p.kval - 0.6180339887 * float(np.median(p.kval[diag])) * diag
    up = p.upper(vals)
    f = CH.LdltFactor.new(CH.UpperColumns(p.n, *up))
    o = O.LdltFactor(p.n, *up)
    assert f.inertia() == o.inertia() and f.inertia()[1] > p.n // 3
    assert f.two_by_two() == o.two_by_two() > 100
    assert np.array_equal(f.values().view(np.uint64), o.values().view(np.uint64))
    x = f.solve(p.rhs)
    assert np.array_equal(x.view(np.uint64), o.solve(p.rhs).view(np.uint64))
    assert np.abs(p.mul(x, vals) - p.rhs).max() < 1e-10 * np.abs(p.rhs).max()


@pytest.mark.parametrize("rows", ["16", "100"])
def test_backward_solve_in_chunks_gives_the_same_bits(gpu, monkeypatch, rows):
    """The backward pass stages at most FC_LDLT_SOLVE_ROWS rows of a column in LDS: supernodes
    with more rows take the chunked branch, whose sums must be the same bits."""
    a, _ = shifted_grid([14, 12, 10], 2.7)
    n, ptr, idx, val = a
    f = device_factor(a)
    b = np.cos(np.arange(n) * 0.3)
    whole = f.solve(b)
    monkeypatch.setenv("FC_LDLT_SOLVE_ROWS", rows)
    chunked = f.solve(b)
    assert np.array_equal(whole.view(np.uint64), chunked.view(np.uint64))
    o = O.LdltFactor(n, *O.from_symmetric(n, ptr, idx, val))
    assert np.array_equal(chunked.view(np.uint64), o.solve(b).view(np.uint64))


@pytest.mark.parametrize("kind", ["real", "complex"])
def test_panel_work_columns_in_device_memory_give_the_same_bits(gpu, monkeypatch, kind):
    """The panel kernels keep a supernode's two work columns in LDS when they fit
    (FC_LDLT_WORK_LDS bytes) and in device memory otherwise. With the budget at 0 every launch
    takes the device-memory branch (production-size supernodes do: two columns of 3072 rows or
    more); its factor and solution must be the oracle's bits, real and complex, on a grid whose
    supernodes are several panels wide."""
    from test_zldlt_oracle import complex_grid, upper as zupper

    monkeypatch.setenv("FC_LDLT_WORK_LDS", "0")
    if kind == "real":
        a, _ = shifted_grid([16, 16, 14], 4.1)
        n, ptr, idx, val = a
        f = device_factor(a)
        o = O.LdltFactor(n, *O.from_symmetric(n, ptr, idx, val))
        b = np.cos(np.arange(n) * 0.3)
        assert f.inertia() == o.inertia()
    else:
        a = complex_grid([16, 16, 14], 4.1)
        n = a[0]
        f = CH.ZLdltFactor.new(n, *zupper(a))
        o = O.ZLdltFactor(n, *zupper(a))
        b = np.cos(np.arange(n) * 0.3) + 0.5j * np.sin(np.arange(n) * 0.7)
    assert f.two_by_two() == o.two_by_two()
    assert np.array_equal(np.ascontiguousarray(f.values()).view(np.uint64),
                          np.ascontiguousarray(o.values()).view(np.uint64))
    assert np.array_equal(np.ascontiguousarray(f.solve(b)).view(np.uint64),
                          np.ascontiguousarray(o.solve(b)).view(np.uint64))
