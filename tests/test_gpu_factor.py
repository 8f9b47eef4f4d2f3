"""GPU parity of the numeric phase and solves against the oracle (the C restatement of
cholesky.rs), through the C ABI. Tolerances: FP64 factor values and solutions within 1e-12
(relative, normwise) of the oracle's; the reference's own solve-error bounds where it states
them."""
import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P
from oracle import oracle as O

from matrices import dense_to_csr, grid_7x5, random_spd, singular_laplacian, two_grids_and_separator, upper

pytestmark = pytest.mark.gpu

REL = 1e-12  # FP64 parity bound: factor values and solutions vs the oracle


def both(n, ptr, idx, val, precision="fp64"):
    up = upper(n, ptr, idx, val)
    ref = O.Factor(n, *up)
    sym = CH.Symbolic(CH.UpperColumns(n, *up))
    f = CH.Factor(sym, up[2], precision=precision)
    return ref, f, up


def rel(a, b):
    d = np.linalg.norm(np.asarray(a) - np.asarray(b))
    s = np.linalg.norm(np.asarray(b))
    return d / s if s > 0 else d


def check_solve(n, ptr, idx, val, f, tol):
    """check_solve (cholesky.rs:1680-1689): x = sin(0.37 i), b = A x, max |solve(b) - x| < tol."""
    x = np.sin(np.arange(n) * 0.37)
    b = O.csr_mul(n, ptr, idx, val, x)
    err = np.abs(f.solve(b) - x).max()
    assert err < tol, err


def test_grid_7x5(gpu):
    n, ptr, idx, val = dense_to_csr(grid_7x5())
    ref, f, _ = both(n, ptr, idx, val)
    check_solve(n, ptr, idx, val, f, 1e-13)
    assert rel(f.values(), ref.values()) <= REL
    assert f.entries() == ref.entries()


def test_one_by_one_and_empty(gpu):
    n, ptr, idx, val = dense_to_csr([[4.0]])
    _, f, _ = both(n, ptr, idx, val)
    assert f.solve(np.array([2.0])).tolist() == [0.5]  # cholesky.rs:1734-1737
    up = (np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0))
    sym = CH.Symbolic(CH.UpperColumns(0, *up), max_entries=10)
    e = CH.Factor(sym, up[2])
    assert e.entries() == 0
    assert e.solve(np.zeros(0)).shape == (0,)


def test_not_positive_definite_names_the_column(gpu):
    n, ptr, idx, val = dense_to_csr(singular_laplacian())
    up = upper(n, ptr, idx, val)
    with pytest.raises(CH.NotPositiveDefinite) as e:
        CH.Factor.new(CH.UpperColumns(n, *up))
    assert e.value.column == n - 1  # cholesky.rs:1714-1720
    with pytest.raises(O.OracleError) as o:
        O.Factor(n, *up)
    assert o.value.column == e.value.column


def test_two_grids_and_dense_separator(gpu):
    n, ptr, idx, val = two_grids_and_separator()
    ref, f, _ = both(n, ptr, idx, val)
    check_solve(n, ptr, idx, val, f, 1e-12)  # cholesky.rs:1815
    assert rel(f.values(), ref.values()) <= REL
    rhs = [np.sin((np.arange(n) * 7 + s * 13) * 0.11) for s in range(5)]
    alone = [f.solve(b) for b in rhs]
    together = f.solve_many(rhs)
    for a, t in zip(alone, together):
        assert np.array_equal(a.view(np.uint64), t.view(np.uint64))  # cholesky.rs:1816-1828
    for a, b in zip(alone, rhs):
        assert rel(a, ref.solve(b)) <= REL


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_spd_against_dense(gpu, seed):
    a = random_spd(120, 0.05, seed)
    n, ptr, idx, val = dense_to_csr(a)
    ref, f, _ = both(n, ptr, idx, val)
    b = np.cos(np.arange(n) * 0.3)
    x = f.solve(b)
    assert rel(x, np.linalg.solve(a, b)) <= 1e-12
    assert rel(x, ref.solve(b)) <= REL
    assert rel(f.values(), ref.values()) <= REL


def test_cantilever_hex8_matches_the_oracle(gpu):
    p = P.cantilever_hex8()
    ref, f, up = both(p.n, p.ptr, p.idx, p.kval)
    assert rel(f.values(), ref.values()) <= REL
    x, xr = f.solve(p.rhs), ref.solve(p.rhs)
    assert rel(x, xr) <= REL
    r = p.mul(x) - p.rhs
    berr = np.abs(r).max() / (p.norm_inf() * np.abs(x).max() + np.abs(p.rhs).max())
    assert berr < 1e-14


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cube_matches_the_oracle(gpu, precision):
    """A 16^3 hex8 cube (14 688 equations): many levels, wide fronts (128-column panels, split
    products). FP64 within 1e-12 of the oracle; FP32 within its own precision."""
    p = P.cube_hex8(16)
    ref, f, up = both(p.n, p.ptr, p.idx, p.kval, precision=precision)
    x, xr = f.solve(p.rhs), ref.solve(p.rhs)
    if precision == "fp64":
        assert rel(f.values(), ref.values()) <= REL
        assert rel(x, xr) <= REL
    else:
        assert rel(f.values(), ref.values()) <= 1e-5
        assert rel(x, xr) <= 1e-3


def test_refined_solve_reaches_the_target(gpu):
    p = P.cantilever_hex8()
    up = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    A = CH.Matrix(p.n, p.ptr, p.idx, p.kval)
    for prec in ["fp64", "fp32"]:
        f = CH.Factor(sym, up[2], precision=prec)
        x, rep = f.solve_refined(A, p.rhs, target_berr=1e-12)
        assert rep["converged"], rep
        r = p.mul(x) - p.rhs
        berr = np.abs(r).max() / (p.norm_inf() * np.abs(x).max() + np.abs(p.rhs).max())
        assert berr <= 1e-12, (prec, berr, rep)


@pytest.mark.parametrize("terms,bound", [(3, 1e-4), (1, 3e-3)])
def test_mixed_factor_is_close_to_the_oracle(gpu, terms, bound):
    """The mixed form on a 16^3 cube: BF16 Schur products (three products: about 16 significant
    bits; one: 8) keep the factor values within `bound` of the oracle's FP64 ones (measured
    1.4e-6 and 2.8e-4 on this cube). A wrong lane layout or scale would be off by O(1)."""
    p = P.cube_hex8(16)
    up = p.upper()
    ref = O.Factor(p.n, *up)
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    f = CH.Factor(sym, up[2], precision="mixed", bf16_terms=terms)
    assert rel(f.values(), ref.values()) <= bound


def test_mixed_rejects_other_term_counts(gpu):
    p = P.cantilever_hex8(4, 2, 2)
    up = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    with pytest.raises(CH.InvalidArgument):
        CH.Factor(sym, up[2], precision="mixed", bf16_terms=2)


def berr_of(p, x, b):
    r = p.mul(x) - b
    return np.abs(r).max() / (p.norm_inf() * np.abs(x).max() + np.abs(b).max())


def test_mixed_refined_solve_reaches_the_target(gpu):
    """Mixed factor + FP64 refinement reaches berr <= 1e-12 on the slender cantilever (its
    condition number makes the single-product form lose definiteness there)."""
    p = P.cantilever_hex8()
    up = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    A = CH.Matrix(p.n, p.ptr, p.idx, p.kval)
    f = CH.Factor(sym, up[2], precision="mixed", bf16_terms=3)
    x, rep = f.solve_refined(A, p.rhs, target_berr=1e-12)
    assert rep["converged"], rep
    assert berr_of(p, x, p.rhs) <= 1e-12
    xr = O.Factor(p.n, *up).solve(p.rhs)
    assert rel(x, xr) <= 1e-8


def test_refined_solve_of_many_columns(gpu):
    """Five load cases refined together: each column reaches the target on its own."""
    import torch
    p = P.cube_hex8(12)
    up = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    A = CH.Matrix(p.n, p.ptr, p.idx, p.kval)
    f = CH.Factor(sym, up[2], precision="fp32")
    rng = np.random.default_rng(7)
    B = np.asfortranarray(rng.standard_normal((p.n, 5)))
    dB = torch.from_numpy(B.T.copy()).cuda()  # row r of dB = column r of B
    dX = torch.zeros_like(dB)
    rep, berr = f.solve_refined_many_device(A, dB.data_ptr(), dX.data_ptr(), 5, target_berr=1e-12)
    assert rep["converged"] and (berr <= 1e-12).all(), (rep, berr)
    X = dX.cpu().numpy().T
    for c in range(5):
        assert berr_of(p, X[:, c], B[:, c]) <= 1e-12


def test_batch_of_models_matches_each_model(gpu):
    """Independent models factored as one block-diagonal matrix: in FP64 each model's block of
    the solution matches the oracle's solve of that model alone (1e-12); in FP32 with
    refinement of 4 load cases per model, every model reaches berr <= 1e-12 measured with its
    own norms."""
    import torch
    models = [P.random_model(s, 1.5e3, 6e3) for s in range(6)]
    bt = P.batch(models, nrhs=4)
    up = bt.upper()
    sym = CH.Symbolic(CH.UpperColumns(bt.n, *up))
    f64 = CH.Factor(sym, up[2], precision="fp64")
    x = f64.solve(bt.rhs[:, 0])
    for s, m in enumerate(models):
        lo, hi = bt.segments[s], bt.segments[s + 1]
        xr = O.Factor(m.n, *m.upper()).solve(bt.rhs[lo:hi, 0])
        assert rel(x[lo:hi], xr) <= REL
    A = CH.Matrix(bt.n, bt.ptr, bt.idx, bt.kval)
    A.set_segments(bt.segments)
    f32 = CH.Factor(sym, up[2], precision="fp32")
    dB = torch.from_numpy(np.ascontiguousarray(bt.rhs.T)).cuda()
    dX = torch.zeros_like(dB)
    rep, berr = f32.solve_refined_many_device(A, dB.data_ptr(), dX.data_ptr(), 4, target_berr=1e-12)
    assert rep["converged"] and (berr <= 1e-12).all(), (rep, berr)
    X = dX.cpu().numpy().T
    for s, m in enumerate(models):
        lo, hi = bt.segments[s], bt.segments[s + 1]
        for c in range(4):
            assert berr_of(m, X[lo:hi, c], bt.rhs[lo:hi, c]) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replays_match_eager_launches(gpu, precision):
    """A factorization captured into a hipGraph and replayed for new values, and solves replayed
    from their own graphs, give the same bits as the eager launches on every replay, with eager
    work in between."""
    p = P.cube_hex8(10)
    up = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    vals = [up[2], up[2] * 2.0, up[2] * 0.5 + 0.0, up[2]]
    g = CH.Factor(sym, vals[0], precision=precision, use_graph=3)  # factor and solves
    for q, v in enumerate(vals):
        if q:
            g.refactor(v)
        e = CH.Factor(sym, v, precision=precision)
        assert np.array_equal(g.values().view(np.uint64), e.values().view(np.uint64)), q
        assert np.array_equal(g.solve(p.rhs).view(np.uint64), e.solve(p.rhs).view(np.uint64)), q
