"""GPU parity of the benchmark workloads' own forms, at sizes the oracle finishes in seconds:

* the plain FP64 factor of the hex8 unit cube at 28^3 cells (70 644 equations, fronts up to
  ~2 400 columns): factor values and solution within 1e-12 of the oracle's (cholesky.rs:1146-1185
  restated), the bound the 1.6M-equation cube is held to;
* the cube's harder variant (E in {210 GPa, 21 MPa} per element, nu = 0.49, same pattern): FP64
  within 1e-12 of the oracle, the mixed form refined to berr <= 1e-12;
* shift-invert: a tet10 steel block's K - sigma M, sigma = (2 pi 50 Hz)^2, factored in the mixed
  form, then 16-vector blocks refined together to berr <= 1e-12 per column, within the
  conditioning's reach of the oracle's FP64 solutions; the FP64 factor's 16-wide block solve
  within 1e-12 of the oracle and bit-equal to single solves (solve_many, cholesky.rs:1568-1570);
* many models with 64 load cases each: the FP32 factor with refinement of all 64 columns, every
  model's own berr <= 1e-12; the FP64 factor's 64-column solve against the oracle per model.
"""
import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-12


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


def berr_cols(n, ptr, idx, val, X, B):
    """Normwise backward error of each column: ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    rows = np.repeat(np.arange(n), np.diff(ptr))
    norm = np.max(np.bincount(rows, np.abs(val), minlength=n))
    out = []
    for c in range(X.shape[1]):
        ax = np.bincount(rows, val * X[idx, c], minlength=n)
        out.append(np.abs(B[:, c] - ax).max() / (norm * np.abs(X[:, c]).max() + np.abs(B[:, c]).max()))
    return np.array(out)


def device_block(B):
    """A host n x k array as a device block, column-major (column c contiguous)."""
    import torch

    return torch.from_numpy(np.ascontiguousarray(B.T)).cuda()


def test_fp64_factor_of_the_28_cube_is_the_oracles(gpu):
    """The plain method at a size where the fronts are wide (the explicit-inverse TRSM with its
    one FP64 refinement step is the risk as fronts widen): every factor entry and the solution
    within 1e-12 (relative, normwise) of the oracle's."""
    O.set_threads(16)
    p = P.cube_hex8(28)
    up = p.upper()
    print(f"cube 28^3: n = {p.n}, oracle factor", flush=True)
    ref = O.Factor(p.n, *up)
    xr = ref.solve(p.rhs)
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    f = CH.Factor(sym, up[2])
    print(f"device factor: {f.seconds():.3f} s, {sym.info()['entries']} entries", flush=True)
    x = f.solve(p.rhs)
    assert rel(x, xr) <= REL
    lv, lr = f.values(), ref.values()
    assert rel(lv, lr) <= REL
    # the largest per-column deviation stays at rounding level too
    assert np.abs(lv - lr).max() <= 1e-10 * np.abs(lr).max()


def test_contrast_variant_fp64_and_mixed(gpu):
    """E contrast 1e4 with nu = 0.49 on the cube's pattern (one analysis for both matrices): FP64
    within 1e-12 of the oracle; the mixed form with FP64 refinement to berr <= 1e-12."""
    p = P.cube_hex8_contrast(16)
    up = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    ref = O.Factor(p.n, *up)
    f = CH.Factor(sym, up[2])
    assert rel(f.values(), ref.values()) <= REL
    xr = ref.solve(p.rhs)
    assert rel(f.solve(p.rhs), xr) <= REL
    A = CH.Matrix(p.n, p.ptr, p.idx, p.kval)
    m = CH.Factor(sym, up[2], precision="mixed")
    x, rep = m.solve_refined(A, p.rhs, target_berr=1e-12)
    assert rep["converged"], rep
    assert berr_cols(p.n, p.ptr, p.idx, p.kval, x[:, None], p.rhs[:, None])[0] <= 1e-12
    print(f"contrast mixed: {rep['iterations']} iterations, rel to FP64 {rel(x, xr):.2e}", flush=True)
    assert rel(x, xr) <= 1e-6  # the matrix's conditioning: berr 1e-12 bounds the forward error at ~1e-6


@pytest.fixture(scope="module")
def shifted_tet10():
    p = P.block_tet10(cells=(8, 8, 8), size=(0.5, 0.5, 0.5))
    sigma = (2 * np.pi * 50.0) ** 2
    vals = p.kval - sigma * p.mval
    up = p.upper(vals)
    rng = np.random.default_rng(20244)
    B = rng.standard_normal((p.n, 16))
    return p, vals, up, B


def test_shift_invert_block_solves_mixed_and_refined(gpu, shifted_tet10):
    """K - sigma M of the tet10 block (first mode far above 50 Hz, so SPD) in the mixed form; one
    16-vector block refined together: every column to berr <= 1e-12, and within 1e-7 of the
    oracle's FP64 solutions."""
    p, vals, up, B = shifted_tet10
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    ref = O.Factor(p.n, *up)
    Xr = np.stack([ref.solve(B[:, c]) for c in range(16)], 1)
    f = CH.Factor(sym, up[2], precision="mixed")
    assert f.precision_used() == "mixed"
    A = CH.Matrix(p.n, p.ptr, p.idx, vals)
    dB = device_block(B)
    dX = dB.clone() * 0
    rep, berr = f.solve_refined_many_device(A, dB.data_ptr(), dX.data_ptr(), 16, target_berr=1e-12)
    X = dX.cpu().numpy().T
    print(f"tet10 8^3 n = {p.n}: {rep['iterations']} iterations, berr {berr.max():.2e}", flush=True)
    assert rep["converged"] and (berr <= 1e-12).all(), (rep, berr)
    assert (berr_cols(p.n, p.ptr, p.idx, vals, X, B) <= 1e-12).all()
    for c in range(16):
        assert rel(X[:, c], Xr[:, c]) <= 1e-7


def test_sixteen_wide_fp64_block_solve_is_the_oracles(gpu, shifted_tet10):
    """The FP64 factor's 16-column block solve (the 16-wide kernels) within 1e-12 of the oracle's
    per-column solves, and each column the same bits as its single solve."""
    p, vals, up, B = shifted_tet10
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    ref = O.Factor(p.n, *up)
    f = CH.Factor(sym, up[2])
    cols = [B[:, c] for c in range(16)]
    many = f.solve_many(cols)
    for c in range(16):
        assert rel(many[c], ref.solve(cols[c])) <= REL
    for c in (0, 7, 15):
        assert np.array_equal(many[c].view(np.uint64), f.solve(cols[c]).view(np.uint64))


def test_sixty_four_load_cases_per_model(gpu):
    """Six models as one block-diagonal factorization with 64 seeded load cases each: the FP64
    factor's 64-column solve per model within 1e-12 of the oracle's, and the FP32 factor with the
    64 columns refined together to every model's own berr <= 1e-12."""
    models = [P.random_model(s, 1.5e3, 6e3) for s in range(6)]
    bt = P.batch(models, nrhs=64)
    up = bt.upper()
    sym = CH.Symbolic(CH.UpperColumns(bt.n, *up))
    f64 = CH.Factor(sym, up[2])
    dB = device_block(bt.rhs)
    dX = dB.clone()
    f64.solve_device(dX.data_ptr(), 64)
    X = dX.cpu().numpy().T
    for s, m in enumerate(models):
        lo, hi = bt.segments[s], bt.segments[s + 1]
        of = O.Factor(m.n, *m.upper())
        for c in (0, 17, 63):
            assert rel(X[lo:hi, c], of.solve(bt.rhs[lo:hi, c])) <= REL
    A = CH.Matrix(bt.n, bt.ptr, bt.idx, bt.kval)
    A.set_segments(bt.segments)
    f32 = CH.Factor(sym, up[2], precision="fp32")
    dX = dB.clone() * 0
    rep, berr = f32.solve_refined_many_device(A, dB.data_ptr(), dX.data_ptr(), 64, target_berr=1e-12)
    assert rep["converged"] and (berr <= 1e-12).all(), (rep, berr)
    X = dX.cpu().numpy().T
    for s, m in enumerate(models):
        lo, hi = bt.segments[s], bt.segments[s + 1]
        b = berr_cols(m.n, m.ptr, m.idx, m.kval, X[lo:hi], bt.rhs[lo:hi])
        assert (b <= 1e-12).all(), (s, b.max())
