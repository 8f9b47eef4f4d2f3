"""The Lanczos oracle (oracle/eigen.py, the restatement of crates/solver/src/eigen.rs) pinned by
the reference's own known answers (eigen.rs:470-563: a diagonal pencil with six zero, two unit and
then evenly spaced eigenvalues, a range query, the basis cap's error, a problem smaller than one
block) and by library routines: numpy's symmetric eigensolver for the dense part, scipy's
shift-invert ARPACK (eigsh) for the generalized eigenvalues of a tet10 stiffness and mass."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from fieldcraft_amd import problems as P
from oracle import eigen as EG
from oracle import oracle as O


def diagonal_op(k, m, shift):
    k, m = np.asarray(k, float), np.asarray(m, float)
    return EG.Operator(len(k), lambda x: m * x, lambda X: X / (k - shift * m)[:, None])


def test_random_matches_xorshift64star():
    # first outputs of xorshift64* from seed 7, computed with explicit 64-bit arithmetic
    r = EG.Random(7)
    s = 7
    for _ in range(5):
        s ^= s >> 12
        s ^= (s << 25) % 2 ** 64
        s ^= s >> 27
        v = (s * 0x2545F4914F6CDD1D) % 2 ** 64
        assert r.uniform() == (v >> 11) / 2.0 ** 52 - 1.0
    assert all(-1.0 <= EG.Random(3).uniform() < 1.0 for _ in range(100))


@pytest.mark.parametrize("n", [1, 2, 9, 40])
def test_symmetric_eigen_matches_numpy(n):
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, (n, n))
    a = (a + a.T) / 2
    values, vectors = EG.symmetric_eigen(a)
    assert np.all(np.diff(values) >= 0)
    assert np.allclose(values, np.linalg.eigvalsh(a), rtol=0, atol=1e-12)
    assert np.abs(a @ vectors - vectors * values).max() < 1e-12
    assert np.abs(vectors.T @ vectors - np.eye(n)).max() < 1e-12


def test_reference_diagonal_pencil():
    """eigen.rs:514-562."""
    n = 200
    lambdas = np.array([0.0 if i < 6 else 1.0 if i < 8 else 1.0 + (i - 7) * 0.5 for i in range(n)])
    m = 1.0 + (np.arange(n) % 7) * 0.1
    k = lambdas * m
    shift = -1e-3
    op = diagonal_op(k, m, shift)
    pairs = EG.lanczos(op, shift, 10, block=6, max_basis=n, tolerance=1e-9)
    want = [0, 0, 0, 0, 0, 0, 1.0, 1.0, 1.5, 2.0]
    assert len(pairs["values"]) == 10
    assert np.abs(pairs["values"] - want).max() < 1e-8
    x = pairs["vectors"]
    assert np.abs(x.T @ (m[:, None] * x) - np.eye(10)).max() < 1e-8
    ranged = EG.lanczos(op, shift, 100, lo=1.2, hi=3.2, block=6, max_basis=n, tolerance=1e-9)
    assert np.abs(ranged["values"] - [1.5, 2.0, 2.5, 3.0]).max() < 1e-8
    with pytest.raises(EG.NotConverged):
        EG.lanczos(op, shift, 50, block=6, max_basis=24, tolerance=1e-9)
    small = diagonal_op([2.0, 1.0, 3.0], [1.0] * 3, 0.0)
    allv = EG.lanczos(small, 0.0, 5, block=6, max_basis=100, tolerance=1e-9)["values"]
    assert len(allv) == 3 and abs(allv[0] - 1.0) < 1e-12 and abs(allv[2] - 3.0) < 1e-12


def tet10_pencil(cells=(3, 2, 2)):
    p = P.block("tet10", cells, (0.3, 0.2, 0.2), 210e9, 0.3, clamp_face=4, jitter=0.1, seed=5, rho=7850.0)
    K = sp.csr_matrix((p.kval, p.idx, p.ptr), shape=(p.n, p.n))
    M = sp.csr_matrix((p.mval, p.idx, p.ptr), shape=(p.n, p.n))
    return p, K, M


def oracle_op(p, sigma):
    up = p.upper(p.kval - sigma * p.mval)
    f = O.Factor(p.n, *up)
    M = sp.csr_matrix((p.mval, p.idx, p.ptr), shape=(p.n, p.n))
    return EG.Operator(p.n, lambda x: M @ x, lambda X: np.stack([f.solve(X[:, c]) for c in range(X.shape[1])], 1))


def test_tet10_modes_match_arpack():
    p, K, M = tet10_pencil()
    sigma = -1.0e3
    pairs = EG.lanczos(oracle_op(p, sigma), sigma, 8, block=6, max_basis=p.n, tolerance=1e-9)
    ref = np.sort(sla.eigsh(K.tocsc(), k=8, M=M.tocsc(), sigma=sigma, which="LM", tol=1e-13)[0])
    assert np.abs(pairs["values"] - ref).max() <= 1e-9 * ref.max()
    x = pairs["vectors"]
    res = K @ x - (M @ x) * pairs["values"]
    assert np.abs(res).max() <= 1e-6 * np.abs(K @ x).max()
    assert np.abs(x.T @ (M @ x) - np.eye(8)).max() < 1e-8
