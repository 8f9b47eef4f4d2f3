"""GPU parity of the shift-invert block Lanczos (fc_lanczos_device through the C ABI) with the
oracle's restatement of crates/solver/src/eigen.rs: the reference's own diagonal pencil (six zero
eigenvalues, a double one, a range, the basis cap's error, a problem smaller than one block), and
a distorted tet10 stiffness and mass assembled and factored on the device against the oracle's
Lanczos over the oracle's factor and against ARPACK. Tolerances: eigenvalues within 1e-10
relative of the oracle's (1e-8 absolute on the reference's pencil, its own bound); eigenvectors
M-orthonormal to 1e-10 and within 1e-6 of the oracle's."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from fieldcraft_amd import assembly as AS
from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P
from oracle import eigen as EG
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def diagonal_factor(k, m, shift):
    n = len(k)
    ptr = np.arange(n + 1)
    idx = np.arange(n)
    sym = CH.Symbolic(CH.UpperColumns(n, ptr, idx, np.zeros(n)))
    f = CH.Factor(sym, np.asarray(k) - shift * np.asarray(m))
    M = CH.Matrix(n, ptr, idx, np.asarray(m, float))
    return f, M


def test_reference_diagonal_pencil(gpu):
    """eigen.rs:514-562 on the device."""
    n = 200
    lambdas = np.array([0.0 if i < 6 else 1.0 if i < 8 else 1.0 + (i - 7) * 0.5 for i in range(n)])
    m = 1.0 + (np.arange(n) % 7) * 0.1
    shift = -1e-3
    f, M = diagonal_factor(lambdas * m, m, shift)
    print("diagonal pencil n=200", flush=True)
    pairs = f.lanczos_device(M, shift, 10, block=6, max_basis=n, tolerance=1e-9)
    assert np.abs(pairs["values"] - [0, 0, 0, 0, 0, 0, 1.0, 1.0, 1.5, 2.0]).max() < 1e-8
    x = pairs["vectors"].cpu().numpy()
    assert np.abs(x.T @ (m[:, None] * x) - np.eye(10)).max() < 1e-8
    ranged = f.lanczos_device(M, shift, 100, lo=1.2, hi=3.2, block=6, max_basis=n, tolerance=1e-9)
    assert np.abs(ranged["values"] - [1.5, 2.0, 2.5, 3.0]).max() < 1e-8
    with pytest.raises(CH.NotConverged):
        f.lanczos_device(M, shift, 50, block=6, max_basis=24, tolerance=1e-9)
    fs, Ms = diagonal_factor([2.0, 1.0, 3.0], [1.0] * 3, 0.0)
    allv = fs.lanczos_device(Ms, 0.0, 5, block=6, max_basis=100, tolerance=1e-9)["values"]
    assert len(allv) == 3 and abs(allv[0] - 1.0) < 1e-12 and abs(allv[2] - 3.0) < 1e-12


def tet10_problem():
    return P.block("tet10", (4, 3, 3), (0.4, 0.3, 0.3), 210e9, 0.3, clamp_face=4, jitter=0.1, seed=5, rho=7850.0)


def test_tet10_modes_on_device_match_the_oracle(gpu):
    """K - sigma M assembled and factored on the device, then Lanczos on the device, against the
    oracle's Lanczos over the oracle's factor of the oracle's assembled values."""
    p = tet10_problem()
    sigma = (2 * np.pi * 50.0) ** 2
    uptr, uidx, _ = p.upper()
    a = AS.Assembly.of_problem(p, uptr, uidx)
    sym = CH.Symbolic(CH.UpperColumns(p.n, uptr, uidx, np.zeros(len(uidx))))
    vals = a.assemble("shifted", sigma)
    f = CH.Factor(sym, device_values=vals.data_ptr())
    mfull = O.assemble_upper("tet10", p.xyz, p.conn, p.young, p.poisson, p.eq_dof, uptr, uidx, density=p.rho, what="M")
    # M by full rows from its upper columns (M's stored pattern is K's)
    Mu = sp.csc_matrix((mfull, uidx, uptr), shape=(p.n, p.n))
    Mrows = (sp.triu(Mu, 1).T + Mu).tocsr()
    Mrows.sort_indices()
    M = CH.Matrix(p.n, Mrows.indptr, Mrows.indices, Mrows.data)
    print("tet10 4x3x3", p.n, flush=True)
    got = f.lanczos_device(M, sigma, 12, block=6, max_basis=600, tolerance=1e-9)
    shifted = O.assemble_upper("tet10", p.xyz, p.conn, p.young, p.poisson, p.eq_dof, uptr, uidx, density=p.rho,
                               what="shifted", shift=sigma)
    of = O.Factor(p.n, uptr, uidx, shifted)
    op = EG.Operator(p.n, lambda x: Mrows @ x, lambda X: np.stack([of.solve(X[:, c]) for c in range(X.shape[1])], 1))
    want = EG.lanczos(op, sigma, 12, block=6, max_basis=600, tolerance=1e-9)
    # the same schedule of convergence checks; an orthogonalization that rounds differently may
    # accept a pair one check later
    assert abs(got["basis"] - want["basis"]) <= max(want["basis"] // 4, 6)
    assert np.abs(got["values"] - want["values"]).max() <= 1e-10 * want["values"].max()
    x = got["vectors"].cpu().numpy()
    assert np.abs(x.T @ (Mrows @ x) - np.eye(12)).max() < 1e-10
    gaps = np.diff(want["values"])
    for c in range(12):
        distinct = (c == 0 or gaps[c - 1] > 1e-6 * want["values"][c]) and (c == 11 or gaps[c] > 1e-6 * want["values"][c])
        if distinct:
            assert np.abs(x[:, c] - want["vectors"][:, c]).max() <= 1e-6 * np.abs(want["vectors"][:, c]).max()
    K = sp.csc_matrix((p.kval, p.idx, p.ptr), shape=(p.n, p.n))
    ref = np.sort(sla.eigsh(K, k=12, M=sp.csc_matrix((p.mval, p.idx, p.ptr), shape=(p.n, p.n)), sigma=sigma,
                            which="LM", tol=1e-12)[0])
    assert np.abs(got["values"] - ref).max() <= 1e-8 * ref.max()


def test_cancel_flag_stops_the_iteration(gpu):
    import ctypes as C

    n = 200
    m = np.ones(n)
    lam = 1.0 + np.arange(n, dtype=float)
    flag = C.c_int32(0)
    ptr, idx = np.arange(n + 1), np.arange(n)
    sym = CH.Symbolic(CH.UpperColumns(n, ptr, idx, np.zeros(n)))
    f = CH.Factor(sym, lam + 1.0, cancel=flag)
    M = CH.Matrix(n, ptr, idx, m)
    flag.value = 1
    with pytest.raises(CH.Cancelled):
        f.lanczos_device(M, -1.0, 5, block=6, max_basis=n)


def test_indefinite_shift_through_ldlt_matches_the_oracle(gpu):
    """A shift between the tet10 block's 4th and 5th modes: K - sigma M is indefinite, so the
    modes come from an LDL^T factor (Symmetric with Kind::Auto under ShiftInvert): the next four
    modes as found ones, and the converged negative Ritz values (some of the four modes below the
    shift, eigen.rs:413-426), against the oracle's Lanczos over its LDL^T solves and scipy's eigsh."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    from fieldcraft_amd import cholesky as CH
    from fieldcraft_amd import problems as P
    from oracle import eigen as EG
    from oracle import oracle as O

    p = P.block("tet10", (3, 2, 2), (0.3, 0.2, 0.2), 210e9, 0.3, clamp_face=4, jitter=0.1, seed=5, rho=7850.0)
    K = sp.csr_matrix((p.kval, p.idx, p.ptr), shape=(p.n, p.n))
    M = sp.csr_matrix((p.mval, p.idx, p.ptr), shape=(p.n, p.n))
    low = np.sort(sla.eigsh(K.tocsc(), k=8, M=M.tocsc(), sigma=-1.0e3, which="LM", tol=1e-13)[0])
    sigma = 0.5 * (low[3] + low[4])
    up = p.upper(p.kval - sigma * p.mval)
    f = CH.LdltFactor.new(CH.UpperColumns(p.n, *up))
    assert f.inertia()[1] == 4
    res = f.lanczos_device(CH.Matrix(p.n, p.ptr, p.idx, p.mval), sigma, 4, block=6, max_basis=p.n)
    of = O.LdltFactor(p.n, *up)
    op = EG.Operator(p.n, lambda x: M @ x, lambda X: np.stack([of.solve(X[:, c]) for c in range(X.shape[1])], 1))
    want = EG.lanczos(op, sigma, 4, block=6, max_basis=p.n, tolerance=1e-9)
    assert np.abs(res["values"] - want["values"]).max() <= 1e-9 * low.max()
    assert np.abs(res["values"] - low[4:8]).max() <= 1e-9 * low.max()
    assert len(want["negative"]) >= 1
    assert len(res["negative"]) == len(want["negative"])
    assert np.abs(np.sort(res["negative"]) - np.sort(want["negative"])).max() <= 1e-9 * low.max()
    for v in res["negative"]:  # each one of the modes below the shift
        assert np.abs(low[:4] - v).min() <= 1e-9 * low.max()
