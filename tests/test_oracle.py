"""The oracle pinned against the reference's known answers, against scipy's dense Cholesky and
numpy's solve (golden fixtures), and against its own invariants. CPU only."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

from matrices import dense_to_csr, grid_7x5, singular_laplacian, two_grids_and_separator, upper

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
KNOWN = json.load(open(os.path.join(GOLDEN, "known_answers.json")))


def check_solve(n, ptr, idx, val, f, tol):
    """check_solve (cholesky.rs:1680-1689)."""
    x = np.sin(np.arange(n) * 0.37)
    b = O.csr_mul(n, ptr, idx, val, x)
    assert np.abs(f.solve(b) - x).max() < tol


@pytest.mark.parametrize("case", KNOWN["postorder"])
def test_postorder_known_answers(case):
    parent = [-1 if p is None else p for p in case["parent"]]
    assert O.postorder(parent).tolist() == case["post"]


def test_one_by_one_and_empty():
    n, ptr, idx, val = dense_to_csr([[KNOWN["one_by_one"]["a"]]])
    f = O.Factor(n, *upper(n, ptr, idx, val), max_entries=10)
    assert f.solve(np.array([KNOWN["one_by_one"]["b"]])).tolist() == [KNOWN["one_by_one"]["x"]]
    e = O.Factor(0, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0), max_entries=10)
    assert e.entries() == KNOWN["empty"]["entries"]


def test_grid_7x5_solve_cap_and_singular_column():
    n, ptr, idx, val = dense_to_csr(grid_7x5())
    up = upper(n, ptr, idx, val)
    check_solve(n, ptr, idx, val, O.Factor(n, *up), KNOWN["grid_7x5"]["solve_error_below"])
    with pytest.raises(O.OracleError) as e:
        O.Factor(n, *up, max_entries=KNOWN["grid_7x5"]["too_large_cap"])
    assert e.value.kind == O.TOO_LARGE
    n2, p2, i2, v2 = dense_to_csr(singular_laplacian())
    with pytest.raises(O.OracleError) as e:
        O.Factor(n2, *upper(n2, p2, i2, v2))
    assert e.value.kind == O.NOT_POSITIVE_DEFINITE and e.value.column == KNOWN["grid_7x5"]["singular_column"]


def test_two_grids_and_separator():
    n, ptr, idx, val = two_grids_and_separator()
    f = O.Factor(n, *upper(n, ptr, idx, val))
    check_solve(n, ptr, idx, val, f, KNOWN["two_grids_separator"]["solve_error_below"])
    s = f.symbolic()
    assert np.diff(s["first"]).max() > KNOWN["two_grids_separator"]["widest_supernode_above"]
    # several right-hand sides at once, each the same bit for bit as alone (cholesky.rs:1816-1828)
    rhs = np.stack([np.sin((np.arange(n) * 7 + k * 13) * 0.11) for k in range(5)], axis=1)
    together = f.solve(rhs)
    for k in range(5):
        assert np.array_equal(together[:, k].view(np.uint64), f.solve(rhs[:, k]).view(np.uint64))


def test_factor_is_the_same_on_any_thread_count():
    """cholesky.rs:1830-1833: the factor's bits do not depend on the number of threads."""
    n, ptr, idx, val = two_grids_and_separator()
    up = upper(n, ptr, idx, val)
    bits = []
    for t in (1, 3, 8):
        O.set_threads(t)
        bits.append(O.Factor(n, *up).values().view(np.uint64))
    O.set_threads(os.cpu_count() or 1)
    assert all(np.array_equal(bits[0], b) for b in bits[1:])


@pytest.mark.parametrize("name", ["grid_7x5", "random_spd_60", "random_spd_150"])
def test_golden_dense_cholesky(name):
    """Factor values against scipy.linalg.cholesky(P A P^T) and solutions against
    numpy.linalg.solve (tests/golden/make_golden.py)."""
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    n = int(g["n"])
    up = upper(n, g["ptr"], g["idx"], g["val"])
    f = O.Factor(n, *up)
    assert np.array_equal(f.symbolic()["perm"], g["perm"])
    v = f.values()
    assert np.abs(v - g["values"]).max() <= 1e-13 * np.abs(g["values"]).max()
    # every nonzero of the dense factor lies in the supernodal structure
    assert np.isclose(np.sum(g["values"] ** 2), np.sum(g["dense_l"] ** 2), rtol=1e-12)
    x = f.solve(g["b"])
    assert np.linalg.norm(x - g["x"]) <= 1e-12 * np.linalg.norm(g["x"])


def test_a_planted_mistake_is_caught():
    """The golden comparison fails for a factor with one wrong entry (the check pins values)."""
    g = np.load(os.path.join(GOLDEN, "random_spd_60.npz"))
    n = int(g["n"])
    v = O.Factor(n, *upper(n, g["ptr"], g["idx"], g["val"])).values().copy()
    v[np.argmax(np.abs(v))] *= 1 + 1e-12
    assert np.abs(v - g["values"]).max() > 1e-13 * np.abs(g["values"]).max()


def test_csr_mul_closed_form():
    """Csr::mul (sparse.rs:42-50) on the 3-node example of sparse.rs:438-447."""
    # two elements [0,1] and [1,2] of 2 on the diagonal and -1 coupling in each direction
    k = np.zeros((9, 9))
    for a, b in [(0, 1), (1, 2)]:
        for d in range(3):
            i, j = 3 * a + d, 3 * b + d
            k[i, i] += 2.0
            k[j, j] += 2.0
            k[i, j] -= 1.0
            k[j, i] -= 1.0
    n, ptr, idx, val = dense_to_csr(k)
    y = O.csr_mul(n, ptr, idx, val, np.ones(9))
    assert y.tolist() == [1.0] * 3 + [2.0] * 3 + [1.0] * 3


def test_cholesky_solution_bits_match_the_reference_fingerprint():
    """ldlt.rs:809-816 freezes the Cholesky solution of shifted_grid([5, 5, 5], -0.5) with the exact
    right-hand side A x, x_i = ((37 i mod 101) - 50) / 50, as an FNV-1a fingerprint of its bits:
    the oracle's analysis, factorization and solve reproduce the reference bit for bit."""
    from matrices import fingerprint, frozen_rhs, shifted_grid

    (n, ptr, idx, val), _ = shifted_grid([5, 5, 5], -0.5)
    f = O.Factor(n, *O.from_symmetric(n, ptr, idx, val))
    b = frozen_rhs(n, lambda x: O.csr_mul(n, ptr, idx, val, x))
    assert fingerprint(f.solve(b)) == 15069623423640047316
