"""Writes the golden fixtures of tests/golden/ (run in the build container; the outputs are
committed). Test infrastructure.

The reference (Rust) holds no vectors for this path beyond the known answers its own unit tests
assert (cholesky.rs:1696-1737, 2201-2206; ordering.rs:288-312, 381-395), which are restated in
known_answers.json. The factor values and solutions are pinned by a library routine instead:
for each matrix A, L = scipy.linalg.cholesky(P A P^T, lower=True) with P the supernodal
postorder, and x = numpy.linalg.solve(A, b). The supernodal structure used to lay L out in the
reference's trapezoid layout (cholesky.rs:393-420) is the symbolic phase under test; the values
at those positions come from scipy alone, and the fixture also stores L's dense lower triangle
so a structure mistake shows up as a missing entry.

Usage: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from matrices import dense_to_csr, grid_7x5, random_spd, upper  # noqa: E402
from oracle import oracle as O  # noqa: E402


def fixture(name, a, seed):
    n, ptr, idx, val = dense_to_csr(a)
    up = upper(n, ptr, idx, val)
    f = O.Factor(n, *up, analyze_only=True)
    s = f.symbolic()
    perm = s["perm"]
    pa = a[np.ix_(perm, perm)]
    Ld = scipy.linalg.cholesky(pa, lower=True)
    # the reference layout: supernodes in order, each column c holding rows c..m of its rows
    vals = []
    for q in range(len(s["first"]) - 1):
        rows = s["rows"][s["rptr"][q]:s["rptr"][q + 1]]
        for k, c in enumerate(range(s["first"][q], s["first"][q + 1])):
            vals.extend(Ld[rows[k:], c])
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(n)
    x = np.linalg.solve(a, b)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), n=n, ptr=ptr, idx=idx, val=val, perm=perm,
                        values=np.array(vals), dense_l=Ld[np.tril_indices(n)], b=b, x=x)


def main():
    fixture("grid_7x5", grid_7x5(), 1)
    fixture("random_spd_60", random_spd(60, 0.08, 5), 2)
    fixture("random_spd_150", random_spd(150, 0.03, 6), 3)
    known = {
        "source": "crates/solver/src/cholesky.rs and ordering.rs unit tests",
        "postorder": [  # cholesky.rs:2201-2206 (None = root)
            {"parent": [2, 2, 4, 4, None, None], "post": [0, 1, 2, 3, 4, 5]},
            {"parent": [4, 3, 3, 4, None], "post": [0, 1, 2, 3, 4]},
            {"parent": [None, 0], "post": [1, 0]},
            {"parent": [3, None, 1, None], "post": [2, 1, 0, 3]},
        ],
        "one_by_one": {"a": 4.0, "b": 2.0, "x": 0.5},  # cholesky.rs:1734-1737
        "grid_7x5": {"solve_error_below": 1e-13, "too_large_cap": 10,  # cholesky.rs:1710-1713
                     "singular_column": 34},  # cholesky.rs:1714-1720 (n - 1)
        "two_grids_separator": {"solve_error_below": 1e-12, "widest_supernode_above": 128},  # :1813-1815
        "empty": {"entries": 0},  # cholesky.rs:1729-1733
        "grid_12x12": {"top_separator_length": 12},  # ordering.rs:288-312
    }
    with open(os.path.join(HERE, "known_answers.json"), "w") as fh:
        json.dump(known, fh, indent=1)


if __name__ == "__main__":
    main()
