"""The symbolic phase: the oracle against brute-force elimination on small graphs, the library's
host symbolic phase (fc_analyze) against the oracle bit for bit, and the reference's invariants.
CPU only (the symbolic phase runs on the host)."""
import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P
from oracle import oracle as O

from matrices import dense_to_csr, grid_7x5, random_spd, two_grids_and_separator, upper


def brute_force(a):
    """Boolean dense elimination of the pattern of a (n x n): the structure of L, the
    elimination tree (parent = first row below the diagonal in the column) and the counts."""
    n = a.shape[0]
    s = a != 0
    for k in range(n):
        rows = np.nonzero(s[k + 1:, k])[0] + k + 1
        for i in rows:
            s[i, rows] = True
            s[rows, i] = True
    lower = np.tril(s)
    parent = np.array([(np.nonzero(lower[j + 1:, j])[0][0] + j + 1) if lower[j + 1:, j].any() else -1 for j in range(n)])
    counts = lower.sum(axis=0)
    return lower, parent, counts


def random_pattern(n, density, seed):
    rng = np.random.default_rng(seed)
    a = (rng.random((n, n)) < density).astype(float)
    a = np.triu(a, 1)
    a = a + a.T
    a[np.arange(n), np.arange(n)] = n + 1.0
    return a


@pytest.mark.parametrize("seed", range(6))
def test_tree_counts_and_structure_by_brute_force(seed):
    a = random_pattern(40 + 7 * seed, 0.06, seed)
    n, ptr, idx, val = dense_to_csr(a)
    up = upper(n, ptr, idx, val)
    lower, parent, counts = brute_force(a)
    p0, c0 = O.symbolic(n, up[0], up[1], by_columns=True)
    assert np.array_equal(p0, parent) and np.array_equal(c0, counts)
    s = O.Factor(n, *up, analyze_only=True).symbolic()
    assert s["entries"] == counts.sum()
    # in postorder: each supernode's rows are its first column's structure, and its columns
    # nest (fundamental supernodes, cholesky.rs:1236-1253)
    perm = s["perm"]
    lp = lower[np.ix_(perm, perm)]
    for q in range(len(s["first"]) - 1):
        f, e = s["first"][q], s["first"][q + 1]
        rows = s["rows"][s["rptr"][q]:s["rptr"][q + 1]]
        assert np.array_equal(rows, np.nonzero(lp[:, f])[0])
        for c in range(f, e):
            assert np.array_equal(np.nonzero(lp[:, c])[0], rows[c - f:])


def analyses(n, ptr, idx, val):
    up = upper(n, ptr, idx, val)
    so = O.Factor(n, *up, analyze_only=True).symbolic()
    sp = CH.Symbolic(CH.UpperColumns(n, *up)).export()
    return so, sp


def assert_same(so, sp):
    for key in ["perm", "first", "rptr", "rows", "uptr", "updates", "tasks", "top"]:
        assert np.array_equal(so[key], sp[key]), key
    assert so["entries"] == sp["entries"]


MATRICES = {
    "grid_7x5": lambda: dense_to_csr(grid_7x5()),
    "two_grids": two_grids_and_separator,
    "random_spd": lambda: dense_to_csr(random_spd(200, 0.02, 9)),
    "random_pattern": lambda: dense_to_csr(random_pattern(300, 0.01, 4)),
}


@pytest.mark.parametrize("name", list(MATRICES))
def test_library_symbolic_is_the_oracle_bit_for_bit(name):
    assert_same(*analyses(*MATRICES[name]()))


@pytest.mark.parametrize("cells", [(40, 8, 8), (12, 12, 12)])
def test_library_symbolic_on_hex8_blocks(cells):
    p = P.block("hex8", cells, (1.0, 1.0, 1.0), 210e9, 0.3, clamp_face=0)
    so, sp = analyses(p.n, p.ptr, p.idx, p.kval)
    assert_same(so, sp)


def test_tet10_block_symbolic():
    p = P.block("tet10", (4, 3, 5), (1.0, 0.75, 1.25), 210e9, 0.3, clamp_face=4, jitter=0.1, seed=3)
    assert_same(*analyses(p.n, p.ptr, p.idx, p.kval))


def test_uniform_blocks_equal_columnwise():
    """the_structure_of_uniform_blocks_is_the_structure_by_columns (cholesky.rs:2021-2062):
    the blocks' analysis equals the columnwise one; a ragged matrix is taken by columns."""
    nodes, B = 40, 3
    seed = [7]

    def nxt():
        seed[0] = (seed[0] * 6364136223846793005 + 1442695040888963407) % 2**64
        return seed[0] >> 33

    dense = np.zeros((B * nodes, B * nodes))

    def couple(a, b):
        for i in range(B):
            for j in range(B):
                dense[B * a + i, B * b + j] = 100.0 if (a == b and i == j) else 0.5
                dense[B * b + j, B * a + i] = dense[B * a + i, B * b + j]

    for a in range(nodes):
        couple(a, a)
        for _ in range(3):
            b = nxt() % nodes
            couple(max(a, b), min(a, b))
    n, ptr, idx, val = dense_to_csr(dense)
    up = upper(n, ptr, idx, val)
    f = O.Factor(n, *up, analyze_only=True)
    assert f.by_blocks()
    pb, cb = O.symbolic(n, up[0], up[1])
    pc, cc = O.symbolic(n, up[0], up[1], by_columns=True)
    assert np.array_equal(pb, pc) and np.array_equal(cb, cc)
    assert CH.Symbolic(CH.UpperColumns(n, *up)).info()["by_blocks"] == 1
    assert_same(*analyses(n, ptr, idx, val))
    other = dense.copy()
    a_, b_ = B * (nodes - 1), 1
    other[a_, b_] = other[b_, a_] = 0.5 if other[a_, b_] == 0 else 0.0
    n2, p2, i2, v2 = dense_to_csr(other)
    up2 = upper(n2, p2, i2, v2)
    assert not O.Factor(n2, *up2, analyze_only=True).by_blocks()
    assert CH.Symbolic(CH.UpperColumns(n2, *up2)).info()["by_blocks"] == 0
    assert_same(*analyses(n2, p2, i2, v2))


def test_invariants_of_the_structure():
    """rows(s) holds count(first(s)) rows (cholesky.rs:1306); the updates of each descendant
    partition its rows below its columns exactly once; tasks and top cover the supernodes in
    order, tasks being whole subtrees."""
    p = P.cube_hex8(10)
    up = p.upper()
    s = CH.Symbolic(CH.UpperColumns(p.n, *up)).export()
    _, counts0 = O.symbolic(p.n, up[0], up[1])
    perm = s["perm"]
    counts = counts0[perm]
    first, rptr, rows = s["first"], s["rptr"], s["rows"]
    ns = len(first) - 1
    assert all(rptr[q + 1] - rptr[q] == counts[first[q]] for q in range(ns))
    covered = [[] for _ in range(ns)]
    for j in range(ns):
        for k, p1, p2 in s["updates"][s["uptr"][j]:s["uptr"][j + 1]]:
            covered[k].append((p1, p2))
            assert np.all((rows[rptr[k] + p1:rptr[k] + p2] >= first[j]) & (rows[rptr[k] + p1:rptr[k] + p2] < first[j + 1]))
    for k in range(ns):
        runs = sorted(covered[k])
        w = first[k + 1] - first[k]
        pos = w
        for p1, p2 in runs:
            assert p1 == pos
            pos = p2
        assert pos == rptr[k + 1] - rptr[k]
    seen = []
    for a, b in s["tasks"]:
        seen.extend(range(a, b))
    seen.extend(s["top"].tolist())
    assert sorted(seen) == list(range(ns))


def test_too_large_and_malformed_input():
    n, ptr, idx, val = dense_to_csr(grid_7x5())
    up = upper(n, ptr, idx, val)
    with pytest.raises(CH.TooLarge) as e:
        CH.Symbolic(CH.UpperColumns(n, *up), max_entries=10)
    assert e.value.entries == 189 and e.value.limit == 10  # the oracle's count for this grid
    bad = (up[0].copy(), up[1].copy(), up[2])
    bad[1][1], bad[1][2] = bad[1][2], bad[1][1]
    with pytest.raises(CH.InvalidArgument):
        CH.Symbolic(CH.UpperColumns(n, *bad))


def test_amalgamated_fronts_cover_the_supernodes():
    """The execution plan's fronts are chains of whole supernodes in order, and its flop count
    is at least the supernodes' (padding only adds)."""
    p = P.cube_hex8(16)
    up = p.upper()
    info = CH.Symbolic(CH.UpperColumns(p.n, *up)).info()
    assert info["fronts"] <= info["supernodes"]
    assert info["front_flops"] >= info["flops"]
    assert info["stored"] >= info["entries"]
    assert abs(sum(info["kind_flops"].values()) - info["front_flops"]) <= 1e-6 * info["front_flops"]


def test_batch_symbolic_is_the_union_of_the_models():
    """Batching keeps the symbolic phase bit-exact: the analysis of the block-diagonal batch is
    each model's own analysis (ordering, supernodes, row structures, update maps) side by side,
    with column, supernode, row and update offsets."""
    models = [P.random_model(s, 1.5e3, 5e3) for s in (3, 4, 5)]
    b = P.batch(models, nrhs=1)
    whole = CH.Symbolic(CH.UpperColumns(b.n, *b.upper())).export()
    parts = [CH.Symbolic(CH.UpperColumns(m.n, *m.upper())).export() for m in models]
    col = sn = nrow = nupd = 0
    want = {k: [] for k in ("perm", "first", "rptr", "rows", "uptr", "updates")}
    for e in parts:
        ns = len(e["first"]) - 1
        want["perm"].append(e["perm"] + col)
        want["first"].append(e["first"][:-1] + col)
        want["rptr"].append(e["rptr"][:-1] + nrow)
        want["rows"].append(e["rows"] + col)
        want["uptr"].append(e["uptr"][:-1] + nupd)
        want["updates"].append(e["updates"] + np.array([sn, 0, 0]))
        col += len(e["perm"])
        sn += ns
        nrow += len(e["rows"])
        nupd += len(e["updates"])
    want["first"].append(np.array([col]))
    want["rptr"].append(np.array([nrow]))
    want["uptr"].append(np.array([nupd]))
    for k, v in want.items():
        np.testing.assert_array_equal(whole[k], np.concatenate(v), err_msg=k)
    assert whole["entries"] == sum(e["entries"] for e in parts)
