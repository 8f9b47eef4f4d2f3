"""Matrices of the reference's own tests, restated (crates/solver/src/cholesky.rs tests and
ordering.rs tests), plus random SPD and graph helpers. Test infrastructure."""
from __future__ import annotations

import numpy as np


def dense_to_csr(a):
    """dense_to_csr (cholesky.rs:1662-1678): nonzeros by rows."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    rows, cols = np.nonzero(a)
    ptr = np.zeros(n + 1, np.int64)
    np.add.at(ptr, rows + 1, 1)
    ptr = np.cumsum(ptr)
    return n, ptr, cols.astype(np.int64), a[rows, cols]


def upper(n, ptr, idx, val):
    """UpperColumns::from_symmetric (cholesky.rs:112-144)."""
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ptr))
    keep = idx <= rows
    counts = np.bincount(rows[keep], minlength=n)
    uptr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=uptr[1:])
    return uptr, np.ascontiguousarray(idx[keep]), np.ascontiguousarray(val[keep])


def grid_7x5(shift: float = 4.5):
    """The 7 x 5 grid matrix of factor_and_solve_a_sparse_spd_matrix (cholesky.rs:1696-1711)."""
    w, h = 7, 5
    n = w * h
    a = np.zeros((n, n))
    for i in range(w):
        for j in range(h):
            k = i * h + j
            a[k, k] = shift
            if i + 1 < w:
                a[k, k + h] = a[k + h, k] = -1.0
            if j + 1 < h:
                a[k, k + 1] = a[k + 1, k] = -1.0
    return a


def singular_laplacian():
    """The same grid with its diagonal the off-diagonal row sum (cholesky.rs:1714-1720)."""
    a = grid_7x5()
    for k in range(a.shape[0]):
        a[k, k] = -(a[k].sum() - a[k, k])
    return a


def two_grids_and_separator():
    """supernodes_panels_and_threads (cholesky.rs:1757-1815): two 8x8x6 grids of 3 coupled
    unknowns, a dense 400-unknown separator ordered last, and a separate 5 x 5 block."""
    nx, ny, nz = 8, 8, 6
    point = lambda i, j, k: i + nx * (j + ny * k)
    per = 3 * nx * ny * nz
    sep = 400
    n = 2 * per + sep + 5
    rows = [dict() for _ in range(n)]

    def pair(a, b, v):
        rows[a][b] = v
        rows[b][a] = v

    for base in (0, per):
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    p = base + 3 * point(i, j, k)
                    for a in range(3):
                        rows[p + a][p + a] = 7.0
                        for b in range(a + 1, 3):
                            pair(p + a, p + b, 0.25)
                    nxt = []
                    if i + 1 < nx:
                        nxt.append(point(i + 1, j, k))
                    if j + 1 < ny:
                        nxt.append(point(i, j + 1, k))
                    if k + 1 < nz:
                        nxt.append(point(i, j, k + 1))
                    for q in nxt:
                        for a in range(3):
                            pair(p + a, base + 3 * q + a, -1.0)
    s0 = 2 * per
    for a in range(sep):
        rows[s0 + a][s0 + a] = 20.0
        for b in range(a + 1, sep):
            pair(s0 + a, s0 + b, 0.5 / (1.0 + (b - a)) ** 2)
        pair(s0 + a, (7 * a) % per, -0.1)
        pair(s0 + a, per + (11 * a) % per, -0.1)
    d0 = s0 + sep
    for a in range(5):
        rows[d0 + a][d0 + a] = 10.0
        for b in range(a + 1, 5):
            pair(d0 + a, d0 + b, 1.0 / (1.0 + (a + b)))
    ptr = [0]
    idx, val = [], []
    for r in rows:
        for c in sorted(r):
            idx.append(c)
            val.append(r[c])
        ptr.append(len(idx))
    return n, np.array(ptr, np.int64), np.array(idx, np.int64), np.array(val)


def random_spd(n: int, density: float, seed: int):
    """A random sparse SPD matrix (diagonally dominant) by full rows."""
    rng = np.random.default_rng(seed)
    a = np.zeros((n, n))
    mask = rng.random((n, n)) < density
    mask = np.triu(mask, 1)
    vals = rng.uniform(-1, 1, size=(n, n)) * mask
    a = vals + vals.T
    a[np.arange(n), np.arange(n)] = np.abs(a).sum(axis=1) + 1.0 + rng.random(n)
    return a


def csr_of(a):
    n, ptr, idx, val = dense_to_csr(a)
    return n, ptr, idx, val


def hex_grid(dims, wrap=False, at=None):
    """hex_grid (ordering.rs:316-348): the 27-point node graph of a structured hex8 grid."""
    ni = dims[0] if wrap else dims[0] + 1
    nj, nk = dims[1] + 1, dims[2] + 1
    idn = lambda i, j, k: (i % ni) + ni * (j + nj * k)
    n = ni * nj * nk
    xyz = np.zeros((n, 3))
    adj = [[] for _ in range(n)]
    for k in range(nk):
        for j in range(nj):
            for i in range(ni):
                xyz[idn(i, j, k)] = at(i, j, k) if at else (i, j, k)
                for t in range(27):
                    di, dj, dk = t % 3, (t // 3) % 3, t // 9
                    a, b, c = i + ni + di - 1, j + dj, k + dk
                    if wrap:
                        a = a % ni
                    elif a < ni or a >= 2 * ni:
                        continue
                    else:
                        a = a - ni
                    if 1 <= b <= nj and 1 <= c <= nk:
                        adj[idn(i, j, k)].append(idn(a, b - 1, c - 1))
    return xyz, to_csr_graph(adj)


def to_csr_graph(adj):
    ptr = [0]
    idx = []
    for row in adj:
        r = sorted(set(row))
        idx.extend(r)
        ptr.append(len(idx))
    return np.array(ptr, np.int64), np.array(idx, np.int64)


def cylinder_wall():
    """cylinder_wall (ordering.rs:357-364): 160 around, 2 thick, 40 high."""
    def at(i, j, k):
        t = i / 160.0 * 2 * np.pi
        r = 50.0 + j * 0.5
        return (r * np.cos(t), r * np.sin(t), k * 2.0)

    return hex_grid((160, 2, 40), wrap=True, at=at)


def grid_12x12():
    """The 12 x 12 grid of a_permutation_with_separators_last (ordering.rs:288-302)."""
    m = 12
    xyz = np.array([[v // m, v % m, 0.0] for v in range(m * m)], dtype=np.float64)
    adj = [[] for _ in range(m * m)]
    for i in range(m):
        for j in range(m):
            for di, dj in [(0, 1), (1, 0), (0, -1), (-1, 0), (0, 0)]:
                a, b = i + di, j + dj
                if 0 <= a < m and 0 <= b < m:
                    adj[i * m + j].append(a * m + b)
    # the reference keeps each row in push order (no sort); the graph algorithms only test
    # membership and the fill is sorted, so sorted rows give the same results
    return xyz, to_csr_graph(adj)


# --- the reference's LDL^T test matrices (crates/solver/src/ldlt.rs:575-632, 981-1021) -------

def rows_to_csr(rows):
    """csr (ldlt.rs:576-590): each row sorted by column (stable), full rows."""
    ptr, idx, val = [0], [], []
    for row in rows:
        for c, v in sorted(row, key=lambda e: e[0]):
            idx.append(c)
            val.append(v)
        ptr.append(len(idx))
    return len(rows), np.array(ptr, np.int64), np.array(idx, np.int64), np.array(val, np.float64)


def shifted_grid(n3, sigma):
    """shifted_grid (ldlt.rs:596-632): a 3D grid Laplacian on 3 coupled unknowns per point,
    shifted by sigma, and the count of its eigenvalues below sigma (Dirichlet closed form)."""
    nx, ny, nz = n3
    point = lambda i, j, k: i + nx * (j + ny * k)
    size = 3 * nx * ny * nz
    rows = [[] for _ in range(size)]
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                p = 3 * point(i, j, k)
                for a in range(3):
                    rows[p + a].append((p + a, 6.0 - sigma))
                nxt = [point(i + 1, j, k) if i + 1 < nx else None, point(i, j + 1, k) if j + 1 < ny else None,
                       point(i, j, k + 1) if k + 1 < nz else None]
                for q in nxt:
                    if q is None:
                        continue
                    for a in range(3):
                        rows[p + a].append((3 * q + a, -1.0))
                        rows[3 * q + a].append((p + a, -1.0))
    axis = lambda n: [2.0 - 2.0 * np.cos(np.pi * m / (n + 1.0)) for m in range(1, n + 1)]
    below = sum(3 for a in axis(nx) for b in axis(ny) for c in axis(nz) if a + b + c < sigma)
    return rows_to_csr(rows), below


def saddle(nk, nb):
    """saddle (ldlt.rs:652-688): [[K, B], [B^T, 0]] with the zero-diagonal columns first."""
    n = nk + nb
    rows = [[] for _ in range(n)]
    for i in range(nk):
        for j in range(nk):
            rows[i].append((j, 0.5 if i == j else 0.05 / (1.0 + (i + j))))
    for c in range(nb):
        for i in range(nk):
            v = float((i * 7 + c * 3) % 5) - 2.0 + (3.0 if i == c else 0.0)
            rows[i].append((nk + c, v))
            rows[nk + c].append((i, v))
        for d in range(nb):
            rows[nk + c].append((nk + d, 0.0))
    order = list(range(nk, n)) + list(range(nk))
    inverse = [order.index(i) for i in range(n)]
    permuted = [None] * n
    for i, row in enumerate(rows):
        permuted[inverse[i]] = [(inverse[c], v) for c, v in row]
    return rows_to_csr(permuted)


def hex_pattern(n, sigma, nested_dissection):
    """hex_pattern (ldlt.rs:981-1021): a 27-point hex8-like block pattern on n^3 nodes, numbered
    by the reference's coordinate nested dissection (passed in: the oracle's restatement)."""
    ident = lambda i, j, k: i + n * (j + n * k)
    nodes = n ** 3
    xyz = np.array([[v % n, (v // n) % n, v // (n * n)] for v in range(nodes)], np.float64)
    adjacency = [[] for _ in range(nodes)]
    for k in range(n):
        for j in range(n):
            for i in range(n):
                for t in range(27):
                    a, b, c = t % 3, (t // 3) % 3, t // 9
                    x, y, z = i + a, j + b, k + c
                    if 1 <= x <= n and 1 <= y <= n and 1 <= z <= n:
                        adjacency[ident(i, j, k)].append(ident(x - 1, y - 1, z - 1))
    adj_ptr = np.zeros(nodes + 1, np.int64)
    adj_ptr[1:] = np.cumsum([len(a) for a in adjacency])
    adj_idx = np.array([w for a in adjacency for w in a], np.int64)
    order = nested_dissection(xyz, adj_ptr, adj_idx)
    place = np.zeros(nodes, np.int64)
    place[order] = np.arange(nodes)
    rows = [[] for _ in range(3 * nodes)]
    for v, adj in enumerate(adjacency):
        for w in adj:
            for a in range(3):
                for b in range(3):
                    if v == w:
                        value = 30.0 - sigma if a == b else 0.5
                    elif a == b:
                        value = -1.0
                    else:
                        value = 0.01 * ((v + w) % 3)
                    rows[3 * place[v] + a].append((3 * place[w] + b, value))
    return rows_to_csr(rows)


def fingerprint(values):
    """FNV-1a over the bit patterns of f64 values (ldlt.rs:756-758)."""
    h = 0xCBF29CE484222325
    for v in np.asarray(values, np.float64).view(np.uint64):
        h = ((h ^ int(v)) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def frozen_rhs(n, mul):
    """The exact right-hand side of the frozen cases (ldlt.rs:785-788): x_i = ((37 i mod 101) - 50) / 50, b = A x."""
    x = np.array([((i * 37 % 101) - 50.0) / 50.0 for i in range(n)])
    return mul(x)
