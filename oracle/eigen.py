"""ORACLE — test infrastructure only (tests/ and bench.py's CPU-baseline leg).

numpy restatement of FieldCraft's shift-invert block Lanczos (crates/solver/src/eigen.rs), the
modal path that calls the factor's solves:

* ``Random`` — xorshift64* start vectors (eigen.rs:159-175);
* ``symmetric_eigen`` — Householder tridiagonalization, then implicit symmetric QR with
  Wilkinson shifts, eigenvalues ascending with their vectors (eigen.rs:26-153);
* ``orthonormalize`` — classical Gram-Schmidt in the M inner product, twice, against the basis
  and the block's earlier vectors, a vanishing column replaced by a fresh random direction
  (eigen.rs:231-279);
* ``lanczos`` — the block iteration with full reorthogonalization, convergence checks on a
  schedule, the wanted count and range, negative Ritz values, the caps' errors, Ritz vectors
  M-normalized with the largest component positive (eigen.rs:281-466).

The inner products here are numpy's (pairwise sums), not the reference's left-to-right sums:
the results agree to rounding, which is the bar a floating-point Lanczos is held to.
"""
from __future__ import annotations

import numpy as np

MASK = (1 << 64) - 1


class NotConverged(Exception):
    pass


class Random:
    """xorshift64* (eigen.rs:159-175): uniform in [-1, 1)."""

    def __init__(self, seed: int):
        self.s = max(int(seed), 1) & MASK

    def uniform(self) -> float:
        s = self.s
        s ^= s >> 12
        s ^= (s << 25) & MASK
        s ^= s >> 27
        self.s = s
        v = (s * 0x2545F4914F6CDD1D) & MASK
        return (v >> 11) / float(1 << 52) - 1.0

    def vector(self, n: int) -> np.ndarray:
        return np.array([self.uniform() for _ in range(n)])


def symmetric_eigen(a: np.ndarray):
    """Eigenvalues ascending and eigenvectors (columns) of a dense symmetric matrix
    (eigen.rs:26-100)."""
    t = np.array(a, dtype=np.float64, copy=True)
    n = t.shape[0]
    q = np.eye(n)
    for k in range(n - 2):
        m = k + 1
        x = t[m:, k]
        norm = np.sqrt(np.sum(x * x))
        if norm == 0.0:
            continue
        alpha = -norm if x[0] >= 0.0 else norm
        v = x.copy()
        v[0] -= alpha
        vnorm = np.sqrt(np.sum(v * v))
        if vnorm == 0.0:
            continue
        v /= vnorm
        p = t[m:, m:] @ v
        p -= (v @ p) * v
        t[m:, m:] -= 2.0 * (np.outer(v, p) + np.outer(p, v))
        t[m, k] = t[k, m] = alpha
        t[m + 1:, k] = 0.0
        t[k, m + 1:] = 0.0
        s = q[:, m:] @ v
        q[:, m:] -= 2.0 * np.outer(s, v)
    d = np.diag(t).copy()
    e = np.array([t[i + 1, i] for i in range(n - 1)])
    _tridiagonal_qr(d, e, q)
    order = sorted(range(n), key=lambda i: (d[i], i))
    return d[order], q[:, order]


def _tridiagonal_qr(d, e, q):
    """Implicit QR steps with Wilkinson shifts on a symmetric tridiagonal matrix, rotations
    applied to the columns of q (eigen.rs:102-153)."""
    n = len(d)
    eps = np.finfo(np.float64).eps

    def small(ev, a, b):
        return abs(ev) <= eps * (abs(a) + abs(b))

    h = n - 1
    iterations = 0
    cap = 30 * max(n, 1) + 30
    while h > 0:
        if small(e[h - 1], d[h - 1], d[h]):
            e[h - 1] = 0.0
            h -= 1
            continue
        l = h - 1
        while l > 0 and not small(e[l - 1], d[l - 1], d[l]):
            l -= 1
        if l > 0:
            e[l - 1] = 0.0
        iterations += 1
        if iterations > cap:
            raise NotConverged("the tridiagonal QR iteration did not converge")
        dd = (d[h - 1] - d[h]) / 2.0
        eh = e[h - 1]
        sign = 1.0 if dd >= 0.0 else -1.0
        mu = d[h] - eh * eh / (dd + sign * np.hypot(dd, eh))
        x = d[l] - mu
        z = e[l]
        for k in range(l, h):
            r = np.hypot(x, z)
            c, s = (1.0, 0.0) if r == 0.0 else (x / r, -z / r)
            if k > l:
                e[k - 1] = c * e[k - 1] - s * z
            a, b, dq = d[k], e[k], d[k + 1]
            d[k] = c * c * a - 2.0 * c * s * b + s * s * dq
            d[k + 1] = s * s * a + 2.0 * c * s * b + c * c * dq
            e[k] = c * s * (a - dq) + (c * c - s * s) * b
            if k + 1 < h:
                z = -s * e[k + 1]
                e[k + 1] *= c
                x = e[k]
            qp, qq = q[:, k].copy(), q[:, k + 1].copy()
            q[:, k] = c * qp - s * qq
            q[:, k + 1] = s * qp + c * qq


class Operator:
    """(K - sigma M)^{-1} M in the M inner product (eigen.rs:187-211): ``mass(x)`` = M x and
    ``solve_many(X)`` = (K - sigma M)^{-1} X for the columns of X."""

    def __init__(self, n, mass, solve_many):
        self.n, self.mass, self.solve_many = n, mass, solve_many


def orthonormalize(op: Operator, w, q, mq, rng: Random):
    """eigen.rs:231-279: returns (vectors, their M products, r) with r[j][i] the coefficient of
    column i on new vector j."""
    k = len(w)
    r = np.zeros((k, k))
    out, mout = [], []
    for i, v in enumerate(w):
        v = np.array(v, dtype=np.float64, copy=True)
        fresh = False
        for _attempt in range(4):
            before = np.sqrt(max(v @ op.mass(v), 0.0))
            for _ in range(2):
                for qj, mqj in zip(q, mq):
                    v -= (mqj @ v) * qj
                for j, (qj, mqj) in enumerate(zip(out, mout)):
                    c = mqj @ v
                    v -= c * qj
                    if not fresh:
                        r[j, i] += c
            mv = op.mass(v)
            norm = np.sqrt(max(v @ mv, 0.0))
            if norm > 1e-8 * before and norm > 0.0:
                if not fresh:
                    r[i, i] = norm
                out.append(v / norm)
                mout.append(mv / norm)
                break
            fresh = True
            v = rng.vector(op.n)
    return out, mout, r


def lanczos(op: Operator, shift: float, count: int, lo: float = -np.inf, hi: float = np.inf, block: int = 6,
            max_basis: int = 10 ** 9, tolerance: float = 1e-9):
    """eigen.rs:281-466. Returns dict(values, vectors (columns), negative, basis)."""
    n = op.n
    if n == 0:
        raise ValueError("there are no free degrees of freedom")
    b = min(max(block, 1), n)
    max_basis = max(min(max_basis, n), b)
    rng = Random(0x9E3779B97F4A7C15)
    q, mq = [], []
    start = [rng.vector(n) for _ in range(b)]
    cur, mcur, _ = orthonormalize(op, start, q, mq, rng)
    if len(cur) < b:
        raise NotConverged("could not build a Lanczos start block")
    prev_b = None
    next_check = max(count + b, 2 * b)
    t = np.zeros((0, 0))
    while True:
        j0 = len(q)
        w = list(op.solve_many(np.array(mcur).T).T)
        a = np.array([[mcur[i] @ w[j] for j in range(b)] for i in range(b)])
        for i in range(b):
            for j in range(i + 1, b):
                s = 0.5 * (a[i, j] + a[j, i])
                a[i, j] = a[j, i] = s
        q += cur
        mq += mcur
        m = len(q)
        t2 = np.zeros((m, m))
        t2[:j0, :j0] = t
        t = t2
        t[j0:j0 + b, j0:j0 + b] = a
        if prev_b is not None:
            for i in range(b):
                for j in range(b):
                    t[j0 + i, j0 - b + j] = prev_b[i, j]
                    t[j0 - b + j, j0 + i] = prev_b[i, j]
        full = m >= max_basis or m + b > n
        if full:
            nxt, mnxt, bnext = [], [], np.zeros((b, b))
        else:
            nxt, mnxt, bnext = orthonormalize(op, w, q, mq, rng)
        if m >= next_check or full or len(nxt) < b:
            next_check = m + max(m // 4, b)
            theta, s = symmetric_eigen(t)
            idx = sorted(range(m), key=lambda i: (-theta[i], i))
            done, accepted, found = False, [], 0
            for i in idx:
                th = theta[i]
                if not th > 0.0:
                    break
                res2 = 0.0 if full else float(np.sum((bnext @ s[m - b:, i]) ** 2))
                if np.sqrt(res2) > tolerance * th:
                    break
                lam = shift + 1.0 / th
                if lam > hi:
                    done = True
                    break
                if lam >= lo:
                    accepted.append(i)
                    found += 1
                    if found == count:
                        done = True
                        break
            if full and not done:
                if m >= n:
                    done = True
                else:
                    raise NotConverged(f"found {found} of {count} modes after {m} Lanczos vectors (the cap)")
            if done:
                negative = []
                for i in reversed(idx):
                    th = theta[i]
                    if not th < 0.0:
                        break
                    res2 = 0.0 if full else float(np.sum((bnext @ s[m - b:, i]) ** 2))
                    if np.sqrt(res2) > tolerance * abs(th):
                        break
                    negative.append(shift + 1.0 / th)
                Q = np.array(q).T
                values, vectors = [], []
                for i in accepted:
                    values.append(shift + 1.0 / theta[i])
                    x = Q @ s[:, i]
                    norm = np.sqrt(max(x @ op.mass(x), 0.0))
                    if norm > 0.0:
                        x = x / norm
                    big = x[np.argmax(np.abs(x))]
                    if big < 0.0:
                        x = -x
                    vectors.append(x)
                vec = np.array(vectors).T if vectors else np.zeros((n, 0))
                return dict(values=np.array(values), vectors=vec, negative=np.array(negative), basis=m)
        if full:
            raise NotConverged("the Lanczos basis is exhausted")
        cur, mcur = nxt, mnxt
        if len(cur) < b:
            raise NotConverged("the Lanczos block lost rank")
        prev_b = bnext
