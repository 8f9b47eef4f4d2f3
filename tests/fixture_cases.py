"""The reference's blessed end-to-end cases (tests/golden/fixture_*.json, written by
tests/golden/make_fixtures.py) as linear systems: mesh, free equations, the upper-column pattern of
K_ff and the load vector, built with numpy from the cases' data. The element matrices and the
solution come from the code under test (the oracle on the CPU, the library on the GPU); this
module only states supports and loads:

* symmetry on a coordinate plane fixes that direction's displacement of every node on it;
* a uniform traction t on a face is its consistent nodal load, sum_i (int_face N_i dA) t: for a
  bilinear quadrilateral (hex8 face) by 2 x 2 Gauss points on its own map; for a straight-sided
  six-node triangle (tet10 face) corners 0 and mid-side nodes A/3 each (the integrals of the
  quadratic triangle's shape functions);
* a force on a set of element faces (the cantilever's tip) is a total split by the faces'
  consistent nodal weights, area-weighted (Distribution::Total, crates/model/src/loads.rs:42-54):
  the same weights as a uniform traction's.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

HEX8_FACES = [[0, 1, 2, 3], [4, 5, 6, 7], [0, 1, 5, 4], [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]]
TET_FACES = [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]
TET10_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]  # mid-side nodes 4..9


def load_case(name):
    with open(os.path.join(HERE, "golden", f"fixture_{name}.json")) as fh:
        return json.load(fh)


def upper_pattern(conn, dof_eq, neq):
    """Upper columns over the free equations: column j holds every row i <= j coupled to j
    through an element (rows sorted)."""
    pairs = set()
    for el in conn:
        eqs = [dof_eq[3 * v + d] for v in el for d in range(3)]
        eqs = [e for e in eqs if e >= 0]
        for a in eqs:
            for b in eqs:
                if a <= b:
                    pairs.add((b, a))
    cols = sorted(pairs)
    ptr = np.zeros(neq + 1, np.int64)
    for c, _ in cols:
        ptr[c + 1] += 1
    ptr = np.cumsum(ptr)
    idx = np.array([r for _, r in cols], np.int64)
    return ptr, idx


def quad_load(x, t):
    """Consistent nodal loads of a uniform traction t on the bilinear quadrilateral x (4 x 3)."""
    g = 1.0 / np.sqrt(3.0)
    out = np.zeros((4, 3))
    for xi in (-g, g):
        for eta in (-g, g):
            n = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
            dxi = 0.25 * np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)])
            deta = 0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])
            da = np.linalg.norm(np.cross(dxi @ x, deta @ x))
            out += np.outer(n, t) * da
    return out


class System:
    """K_ff u = f of a case: mesh, free equations (eq -> 3 node + direction), pattern and load."""

    def __init__(self, kind, xyz, conn, young, poisson, fixed, f_full):
        self.kind, self.young, self.poisson = kind, young, poisson
        self.xyz = np.asarray(xyz, np.float64)
        self.conn = np.asarray(conn, np.int64)
        nd = 3 * len(self.xyz)
        self.dof_eq = -np.ones(nd, np.int64)
        free = [d for d in range(nd) if d not in fixed]
        self.dof_eq[free] = np.arange(len(free))
        self.eq_dof = np.array(free, np.int64)
        self.n = len(free)
        self.ptr, self.idx = upper_pattern(self.conn, self.dof_eq, self.n)
        self.f = np.asarray(f_full)[self.eq_dof]
        self.fixed = sorted(fixed)

    def node_at(self, p):
        d = np.linalg.norm(self.xyz - np.asarray(p), axis=1)
        i = int(np.argmin(d))
        assert d[i] < 1e-12, p
        return i

    def displacement(self, u_eq, node):
        out = np.zeros(3)
        for d in range(3):
            e = self.dof_eq[3 * node + d]
            out[d] = u_eq[e] if e >= 0 else 0.0
        return out

    def full(self, u_eq):
        u = np.zeros(3 * len(self.xyz))
        u[self.eq_dof] = u_eq
        return u.reshape(-1, 3)


def tet10_face_weights(xyz, conn, axis, side):
    """Consistent nodal weights (int_face N_i dA) of the straight-sided tet10 faces lying on the
    plane x_axis = side: mid-side nodes A/3 per face, corners 0."""
    w = np.zeros(len(xyz))
    for el in conn:
        for face in TET_FACES:
            corners = [el[q] for q in face]
            if all(abs(xyz[v, axis] - side) <= 1e-12 * max(1.0, abs(side)) for v in corners):
                area = 0.5 * np.linalg.norm(np.cross(xyz[corners[1]] - xyz[corners[0]], xyz[corners[2]] - xyz[corners[0]]))
                for e, (p, q) in enumerate(TET10_EDGES):
                    if p in face and q in face:
                        w[el[4 + e]] += area / 3.0
    return w


def patch_system(name):
    c = load_case(name)
    xyz = np.asarray(c["xyz"])
    conn = c["conn"]
    fixed = set()
    for axis in c["symmetry"]:
        a = "xyz".index(axis)
        for v in np.where(xyz[:, a] == 0.0)[0]:
            fixed.add(3 * int(v) + a)
    f = np.zeros((len(xyz), 3))
    for tr in c["tractions"]:
        a = "xyz".index(tr["set"][0])
        side = float(tr["set"][1])
        t = np.asarray(tr["vector"])
        for el in conn:
            if c["kind"] == "hex8":
                for face in HEX8_FACES:
                    nodes = [el[q] for q in face]
                    if all(xyz[v, a] == side for v in nodes):
                        f[nodes] += quad_load(xyz[nodes], t)
            else:
                for face in TET_FACES:
                    corners = [el[q] for q in face]
                    if all(xyz[v, a] == side for v in corners):
                        area = 0.5 * np.linalg.norm(np.cross(xyz[corners[1]] - xyz[corners[0]],
                                                             xyz[corners[2]] - xyz[corners[0]]))
                        for e, (p, q) in enumerate(TET10_EDGES):
                            if p in face and q in face:
                                f[el[4 + e]] += t * area / 3.0
    return c, System(c["kind"], xyz, conn, c["young"], c["poisson"], fixed, f.reshape(-1))


def linear_fit_stress(sys_, u_eq, node, young, poisson):
    """The stress of the linear field fitted to the displacements of the elements around `node`
    (exact for a patch test's linear solution): u(x) = a + G x by least squares, eps = sym(G),
    sigma = lambda tr(eps) I + 2 mu eps."""
    els = [el for el in sys_.conn if node in el]
    nodes = sorted({int(v) for el in els for v in el})
    u = sys_.full(u_eq)[nodes]
    X = np.hstack([np.ones((len(nodes), 1)), sys_.xyz[nodes]])
    coef, *_ = np.linalg.lstsq(X, u, rcond=None)
    G = coef[1:].T  # G[i, j] = du_i / dx_j
    eps = 0.5 * (G + G.T)
    lam = young * poisson / ((1 + poisson) * (1 - 2 * poisson))
    mu = young / (2 * (1 + poisson))
    return lam * np.trace(eps) * np.eye(3) + 2 * mu * eps


def check_patch(c, sys_, u_eq):
    """The blessed corner displacements and inner stresses within their bands."""
    b = c["blessed"]
    u = sys_.displacement(u_eq, sys_.node_at(c["probe_u"]))
    got = {"ux": u[0], "uy": u[1], "uz": u[2]}
    s = linear_fit_stress(sys_, u_eq, sys_.node_at(c["probe_stress"]), c["young"], c["poisson"])
    got.update(sxx=s[0, 0], syy=s[1, 1], szz=s[2, 2], sxy=s[0, 1])
    for k, v in got.items():
        assert abs(v - b[k]["value"]) <= b[k]["band"], (k, v, b[k])
    return got
