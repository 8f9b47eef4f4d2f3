"""The assembly oracle (oracle/assembly_oracle.c: the reference's element matrices, assembly in
element order, reduction and upper columns) pinned by what fixes it independently of a
restatement:

* the host generator's element kernels (fieldcraft_amd/csrc/problems.cpp), written from the
  mathematics (B^T D B blocks as lambda g g^T + mu g g^T + mu (g.g) I; the tet10 mass from the
  exact barycentric monomial integral, no quadrature), agree entry by entry to rounding on
  distorted meshes, per-element materials and the consistent mass;
* closed forms of the unit cube's trilinear stiffness (2x2x2 Gauss integrates it exactly);
* rigid-body modes in the null space of every element stiffness, the exact symmetry of the
  stiffness, the mass of a tet10 summing to rho V;
* the first degenerate element in element order is the one reported.
The C ABI's assembly symbols are checked with the other exports (test_problems_and_abi.py)."""
import numpy as np
import pytest

from fieldcraft_amd import problems as P
from oracle import oracle as O


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(b)))


def oracle_upper(p, what="K", shift=0.0, xyz=None):
    uptr, uidx, _ = p.upper()
    return O.assemble_upper(p.kind, p.xyz if xyz is None else xyz, p.conn, p.young, p.poisson, p.eq_dof, uptr, uidx,
                            density=p.rho, what=what, shift=shift)


def test_hex8_distorted_contrast_matches_independent_kernels():
    rng = np.random.default_rng(5)
    young = np.where(rng.random(5 * 4 * 3) < 0.3, 21e6, 210e9)
    p = P.block("hex8", (5, 4, 3), (0.5, 0.4, 0.3), young, 0.49, clamp_face=0, jitter=0.2, seed=11)
    _, _, kval = p.upper()
    assert rel(oracle_upper(p), kval) < 5e-14


def test_tet10_stiffness_and_mass_match_independent_kernels():
    p = P.block("tet10", (3, 3, 2), (0.3, 0.3, 0.2), 210e9, 0.3, clamp_face=4, jitter=0.15, seed=7, rho=7850.0)
    _, _, kval = p.upper()
    _, _, mval = p.upper(p.mval)
    assert rel(oracle_upper(p, "K"), kval) < 5e-14
    assert rel(oracle_upper(p, "M"), mval) < 5e-14
    sigma = 98696.04
    assert rel(oracle_upper(p, "shifted", sigma), kval - sigma * mval) < 5e-14


UNIT_CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


@pytest.mark.parametrize("poisson", [0.0, 0.25])
def test_unit_cube_closed_forms(poisson):
    # N0 = (1-x)(1-y)(1-z), N1 = x(1-y)(1-z): K_xx(0,0) = (l + 2m)/9 + 2m/9, K_xx(0,1) =
    # -(l + 2m)/9 + m/9, K_xy(0,0) = (l + m)/12 (products of 1D integrals 1/3, 1/6, 1/2)
    E = 1.0
    lam = E * poisson / ((1 + poisson) * (1 - 2 * poisson))
    mu = E / (2 * (1 + poisson))
    ke = O.solid_stiffness("hex8", UNIT_CUBE, E, poisson)
    assert ke[0, 0] == pytest.approx((lam + 4 * mu) / 9, rel=1e-14)
    assert ke[0, 3] == pytest.approx(-(lam + 2 * mu) / 9 + mu / 9, rel=1e-14)
    assert ke[0, 1] == pytest.approx((lam + mu) / 12, rel=1e-14)


def rigid_modes(x):
    c = x.mean(axis=0)
    r = x - c
    modes = []
    for d in range(3):
        u = np.zeros_like(x)
        u[:, d] = 1.0
        modes.append(u.reshape(-1))
    for a, b in [(0, 1), (1, 2), (2, 0)]:
        u = np.zeros_like(x)
        u[:, a], u[:, b] = -r[:, b], r[:, a]
        modes.append(u.reshape(-1))
    return np.array(modes).T


def distorted(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "hex8":
        return UNIT_CUBE * [0.3, 0.2, 0.25] + rng.uniform(-0.03, 0.03, (8, 3))
    c = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.2 + rng.uniform(-0.02, 0.02, (4, 3))
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    return np.vstack([c] + [(c[a] + c[b]) / 2 for a, b in edges])


@pytest.mark.parametrize("kind", ["hex8", "tet10"])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_rigid_body_modes_and_symmetry(kind, seed):
    x = distorted(kind, seed)
    ke = O.solid_stiffness(kind, x, 210e9, 0.3)
    R = rigid_modes(x)
    assert np.abs(ke @ R).max() <= 1e-12 * np.abs(ke).max() * np.abs(R).max() * ke.shape[0]
    assert np.array_equal(ke, ke.T)  # the reference mirrors every block: exactly symmetric
    assert np.all(np.linalg.eigvalsh(ke)[6:] > 0)


def test_tet10_mass_sums_to_rho_v():
    x = distorted("tet10", 4)
    rho = 7850.0
    me = O.tet10_mass(x, rho)
    vol = abs(np.linalg.det((x[1:4] - x[0]))) / 6
    assert me.sum() == pytest.approx(rho * vol, rel=1e-14)
    assert np.allclose(me, me.T, rtol=0, atol=1e-15 * np.abs(me).max())
    assert np.all(np.linalg.eigvalsh((me + me.T) / 2) > 0)


def flattened(p, element):
    """Coordinates with one element folded: its first node moved through the opposite face."""
    xyz = p.xyz.copy()
    nodes = p.conn[element]
    far = xyz[nodes[6]] if p.kind == "hex8" else xyz[nodes[1]]
    xyz[nodes[0]] = far + (far - xyz[nodes[0]])
    return xyz


def integration_points(kind, mass):
    """Natural coordinates of the points the reference checks the Jacobian at: 2x2x2 Gauss for
    hex8; the 4-point rule, and with a mass Stroud's 4x4x4 conical points, for tet10."""
    if kind == "hex8":
        g = np.array([-1, 1]) / np.sqrt(3)
        return [(x, y, z) for z in g for y in g for x in g]
    a, b = (5 + 3 * np.sqrt(5)) / 20, (5 - np.sqrt(5)) / 20
    pts = [(b, b, b), (a, b, b), (b, a, b), (b, b, a)]
    if mass:
        x, _ = np.polynomial.legendre.leggauss(4)
        u = (x + 1) / 2
        pts += [(p, q * (1 - p), r * (1 - p) * (1 - q)) for p in u for q in u for r in u]
    return pts


def natural_derivatives(kind, xi):
    """dN/dxi, written from the shape functions' definitions (trilinear; quadratic tet10 with
    corners L(2L - 1) and mid-side nodes 4 La Lb on the model's edges)."""
    x, y, z = xi
    if kind == "hex8":
        c = UNIT_CUBE * 2 - 1
        f = 1 + np.array(xi) * c
        return np.stack([c[:, 0] * f[:, 1] * f[:, 2], c[:, 1] * f[:, 0] * f[:, 2], c[:, 2] * f[:, 0] * f[:, 1]], 1) / 8
    lam = np.array([1 - x - y - z, x, y, z])
    dl = np.array([[-1, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    dn = [(4 * lam[i] - 1) * dl[i] for i in range(4)]
    for a, b in [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]:
        dn.append(4 * (dl[a] * lam[b] + lam[a] * dl[b]))
    return np.array(dn)


def first_bad(p, xyz, mass):
    """The first element (in element order) whose normalized Jacobian det J / prod |rows of J|
    is not above 1e-12 at one of its integration points (the mass's too when it is assembled):
    the brute-force form of the reference's gate (model mesh.rs:254-269), with numpy's
    determinant."""
    pts = integration_points(p.kind, mass)
    for e, nodes in enumerate(p.conn):
        x = xyz[nodes]
        for xi in pts:
            j = natural_derivatives(p.kind, xi).T @ x
            if not np.linalg.det(j) / np.prod(np.linalg.norm(j, axis=1)) > 1e-12:
                return e
    return None


@pytest.mark.parametrize("kind,what", [("hex8", "K"), ("tet10", "K"), ("tet10", "shifted")])
def test_first_degenerate_element_is_named(kind, what):
    # with the mass, element 6 of this mesh fails only at a mass point: K alone names 12
    p = P.block(kind, (3, 3, 3), (0.3, 0.3, 0.3), 210e9, 0.3, clamp_face=4, rho=7850.0 if kind == "tet10" else 0.0)
    xyz = flattened(p, 13)
    bad = first_bad(p, xyz, what != "K")
    assert bad is not None
    with pytest.raises(O.DegenerateElement) as info:
        oracle_upper(p, what, 1e5, xyz=xyz)
    assert info.value.element == bad


def test_pattern_must_cover_the_couplings():
    p = P.block("hex8", (2, 2, 2), (0.2, 0.2, 0.2), 210e9, 0.3, clamp_face=0)
    uptr, uidx, _ = p.upper()
    # drop the first off-diagonal entry of the last column
    k = len(uptr) - 2
    keep = np.ones(len(uidx), bool)
    keep[uptr[k]] = False
    uptr2 = uptr.copy()
    uptr2[k + 1:] -= 1
    with pytest.raises(ValueError):
        O.assemble_upper("hex8", p.xyz, p.conn, p.young, p.poisson, p.eq_dof, uptr2, uidx[keep])
