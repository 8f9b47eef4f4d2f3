"""The oracle's element chain, factor and solve against the reference's blessed end-to-end results
(tests/oracle/{patch-hex8,patch-tet10,cantilever-tet10}/fixture.json in the reference, copied as
data into tests/golden/fixture_*.json): the constant-stress patch tests on distorted hex8 and tet10
meshes (corner displacements within 2.5e-13..2.25e-12, inner stresses within 1e-9..3e-9) and the
tet10 cantilever's tip deflection -4.0232748685235595 within 4.0e-9. The assembly oracle
(oracle/assembly_oracle.c) and the factor oracle (oracle/chol_oracle.c) are pinned here by the
reference's own numbers, beyond the mathematics they were pinned by before."""
import numpy as np

from fieldcraft_amd import problems as P
from oracle import oracle as O

from fixture_cases import check_patch, patch_system

import pytest


@pytest.mark.parametrize("name", ["patch_hex8", "patch_tet10"])
def test_patch_oracle_chain_matches_the_blessed_results(name):
    c, s = patch_system(name)
    vals = O.assemble_upper(c["kind"], s.xyz, s.conn, c["young"], c["poisson"], s.eq_dof, s.ptr, s.idx)
    u = O.Factor(s.n, s.ptr, s.idx, vals).solve(s.f)
    got = check_patch(c, s, u)
    # the whole field is the linear one: eps_i = (sigma_i - nu (sigma_j + sigma_k)) / E
    e = np.array([(1 - 0.25 * 5) / 1000, (2 - 0.25 * 4) / 1000, (3 - 0.25 * 3) / 1000])
    assert np.abs(s.full(u) - s.xyz * e).max() <= 1e-12
    assert abs(got["ux"] + 0.00025) < 1e-15


def cantilever():
    c = __import__("fixture_cases").load_case("cantilever_tet10")
    p = P.block("tet10", tuple(c["cells"]), tuple(c["size"]), c["young"], c["poisson"], clamp_face=0)
    tip = np.where(np.abs(p.xyz[:, 0] - c["size"][0]) < 1e-12)[0]
    w = __import__("fixture_cases").tet10_face_weights(p.xyz, p.conn, 0, c["size"][0])
    assert set(np.where(w > 0)[0]) <= set(tip)
    f_full = np.outer(w / w.sum(), np.asarray(c["force_total"]))
    rhs = f_full.reshape(-1)[p.eq_dof]
    probe = int(np.argmin(np.linalg.norm(p.xyz - np.asarray(c["probe_u"]), axis=1)))
    eq_z = int(np.where(p.eq_dof == 3 * probe + 2)[0][0])
    return c, p, rhs, tip, eq_z


def test_cantilever_tet10_oracle_matches_the_blessed_tip_deflection():
    c, p, rhs, tip, eq_z = cantilever()
    assert len(tip) == 25  # the quadratic lattice of the 2 x 2 tip face
    up = p.upper()
    vals = O.assemble_upper("tet10", p.xyz, p.conn, p.young, p.poisson, p.eq_dof, up[0], up[1])
    assert np.array_equal(vals, up[2]) or np.abs(vals - up[2]).max() <= 1e-12 * np.abs(up[2]).max()
    u = O.Factor(p.n, up[0], up[1], vals).solve(rhs)
    b = c["blessed"]["tip_uz"]
    assert abs(u[eq_z] - b["value"]) <= b["band"], (u[eq_z], b)
