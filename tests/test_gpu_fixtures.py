"""The device chain (assembly of K_ff on the GPU, the numeric factorization and the solve through the
C ABI) against the reference's blessed end-to-end results (tests/golden/fixture_*.json, data from
the reference's tests/oracle cases): the constant-stress patch tests on distorted hex8 and tet10
meshes and the tet10 cantilever's tip deflection, in FP64 and in the mixed form refined to
berr <= 1e-15, each within the fixture's own band."""
import numpy as np
import pytest

from fieldcraft_amd import assembly as AS
from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P

from fixture_cases import check_patch, load_case, patch_system, tet10_face_weights

pytestmark = pytest.mark.gpu


def device_solve(kind, xyz, conn, young, poisson, eq_dof, uptr, uidx, rhs, precision):
    n = len(eq_dof)
    a = AS.Assembly(kind, xyz, conn, young, poisson, eq_dof, uptr, uidx)
    vals = a.assemble("K")
    sym = CH.Symbolic(CH.UpperColumns(n, uptr, uidx, np.zeros(len(uidx))))
    f = CH.Factor(sym, device_values=vals.data_ptr(), precision=precision)
    if precision == "fp64":
        return f.solve(rhs)
    # refinement needs K by full rows: from the device's upper columns
    up = vals.cpu().numpy()
    cols = np.repeat(np.arange(n), np.diff(uptr))
    rows = np.concatenate([uidx, cols[uidx != cols]])
    cc = np.concatenate([cols, uidx[uidx != cols]])
    vv = np.concatenate([up, up[uidx != cols]])
    order = np.lexsort((cc, rows))
    ptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=ptr[1:])
    # berr <= 1e-12 is the mixed form's bar, but the slender cantilever's forward error at that
    # backward error (~5e-7 relative) is wider than its fixture's band: refine to 1e-15, where
    # the refined solution is as accurate as the FP64 factor's
    x, rep = f.solve_refined(CH.Matrix(n, ptr, cc[order], vv[order]), rhs, target_berr=1e-15)
    assert rep["converged"], rep
    return x


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["patch_hex8", "patch_tet10"])
def test_patch_device_chain_matches_the_blessed_results(gpu, name, precision):
    c, s = patch_system(name)
    u = device_solve(c["kind"], s.xyz, s.conn, c["young"], c["poisson"], s.eq_dof, s.ptr, s.idx, s.f, precision)
    check_patch(c, s, u)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cantilever_tet10_device_chain_matches_the_blessed_tip_deflection(gpu, precision):
    c = load_case("cantilever_tet10")
    p = P.block("tet10", tuple(c["cells"]), tuple(c["size"]), c["young"], c["poisson"], clamp_face=0)
    w = tet10_face_weights(p.xyz, p.conn, 0, c["size"][0])
    rhs = np.outer(w / w.sum(), np.asarray(c["force_total"])).reshape(-1)[p.eq_dof]
    up = p.upper()
    u = device_solve("tet10", p.xyz, p.conn, p.young, p.poisson, p.eq_dof, up[0], up[1], rhs, precision)
    probe = int(np.argmin(np.linalg.norm(p.xyz - np.asarray(c["probe_u"]), axis=1)))
    uz = u[int(np.where(p.eq_dof == 3 * probe + 2)[0][0])]
    b = c["blessed"]["tip_uz"]
    assert abs(uz - b["value"]) <= b["band"], (uz, b)
