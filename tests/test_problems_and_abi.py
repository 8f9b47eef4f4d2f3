"""The synthetic inputs (element matrices pinned by rigid-body modes and total mass, the
configurations' sizes) and the C ABI (the library loads and exports every symbol its header
declares). CPU only: no call here computes on a GPU."""
import os
import re

import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rigid_modes(xyz, eq_dof):
    """The six rigid-body motions of the nodes, restricted to the equations."""
    modes = []
    for d in range(3):
        u = np.zeros((len(xyz), 3))
        u[:, d] = 1.0
        modes.append(u)
    for a, b in [(0, 1), (1, 2), (2, 0)]:
        u = np.zeros((len(xyz), 3))
        u[:, a] = -xyz[:, b]
        u[:, b] = xyz[:, a]
        modes.append(u)
    return [m.reshape(-1)[eq_dof] for m in modes]


@pytest.mark.parametrize("kind,jitter", [("hex8", 0.0), ("tet10", 0.0), ("tet10", 0.1)])
def test_element_matrices_have_the_rigid_modes_in_their_null_space(kind, jitter):
    p = P.block(kind, (3, 2, 2), (1.5, 1.0, 0.8), 210e9, 0.3, clamp_face=-1, jitter=jitter, seed=5, rho=7850.0)
    for u in rigid_modes(p.xyz, p.eq_dof):
        assert np.abs(p.mul(u)).max() <= 1e-6 * p.norm_inf() * np.abs(u).max()
    k = np.zeros((p.n, p.n))
    k[np.repeat(np.arange(p.n), np.diff(p.ptr)), p.idx] = p.kval
    assert np.allclose(k, k.T, rtol=0, atol=1e-9 * np.abs(k).max())
    if kind == "tet10":
        # the consistent mass sums to rho V in each direction, V the tetrahedra's volume
        x = p.xyz[p.conn[:, :4]]
        vol = np.abs(np.linalg.det(np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0], x[:, 3] - x[:, 0]], axis=1))).sum() / 6
        if jitter == 0.0:
            assert np.isclose(vol, 1.5 * 1.0 * 0.8, rtol=1e-12)
        assert np.isclose(p.mval.sum(), 3 * 7850.0 * vol, rtol=1e-12)


def test_the_clamped_cantilever_sizes():
    """config[0]: 41 x 9 x 9 nodes, x = 0 clamped: 9 720 equations, 663 750 entries of K_ff."""
    p = P.cantilever_hex8()
    assert p.n == 9720 and p.nnz == 663750
    assert np.isclose(p.rhs.sum(), -1000.0)
    assert np.all(np.diff(p.ptr) > 0)


def test_header_declarations_are_exported():
    header = open(os.path.join(ROOT, "include", "fc_chol.h")).read()
    declared = sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(fc_\w+)\(", header, re.M)))
    assert sorted(CH.EXPORTS) == declared
    L = CH.lib()
    missing = [name for name in declared if not hasattr(L, name)]
    assert not missing
    assert CH.version().startswith("fieldcraft_amd")
