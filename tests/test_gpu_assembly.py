"""GPU parity of the assembly (fieldcraft_amd/csrc/assembly.hip through the C ABI) with the
oracle's restatement of the reference's assembly chain: every value of K_ff, M_ff and
K_ff - sigma M_ff the same bit for bit, on distorted meshes, per-element materials and the full
80^3 hex8 cube; the first degenerate element named as the reference's sequential loop names
it; and the device values feeding the factor directly."""
import numpy as np
import pytest

from fieldcraft_amd import assembly as AS
from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P
from oracle import oracle as O

from test_assembly import first_bad, flattened

pytestmark = pytest.mark.gpu


def device_values(p, what="K", shift=0.0, xyz=None):
    uptr, uidx, _ = p.upper()
    a = AS.Assembly(p.kind, p.xyz if xyz is None else xyz, p.conn, p.young, p.poisson, p.eq_dof, uptr, uidx,
                    density=p.rho)
    return a.assemble(what, shift).cpu().numpy()


def oracle_values(p, what="K", shift=0.0):
    uptr, uidx, _ = p.upper()
    return O.assemble_upper(p.kind, p.xyz, p.conn, p.young, p.poisson, p.eq_dof, uptr, uidx, density=p.rho,
                            what=what, shift=shift)


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def test_hex8_distorted_contrast_bitwise(gpu):
    rng = np.random.default_rng(5)
    young = np.where(rng.random(6 * 5 * 4) < 0.3, 21e6, 210e9)
    p = P.block("hex8", (6, 5, 4), (0.6, 0.5, 0.4), young, 0.49, clamp_face=0, jitter=0.2, seed=11)
    print("hex8 6x5x4 jitter 0.2 contrast", flush=True)
    assert same_bits(device_values(p), oracle_values(p))


@pytest.mark.parametrize("what", ["K", "M", "shifted"])
def test_tet10_distorted_bitwise(gpu, what):
    p = P.block("tet10", (4, 3, 3), (0.4, 0.3, 0.3), 210e9, 0.3, clamp_face=4, jitter=0.15, seed=7, rho=7850.0)
    sigma = (2 * np.pi * 50) ** 2
    print(f"tet10 4x3x3 jitter 0.15 {what}", flush=True)
    assert same_bits(device_values(p, what, sigma), oracle_values(p, what, sigma))


def test_repeated_calls_and_new_shift(gpu):
    p = P.block("tet10", (3, 3, 2), (0.3, 0.3, 0.2), 210e9, 0.3, clamp_face=4, jitter=0.1, seed=3, rho=7850.0)
    uptr, uidx, _ = p.upper()
    a = AS.Assembly.of_problem(p, uptr, uidx)
    for what, s in [("shifted", 1e5), ("K", 0.0), ("shifted", 4e6), ("M", 0.0)]:
        assert same_bits(a.assemble(what, s).cpu().numpy(), oracle_values(p, what, s))


@pytest.mark.parametrize("kind,what", [("hex8", "K"), ("tet10", "K"), ("tet10", "shifted")])
def test_first_degenerate_element_is_named(gpu, kind, what):
    p = P.block(kind, (3, 3, 3), (0.3, 0.3, 0.3), 210e9, 0.3, clamp_face=4, rho=7850.0 if kind == "tet10" else 0.0)
    xyz = flattened(p, 13)
    bad = first_bad(p, xyz, what != "K")
    print(f"degenerate {kind} {what}: expect element {bad}", flush=True)
    with pytest.raises(CH.DegenerateElement) as info:
        device_values(p, what, 1e5, xyz=xyz)
    assert info.value.column == bad


def test_bad_material_and_pattern_are_refused(gpu):
    p = P.block("hex8", (2, 2, 2), (0.2, 0.2, 0.2), 210e9, 0.3, clamp_face=0)
    uptr, uidx, _ = p.upper()
    with pytest.raises(CH.InvalidArgument):
        AS.Assembly("hex8", p.xyz, p.conn, 210e9, 0.5, p.eq_dof, uptr, uidx)
    k = len(uptr) - 2
    keep = np.ones(len(uidx), bool)
    keep[uptr[k]] = False
    uptr2 = uptr.copy()
    uptr2[k + 1:] -= 1
    with pytest.raises(CH.InvalidArgument):
        AS.Assembly("hex8", p.xyz, p.conn, 210e9, 0.3, p.eq_dof, uptr2, uidx[keep])
    a = AS.Assembly("hex8", p.xyz, p.conn, 210e9, 0.3, p.eq_dof, uptr, uidx)
    with pytest.raises(CH.InvalidArgument):  # no mass for hexahedra
        a.assemble("M")


def test_cube80_bitwise(gpu):
    """The 1.6M-equation hex8 cube, the whole K_ff (62M entries), bit for bit."""
    p = P.cube_hex8(80)
    print("hex8 80^3", p.n, flush=True)
    assert same_bits(device_values(p), oracle_values(p))


def test_device_values_feed_the_factor(gpu):
    """K - sigma M assembled on the device and factored from device memory: the same factor and
    solution, bit for bit, as from the oracle's values uploaded from the host."""
    p = P.block("tet10", (5, 4, 4), (0.5, 0.4, 0.4), 210e9, 0.3, clamp_face=4, jitter=0.1, seed=9, rho=7850.0)
    sigma = (2 * np.pi * 50) ** 2
    uptr, uidx, _ = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, uptr, uidx, np.zeros(len(uidx))))
    a = AS.Assembly.of_problem(p, uptr, uidx)
    vals = a.assemble("shifted", sigma)
    f_dev = CH.Factor(sym, device_values=vals.data_ptr())
    f_host = CH.Factor(sym, oracle_values(p, "shifted", sigma))
    b = np.sin(np.arange(p.n) * 0.37)
    assert same_bits(f_dev.values(), f_host.values())
    assert same_bits(f_dev.solve(b.copy()), f_host.solve(b.copy()))
