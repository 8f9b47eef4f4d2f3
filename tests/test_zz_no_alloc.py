"""The allocation-free solve (fc_solve_device_noalloc, through the C ABI) against the solve it
varies (fc_solve_device): the same bits for every right-hand side, on two calls in a row (the
second on a smaller system with fewer right-hand sides, after the first left its values in the
shared work array), in FP64 and FP32 storage; passes split to fit a lowered limit; a factor whose
one right-hand side does not fit refused with TooLarge."""
import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P

pytestmark = pytest.mark.gpu


def block(p, k, seed):
    rng = np.random.default_rng(seed)
    return np.stack([p.rhs * (1.0 + 0.1 * c) + rng.standard_normal(p.n) for c in range(k)], axis=1)


def both(f, B):
    """B solved by fc_solve_device and by fc_solve_device_noalloc (device blocks, column-major)."""
    import torch

    a = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()
    b = a.clone()
    f.solve_device(a.data_ptr(), B.shape[1])
    f.solve_device_noalloc(b.data_ptr(), B.shape[1])
    torch.cuda.synchronize()
    return a.cpu().numpy().T, b.cpu().numpy().T


def bits(x):
    return np.ascontiguousarray(x).view(np.uint64)


def factor(k, precision):
    p = P.cube_hex8(k)
    up = p.upper()
    sym = CH.Symbolic(CH.UpperColumns(p.n, *up))
    return p, sym, CH.Factor(sym, up[2], precision=precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_solve_is_the_solve(gpu, precision):
    p, sym, f = factor(12, precision)
    print(f"cube 12^3 {precision}: n = {p.n}, 20 right-hand sides", flush=True)
    want, got = both(f, block(p, 20, 1))
    assert np.array_equal(bits(got), bits(want))
    q, qsym, g = factor(8, precision)  # the second call: a smaller system, fewer right-hand sides
    print(f"cube 8^3 {precision}: n = {q.n}, 3 right-hand sides", flush=True)
    want, got = both(g, block(q, 3, 2))
    assert np.array_equal(bits(got), bits(want))
    assert np.isfinite(got).all()


def test_noalloc_splits_passes_and_refuses_what_does_not_fit(gpu, monkeypatch):
    p, sym, f = factor(10, "fp64")
    need = sym.info()["solve_bytes"]
    monkeypatch.setenv("FC_NOALLOC_LIMIT", str(3 * need + 1024))  # three right-hand sides a pass
    print(f"cube 10^3: n = {p.n}, {need} bytes a right-hand side, limit {3 * need + 1024}", flush=True)
    want, got = both(f, block(p, 20, 3))
    assert np.array_equal(bits(got), bits(want))
    monkeypatch.setenv("FC_NOALLOC_LIMIT", str(need - 1))
    import torch

    a = torch.zeros(p.n, dtype=torch.float64, device="cuda")
    with pytest.raises(CH.TooLarge) as e:
        f.solve_device_noalloc(a.data_ptr(), 1)
    assert e.value.entries >= need and e.value.limit == need - 1
