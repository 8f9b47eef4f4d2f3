"""The small-front launches on the host (no GPU): a front of at most 64 columns, with at most 768
rows below them and an update matrix of at most 2e6 multiply-adds, is taken whole by one
workgroup of the level's small-front launch, and its work is counted once, under the "small"
kind, with the flops the tiled launches would count for it. A dense matrix is one front: 64
columns make it small, 65 do not."""
import numpy as np
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P

from matrices import dense_to_csr, grid_7x5, random_spd, upper

TILED = ("potrf", "trsm", "syrk", "update")


def analyse_dense(a):
    n, ptr, idx, val = dense_to_csr(a)
    return CH.Symbolic(CH.UpperColumns(n, *upper(n, ptr, idx, val)))


@pytest.mark.parametrize("n, small", [(64, True), (65, False), (200, False)])
def test_a_dense_front_is_small_up_to_64_columns(n, small):
    i = analyse_dense(random_spd(n, 1.0, 7)).info()
    assert i["fronts"] == 1
    kf = i["kind_flops"]
    if small:  # the diagonal block alone, sum_{t <= w} t^2: nothing below it
        assert kf["small"] == n * (n + 1) * (2 * n + 1) / 6 and all(kf[k] == 0 for k in TILED)
    else:
        assert kf["small"] == 0 and sum(kf[k] for k in TILED) == i["front_flops"]


def test_a_grid_of_small_fronts_takes_no_tiled_launch():
    i = analyse_dense(grid_7x5()).info()
    kf = i["kind_flops"]
    assert kf["small"] > 0 and all(kf[k] == 0 for k in TILED)
    assert abs(kf["small"] - i["front_flops"]) <= 1e-9 * i["front_flops"]


@pytest.mark.parametrize("maker", [lambda: P.cube_hex8(16), lambda: P.cantilever_hex8()])
def test_small_and_tiled_fronts_share_a_model(maker):
    p = maker()
    i = CH.Symbolic(CH.UpperColumns(p.n, *p.upper())).info()
    kf = i["kind_flops"]
    assert kf["small"] > 0 and kf["syrk"] > 0
    total = kf["small"] + sum(kf[k] for k in TILED)
    assert abs(total - i["front_flops"]) <= 1e-9 * i["front_flops"]
