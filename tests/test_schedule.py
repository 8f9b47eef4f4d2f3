"""The numeric phase's schedule on the host (no GPU): setting it rebuilds the launch program,
and every schedule covers the same algorithmic work. A dropped or repeated panel product, a
slab update that missed columns or a split piece counted twice would change a kind's flop
total, so the totals must not depend on the slab width or the split targets."""
import pytest

from fieldcraft_amd import cholesky as CH
from fieldcraft_amd import problems as P

DEFAULT = {"slab": 4, "split_target": 1024, "split_side_fp64": 384, "split_side_fp32": 256}


def analyse(p):
    return CH.Symbolic(CH.UpperColumns(p.n, *p.upper()))


def test_default_schedule_and_round_trip():
    s = analyse(P.cube_hex8(8))
    assert s.schedule() == DEFAULT
    s.set_schedule(slab=2, split_side_fp32=128)
    assert s.schedule() == dict(DEFAULT, slab=2, split_side_fp32=128)


@pytest.mark.parametrize("maker", [lambda: P.cube_hex8(16), lambda: P.cantilever_hex8()])
def test_every_schedule_covers_the_same_work(maker):
    s = analyse(maker())
    base = s.info()
    seen = set()
    for slab in (1, 2, 3, 4, 8, 16):
        for split in (1, 64, 1024, 65536):
            s.set_schedule(slab=slab, split_target=split, split_side_fp64=split, split_side_fp32=split)
            i = s.info()
            assert i["kind_flops"] == base["kind_flops"], (slab, split)
            assert i["kind_entries"] == base["kind_entries"], (slab, split)
            seen.add(i["launches"])
    assert len(seen) > 1  # the schedules do change the launches


def test_wider_slabs_need_fewer_launches():
    s = analyse(P.cube_hex8(16))
    counts = []
    for slab in (1, 2, 4):
        s.set_schedule(slab=slab)
        counts.append(s.info()["launches"])
    assert counts[0] > counts[1] > counts[2], counts


@pytest.mark.parametrize("bad", [dict(slab=0), dict(slab=17), dict(split_target=0), dict(split_side_fp32=70000)])
def test_out_of_range_schedules_are_refused(bad):
    s = analyse(P.cube_hex8(4))
    with pytest.raises(CH.InvalidArgument):
        s.set_schedule(**bad)
    assert s.schedule() == DEFAULT


def test_unknown_keys_are_refused():
    s = analyse(P.cube_hex8(4))
    with pytest.raises(ValueError):
        s.set_schedule(panels=2)
