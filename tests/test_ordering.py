"""Fill-reducing ordering (ordering.rs): the reference's own property tests on the oracle and on
the library's generator, the two bit for bit equal, and fill() against brute force. CPU only."""
import numpy as np
import pytest

from fieldcraft_amd import problems as P
from oracle import oracle as O

from matrices import cylinder_wall, grid_12x12, hex_grid, to_csr_graph
from test_symbolic import brute_force

IMPLS = {"oracle": O, "library": P}


@pytest.mark.parametrize("impl", list(IMPLS))
def test_a_permutation_with_separators_last(impl):
    """ordering.rs:288-312: a permutation whose top separator (last 12) is one grid line."""
    m = 12
    xyz, (ap, ai) = grid_12x12()
    order = IMPLS[impl].nested_dissection(xyz, ap, ai)
    assert sorted(order.tolist()) == list(range(m * m))
    last = order[-m:]
    assert np.all(last // m == last[0] // m) or np.all(last % m == last[0] % m)
    assert np.array_equal(order, IMPLS[impl].nested_dissection(xyz, ap, ai))


def test_orders_are_the_oracles_bit_for_bit():
    for xyz, (ap, ai) in [grid_12x12(), hex_grid((6, 5, 7)), cylinder_wall()]:
        assert np.array_equal(O.nested_dissection(xyz, ap, ai), P.nested_dissection(xyz, ap, ai))
        assert np.array_equal(O.graph_dissection(ap, ai), P.graph_dissection(ap, ai))
        oo, go = O.fill_reducing_order(xyz, ap, ai)
        op, gp = P.fill_reducing_order(xyz, ap, ai)
        assert np.array_equal(oo, op) and go == gp


def test_graph_dissection_is_a_deterministic_permutation():
    """ordering.rs:366-379, including two disconnected blocks and an isolated node."""
    _, (ap, ai) = cylinder_wall()
    for impl in IMPLS.values():
        order = impl.graph_dissection(ap, ai)
        assert sorted(order.tolist()) == list(range(len(ap) - 1))
        assert np.array_equal(order, impl.graph_dissection(ap, ai))
    xyz, (bp, bi) = hex_grid((6, 6, 6))
    m = len(bp) - 1
    adj = [bi[bp[v]:bp[v + 1]].tolist() for v in range(m)]
    two = adj + [[w + m for w in row] for row in adj] + [[]]
    tp, ti = to_csr_graph(two)
    for impl in IMPLS.values():
        assert sorted(impl.graph_dissection(tp, ti).tolist()) == list(range(2 * m + 1))
    assert np.array_equal(O.graph_dissection(tp, ti), P.graph_dissection(tp, ti))


def test_the_graph_cuts_a_thin_curved_wall_with_less_fill():
    """ordering.rs:381-387."""
    xyz, (ap, ai) = cylinder_wall()
    for impl in IMPLS.values():
        planes = impl.fill(ap, ai, impl.nested_dissection(xyz, ap, ai))
        graph = impl.fill(ap, ai, impl.graph_dissection(ap, ai))
        assert graph < planes
        order, _ = impl.fill_reducing_order(xyz, ap, ai)
        assert impl.fill(ap, ai, order) == graph


def test_a_block_keeps_the_better_ordering():
    """ordering.rs:390-395: chosen fill <= planes fill on a 16^3 block."""
    xyz, (ap, ai) = hex_grid((16, 16, 16))
    for impl in IMPLS.values():
        planes = impl.fill(ap, ai, impl.nested_dissection(xyz, ap, ai))
        order, _ = impl.fill_reducing_order(xyz, ap, ai)
        assert impl.fill(ap, ai, order) <= planes


@pytest.mark.parametrize("seed", range(4))
def test_fill_is_the_brute_force_factor_size(seed):
    rng = np.random.default_rng(seed)
    n = 30 + 10 * seed
    adj = [[v] for v in range(n)]
    for _ in range(2 * n):
        a, b = rng.integers(0, n, 2)
        adj[a].append(b)
        adj[b].append(a)
    ap, ai = to_csr_graph(adj)
    order = rng.permutation(n)
    a = np.zeros((n, n))
    pos = np.argsort(order)
    for v in range(n):
        for w in ai[ap[v]:ap[v + 1]]:
            a[pos[v], pos[w]] = 1.0
    lower, _, counts = brute_force(a)
    assert O.fill(ap, ai, order) == counts.sum() == P.fill(ap, ai, order)


def test_the_cantilever_order_is_the_oracles():
    p = P.cantilever_hex8()
    order, graph = O.fill_reducing_order(p.xyz, p.adj_ptr, p.adj_idx)
    assert np.array_equal(order, p.order) and graph == p.graph_chosen
