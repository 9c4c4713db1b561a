"""Environment-cache invalidation events on the host (CPU).

env(u -> v) contracts every tensor on u's side of the edge (u, v).
invalidate_node!(n) drops exactly the environments whose side holds n, and
invalidate_edge!(n, m) those whose side holds n or m, bumping the node
generations (envcache.jl:812-844).  The sides here come from a breadth-first
search over the oracle's tree, independent of the package's subtree helper.
"""
import itertools

import pytest

from oracle import heff_oracle as ref

from graft_amd import contractions as C, trees as Tr


def _side(t, u, v):
    seen, todo = {u}, [u]
    while todo:
        x = todo.pop()
        for y in ref.neighbors(t, x):
            if (x, y) != (u, v) and y not in seen:
                seen.add(y)
                todo.append(y)
    return seen


def _full_cache(topo, t):
    cache = C.EnvCache(topo)
    for u in range(t.n):
        for v in ref.neighbors(t, u):
            cache.envs[(u, v)] = object()
    return cache


@pytest.mark.parametrize("topo", [Tr.binary_topology(3), Tr.star_topology(3, 2),
                                  Tr.TreeTopology("0", [(str(i - 1), str(i)) for i in range(1, 6)])],
                         ids=["binary3", "star3x2", "chain6"])
def test_invalidation_drops_exactly_the_environments_that_see_the_node(topo):
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    all_keys = {(u, v) for u in range(t.n) for v in ref.neighbors(t, u)}
    assert len(all_keys) == 2 * (t.n - 1)
    for n in range(t.n):
        cache = _full_cache(topo, t)
        cache.invalidate_node_(n)
        assert set(cache.envs) == {k for k in all_keys if n not in _side(t, *k)}, n
        assert cache.node_generations[n] == 1 and sum(cache.node_generations) == 1
    edges = [(c, t.parent[c]) for c in range(t.n) if t.parent[c] >= 0]
    for n, m in edges:
        cache = _full_cache(topo, t)
        cache.invalidate_edge_(n, m)
        keep = {k for k in all_keys if n not in _side(t, *k) and m not in _side(t, *k)}
        assert set(cache.envs) == keep, (n, m)
        assert (n, m) not in keep and (m, n) not in keep     # both directions of the moved edge die
        assert cache.node_generations[n] == 1 and cache.node_generations[m] == 1
    # invalidations compose: two node events drop the union
    for a, b in itertools.combinations(range(min(t.n, 5)), 2):
        cache = _full_cache(topo, t)
        cache.invalidate_node_(a).invalidate_node_(b)
        assert set(cache.envs) == {k for k in all_keys if a not in _side(t, *k) and b not in _side(t, *k)}
