"""Host planner: the reference's known answers and optimality against brute force (CPU only)."""
import itertools
import math

import numpy as np
import pytest

from graft_amd import planning as P
from graft_amd import trees as Tr


def test_spec_validation():
    with pytest.raises(ValueError):
        P.ContractionSpec([[-1, 1]], [], 1, (1, 0), 0)
    with pytest.raises(ValueError):
        P.ContractionSpec([[-1, 1], [1, -2]], [False] * 2, 2, (2, 1), 0)
    with pytest.raises(ValueError):
        P.ContractionSpec([[-1, 1], [1, -2]], [False] * 2, 2, (1, 1), 0, preferred_slots=[0])


def test_dense_cost_fixture():
    m = P.dense_cost([-1, 1], [2, 3], [1, -2], [3, 4])
    assert (m["oindA"], m["cindA"], m["oindB"], m["cindB"]) == ([0], [1], [1], [0])
    assert m["labels"] == [-1, -2] and m["dims"] == [2, 4] and m["peak_elements"] == 8
    with pytest.raises(ValueError):
        P.dense_cost([-1, 1], [2, 3], [1, -2], [4, 5])


def _kat_legs():
    L = lambda *s: P.SectorLeg([((q,), d) for q, d in s])
    left = [L((0, 1), (1, 1)), L((0, 1), (-1, 2))]
    mid = [L((0, 1), (1, 2)), L((0, 1), (-1, 4))]
    right = [L((0, 1), (1, 4)), L((0, 2), (-1, 1))]
    return [left, mid, right]


def test_reference_known_answers_60_30_63_28():
    """lib/GraftPlanning/test/runtests.jl:167-204."""
    spec = P.ContractionSpec([[-1, 1], [1, 2], [2, -2]], [False] * 3, 2, (1, 1), 0, [1, 2])
    dims = [[2, 3], [3, 5], [5, 3]]
    dense = P.plan_contraction(spec, dims, _kat_legs(), sector_aware=False, memory_weight=0)
    sector = P.plan_contraction(spec, dims, _kat_legs(), sector_aware=True, memory_weight=0)
    assert dense.strategy == "env_first" and dense.flops == 60 and dense.sector_flops == 30
    assert sector.strategy == "sector_exact" and sector.flops == 63 and sector.sector_flops == 28
    assert sorted([sector.steps[0].a, sector.steps[0].b]) == [1, 2]
    assert sector.sector_operand_bytes < sector.operand_bytes
    assert sector.sector_live_peak_bytes <= sector.live_peak_bytes


def test_env_first_fixture_and_cap():
    spec = P.ContractionSpec([[-1, 1], [1, 2], [2, -2]], [False] * 3, 2, (1, 1), 0, [1, 2])
    dims = [[2, 3], [3, 4], [4, 5]]
    plan = P.plan_contraction(spec, dims, optimize=False)
    assert plan.strategy == "env_first" and len(plan.steps) == 2 and plan.scalar_bytes == 16
    assert plan.peak_elements >= 10 and plan.live_peak_bytes >= plan.operand_bytes
    with pytest.raises(ValueError):
        P.plan_contraction(spec, dims, optimize=False, memory_cap_bytes=plan.live_peak_bytes - 1)


def _all_trees(items):
    if len(items) == 1:
        yield items[0]
        return
    first, rest = items[0], items[1:]
    for r in range(len(rest)):
        for left in itertools.combinations(rest, r):
            right = [x for x in rest if x not in left]
            for lt in _all_trees([first] + list(left)):
                for rt in _all_trees(right):
                    yield (lt, rt)


def _tree_flops(tree, spec, dims):
    if isinstance(tree, int):
        return 0.0, list(spec.labels[tree]), list(dims[tree])
    fa, la, da = _tree_flops(tree[0], spec, dims)
    fb, lb, db = _tree_flops(tree[1], spec, dims)
    m = P.dense_cost(la, da, lb, db)
    return fa + fb + m["flops"], m["labels"], m["dims"]


def test_optimal_tree_matches_brute_force():
    """Brute-force enumeration of every binary tree (contraction_planning.jl:43-104).

    Network: the one-site map of a degree-3 node, x(c1, c2, p, parent),
    W(u1, u2, p_out, p_in, up) and three environments."""
    # x(c1,c2,p,par=1) W(u1,u2,pout,pin,up) E1 E2 E3(parent)
    D, w, d = 8, 3, 2
    spec = P.ContractionSpec(
        [[1, 3, 5, 7], [2, 4, -3, 5, 6], [1, 2, -1], [3, 4, -2], [7, 6, -4]],
        [False] * 5, 4, (3, 1), 0, [2, 3, 1, 4])
    dims = [[D, D, d, D], [w, w, d, d, w], [D, w, D], [D, w, D], [D, w, D]]
    best = min(_tree_flops(t, spec, dims)[0] for t in _all_trees(list(range(5))))
    plan = P.plan_contraction(spec, dims, memory_weight=0)
    opt = P._optimal_tree(spec, dims)
    assert _tree_flops(opt, spec, dims)[0] == best
    assert plan.flops <= P.plan_contraction(spec, dims, optimize=False).flops
