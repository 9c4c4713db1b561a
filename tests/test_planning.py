"""Host planner: the reference's known answers and optimality against brute force (CPU only)."""
import itertools
import math

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


def _connected(tree, spec):
    """Every merge of the tree contracts at least one label (the sector search's rule)."""
    def labels(t):
        if isinstance(t, int):
            return set(l for l in spec.labels[t] if l > 0), True
        la, ca = labels(t[0])
        lb, cb = labels(t[1])
        return la ^ lb, ca and cb and bool(la & lb)
    return labels(tree)[1]


def test_device_time_model():
    """The device executor's roofline time model: a single GEMM against its
    closed form, and on a four-tensor U(1) chain the device plan is the
    admissible tree of least modelled time over every binary tree (brute
    force), while the reference objective (fused_permutes=False) is unchanged."""
    m, k, n = 300, 20, 7
    spec = P.ContractionSpec([[-1, 1], [1, -2]], [False] * 2, 2, (1, 1), 0)
    plan = P.plan_contraction(spec, [[m, k], [k, n]], fused_permutes=True)
    want = max(8.0 * m * k * n / P.DEVICE_FLOPS, 16.0 * (m * k + k * n + m * n) / P.DEVICE_BYTES)
    assert math.isclose(plan.device_time, want, rel_tol=1e-12)
    # the same GEMM bandwidth-bound and compute-bound
    assert P.step_device_time(1.0, 1e6) == 16e6 / P.DEVICE_BYTES
    assert P.step_device_time(1e12, 1.0) == 4e12 / P.DEVICE_FLOPS

    L = lambda *s: P.SectorLeg([((q,), d) for q, d in s])
    # one U(1) block GEMM by hand: charge q of o1 pairs a at -q and o2 at -q
    for scale in (1, 40):
        o1 = L((0, 4 * scale), (1, 3 * scale), (-1, 2 * scale))
        a = L((0, 5 * scale), (1, 1 * scale), (-1, 6 * scale))
        o2 = L((0, 3), (1, 2), (-1, 1))
        deg = lambda leg, q: dict((c[0], d) for c, d in leg.sectors)[q]
        sfl = sum(2.0 * deg(o1, q) * deg(a, -q) * deg(o2, -q) for q in (0, 1, -1))
        els = sum(deg(o1, q) * deg(a, -q) + deg(a, -q) * deg(o2, -q) + deg(o1, q) * deg(o2, -q)
                  for q in (0, 1, -1))
        spec = P.ContractionSpec([[-1, 1], [1, -2]], [False] * 2, 2, (1, 1), 0)
        plan = P.plan_contraction(spec, [[o1.dim, a.dim], [a.dim, o2.dim]],
                                  [[o1, a], [a.dual(), o2]], fused_permutes=True)
        assert plan.sector_flops == sfl
        want = max(4.0 * sfl / P.DEVICE_FLOPS, 16.0 * els / P.DEVICE_BYTES)
        assert math.isclose(plan.device_time, want, rel_tol=1e-12), scale

    a = L((0, 3), (1, 5), (-1, 4))
    b = L((0, 6), (1, 2), (-1, 2), (2, 1))
    c = L((0, 2), (1, 1), (-1, 1))
    o1 = L((0, 7), (1, 3), (-1, 3))
    o2 = L((0, 2), (1, 1), (-1, 1))
    legs = [[o1, a], [a.dual(), b], [b.dual(), c], [c.dual(), o2]]
    spec = P.ContractionSpec([[-1, 1], [1, 2], [2, 3], [3, -2]], [False] * 4, 2, (1, 1), 0)
    dims = [[l.dim for l in ls] for ls in legs]
    heur = P.compile_plan(P._heuristic_tree(spec), spec, dims, legs, "env_first", fused_permutes=True)
    heur_ref = P.compile_plan(P._heuristic_tree(spec), spec, dims, legs, "env_first")
    times, flops = [], []
    for t in _all_trees(list(range(4))):
        if not _connected(t, spec):
            continue
        p = P.compile_plan(t, spec, dims, legs, "brute", fused_permutes=True)
        if P._within_floors(p, heur):
            times.append(p.device_time)
        q = P.compile_plan(t, spec, dims, legs, "brute")
        if P._within_floors(q, heur_ref):
            flops.append(q.sector_flops)
    assert max(times) > 1.2 * min(times)        # the trees do differ
    chosen = P.plan_contraction(spec, dims, legs, memory_weight=0, fused_permutes=True)
    assert math.isclose(chosen.device_time, min(times), rel_tol=1e-12)
    ref = P.plan_contraction(spec, dims, legs, memory_weight=0)
    assert ref.sector_flops == min(flops)
