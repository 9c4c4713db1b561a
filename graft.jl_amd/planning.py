"""Shape-only contraction planning (host side) — the reference's Planning module.

Restates lib/GraftPlanning/src/Planning/{types,planner,cost}.jl: a labelled
``ContractionSpec`` is compiled into a list of binary ``PairStep``s.  The
candidates are the env-first fold (the memory floor), an exact FLOP-optimal
tree (exhaustive subset DP, the objective of TensorOperations.optimaltree),
a memory-greedy tree and, for abelian sector spaces, an exact sector DP over
a Pareto frontier.  A candidate is admissible only within the env-first
plan's peak and live-byte floors and an optional hard byte cap; the score is
``flops + memory_weight * live_peak_bytes`` (sector_flops / sector bytes when
the sector model applies).  Sector plans for the device executor are ranked
by a roofline time model instead (``device_time``: per step the larger of its
FLOPs over an effective MFMA rate and its stored operand + result bytes over
an effective HBM rate), because a block-sparse step with a short contraction
is HBM-bound and the FLOP-optimal order can materialize a far larger
intermediate.  Indices are 0-based; labels follow ncon (> 0 contracted, < 0
open, output ordered -1, -2, ...).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

EXACT_TENSOR_LIMIT = 10   # planner.jl:232 (_EXACT_TENSOR_LIMIT)

# Roofline time model of the device executor (MI355X, complex FP64): the
# rates sector steps reach in practice, not the peaks (8 TB/s, 78.6 TFLOP/s).
# Fitted to measurement: every admissible pairwise order of the root-edge
# Hubbard eff_h2 map at D = 256 and D = 1024 (10 orders, U(1) x U(1)) was
# compiled and timed on MI355X, and these two rates minimize the mean
# |log(modelled / measured time)| (0.18; 0.33 for 40 TFLOP/s and 3 TB/s);
# at both sizes the order the model ranks first is the measured fastest.
DEVICE_FLOPS = 46e12      # real FLOP/s of a block GEMM step (coupled-sector blocks counted dense)
DEVICE_BYTES = 2e12       # B/s of a streaming step (stored A + B + C, each once)


def step_device_time(sector_flops, elements, scalar_bytes=16):
    """Modelled seconds of one pair step: ``sector_flops`` is 2 m k n summed
    over blocks (real FLOPs = 4 x that for complex scalars); ``elements`` the
    stored elements of A, B and C."""
    return max(4.0 * sector_flops / DEVICE_FLOPS, scalar_bytes * elements / DEVICE_BYTES)


# ---------------------------------------------------------------------------
# Types — types.jl:12-115
# ---------------------------------------------------------------------------


class ContractionSpec:
    """ContractionSpec(labels, conjs, nopen, out_partition, dynamic_slot; preferred_slots)."""

    def __init__(self, labels, conjs, nopen, out_partition, dynamic_slot=None,
                 preferred_slots=None):
        self.labels = [list(map(int, l)) for l in labels]
        self.conjs = [bool(c) for c in conjs]
        self.nopen = int(nopen)
        self.out_partition = (int(out_partition[0]), int(out_partition[1]))
        self.dynamic_slot = None if dynamic_slot is None else int(dynamic_slot)
        if len(self.labels) != len(self.conjs):
            raise ValueError("ContractionSpec: labels/conjs length mismatch")
        if self.dynamic_slot is not None and self.dynamic_slot != 0:
            raise ValueError("ContractionSpec: only dynamic slot 0 is supported")
        if sum(self.out_partition) != self.nopen:
            raise ValueError(f"ContractionSpec: output partition {self.out_partition} does not "
                             f"contain {self.nopen} open legs")
        inputs = [i for i in range(len(self.labels)) if i != self.dynamic_slot]
        order = list(preferred_slots) if preferred_slots else list(inputs)
        if sorted(order) != inputs:
            raise ValueError("ContractionSpec: preferred slots must contain every non-dynamic slot once")
        self.preferred_slots = order

    def key(self):
        return (tuple(map(tuple, self.labels)), tuple(self.conjs), self.nopen,
                self.out_partition, self.dynamic_slot, tuple(self.preferred_slots))


@dataclass(frozen=True)
class PairStep:
    """One compiled binary contraction (types.jl:54-65), 0-based."""
    a: int
    b: int
    dst: int
    oindA: tuple
    cindA: tuple
    oindB: tuple
    cindB: tuple
    conjA: bool
    conjB: bool
    out: tuple   # (codomain positions, domain positions) over (A open..., B open...)


@dataclass
class ContractionPlan:
    """types.jl:94-115."""
    nslots: int
    output_slot: int
    steps: list
    strategy: str = "env_first"
    flops: float = math.nan
    peak_elements: float = math.nan
    sector_flops: float = math.nan
    sector_peak_elements: float = math.nan
    sector_peak_block_elements: float = math.nan
    scalar_bytes: int = 16
    operand_bytes: float = math.nan
    live_peak_bytes: float = math.nan
    known_temporary_peak_bytes: float = math.nan
    known_permutation_peak_bytes: float = math.nan
    sector_operand_bytes: float = math.nan
    sector_live_peak_bytes: float = math.nan
    sector_known_temporary_peak_bytes: float = math.nan
    sector_known_permutation_peak_bytes: float = math.nan
    scalar_output: bool = False
    candidate_failures: tuple = ()
    device_time: float = math.nan       # step_device_time summed over the steps

    @property
    def ninputs(self):
        return self.nslots - len(self.steps)


# ---------------------------------------------------------------------------
# Abelian sector legs (flat-leg picture).  A flat leg carries a list of
# (charge tuple, degeneracy); a block of an invariant tensor is a tuple of
# sector indices whose charges sum to zero.  Contracted legs are duals of
# each other: same order, negated charges.  Conjugating a tensor negates the
# charges of all its flat legs.
# ---------------------------------------------------------------------------


class SectorLeg:
    def __init__(self, sectors):
        self.sectors = [(tuple(int(x) for x in q), int(d)) for q, d in sectors]
        self._hash = hash(tuple(self.sectors))

    @property
    def dim(self):
        return sum(d for _, d in self.sectors)

    def dual(self):
        return SectorLeg([(tuple(-x for x in q), d) for q, d in self.sectors])

    def charges(self):
        return [q for q, _ in self.sectors]

    def degs(self):
        return [d for _, d in self.sectors]

    def __eq__(self, other):
        return isinstance(other, SectorLeg) and self._hash == other._hash and self.sectors == other.sectors

    def __hash__(self):
        return self._hash

    def __repr__(self):
        return f"SectorLeg({self.sectors})"


def _zero_like(q):
    return tuple(0 for _ in q)


_CHARGE_DIST = {}
_CODE_BASE = 1 << 21        # charge components stay far inside (-2^20, 2^20)


def _leg_codes(leg):
    """(charge code, degeneracy) per sector: code(q) = sum_i q_i B^i is linear
    in q and zero only for q = 0, so charge sums become integer sums."""
    c = getattr(leg, "_codes", None)
    if c is None:
        c = leg._codes = [(sum(int(x) * _CODE_BASE ** i for i, x in enumerate(q)), d)
                          for q, d in leg.sectors]
    return c


_CHARGE_ARR = {}             # leg tuple -> (sorted codes, counts) as int64 arrays
_EXACT_INT64 = 1 << 62


def _leg_arrays(leg):
    a = getattr(leg, "_code_arrays", None)
    if a is None:
        codes = _leg_codes(leg)
        a = leg._code_arrays = (np.array([c for c, _ in codes], dtype=np.int64),
                                np.array([d for _, d in codes], dtype=np.int64))
    return a


def charge_distribution(legs):
    """dict total-charge code -> number of elements over all sector tuples.

    Charges are integer codes (``_leg_codes``; the zero charge is 0, negation
    is integer negation).  Memoized on the leg tuple; a tuple extends the
    distribution of its prefix, so the planner's overlapping sub-networks
    share the work.  Each extension is an outer sum of codes and outer
    product of counts aggregated by code in numpy (exact int64; Python
    integers when a count could pass 2^62).  The returned dict is shared:
    callers must not modify it."""
    key = tuple(legs)
    hit = _CHARGE_DIST.get(key)
    if hit is not None:
        return hit
    if not key:
        return {}
    if len(_CHARGE_DIST) > 200000:
        _CHARGE_DIST.clear()
        _CHARGE_ARR.clear()
    if len(key) > 1:
        base = charge_distribution(key[:-1])
        bc, bw = _CHARGE_ARR.get(key[:-1], (None, None))
    else:
        base = {0: 1}
        bc, bw = np.zeros(1, dtype=np.int64), np.ones(1, dtype=np.int64)
    lc, lw = _leg_arrays(key[-1])
    if bc is not None and int(bw.max()) * int(lw.max()) < _EXACT_INT64:
        c = (bc[:, None] + lc[None, :]).ravel()
        w = (bw[:, None] * lw[None, :]).ravel()
        uc, inv = np.unique(c, return_inverse=True)
        uw = np.zeros(len(uc), dtype=np.int64)
        np.add.at(uw, inv, w)
        _CHARGE_ARR[key] = (uc, uw)
        new = dict(zip(uc.tolist(), uw.tolist()))
    else:
        new = {}
        get = new.get
        for c0, w0 in base.items():
            for q, d in _leg_codes(key[-1]):
                t = c0 + q
                new[t] = get(t, 0) + w0 * d
    _CHARGE_DIST[key] = new
    return new


def stored_elements(legs):
    if not legs:
        return 1.0
    return float(charge_distribution(legs).get(0, 0))


def block_count(legs):
    """Number of coupled-sector blocks of a rank-2 view (codomain = legs[:-1])."""
    if len(legs) < 2:
        return 1
    rows = charge_distribution(legs[:-1])
    cols = charge_distribution([legs[-1]])
    return sum(1 for c in rows if -c in cols)


def conj_legs(legs, conj):
    return [l.dual() for l in legs] if conj else list(legs)


def sector_pair_profile(legsA, oA, cA, legsB, oB):
    """Backend.pair_cost block loop (Backend.jl:328-368) in the flat-leg model.

    Returns (sector_flops = sum_q 2 m_q k_q n_q, output legs, output stored
    elements, largest GEMM block, number of blocks).  Memoized on the three
    leg tuples it depends on: the planner's search meets most of them again."""
    mlegs = tuple(legsA[i] for i in oA)
    klegs = tuple(legsA[i] for i in cA)
    nlegs = tuple(legsB[i] for i in oB)
    key = (mlegs, klegs, nlegs)
    hit = _PAIR_PROFILE.get(key)
    if hit is None:
        if len(_PAIR_PROFILE) > 200000:
            _PAIR_PROFILE.clear()
        hit = _PAIR_PROFILE[key] = _pair_profile(mlegs, klegs, nlegs)
    flops, out_legs, out_el, largest, nblocks = hit
    return flops, list(out_legs), out_el, largest, nblocks


_PAIR_PROFILE = {}


def _pair_profile(mlegs, klegs, nlegs):
    mdist = charge_distribution(mlegs) if mlegs else None
    kdist = charge_distribution(klegs) if klegs else None
    ndist = charge_distribution(nlegs) if nlegs else None
    mdist = mdist or {0: 1}
    kdist = kdist or {0: 1}
    ndist = ndist or {0: 1}
    flops = 0.0
    out_el = 0.0
    largest = 0.0
    nblocks = 0
    for c, m in mdist.items():
        neg = -c
        k = kdist.get(neg, 0)
        n = ndist.get(neg, 0)
        if k and n:
            flops += 2.0 * m * k * n
            nblocks += 1
        if n:
            out_el += m * n
            largest = max(largest, float(m * n))
    return flops, mlegs + nlegs, out_el, largest, nblocks


# ---------------------------------------------------------------------------
# Dense cost — cost.jl:148-174, Backend.jl:432-441
# ---------------------------------------------------------------------------


def _prod(xs):
    p = 1.0
    for x in xs:
        p *= float(x)
    return p


def dense_cost(labelsA, dimsA, labelsB, dimsB):
    shared = set(labelsA) & set(labelsB)
    oa = [i for i, l in enumerate(labelsA) if l not in shared]
    ca = [i for i, l in enumerate(labelsA) if l in shared]
    ob = [i for i, l in enumerate(labelsB) if l not in shared]
    cb = []
    for ia in ca:
        ib = labelsB.index(labelsA[ia])
        if dimsA[ia] != dimsB[ib]:
            raise ValueError(f"incompatible contracted dimensions for label {labelsA[ia]}")
        cb.append(ib)
    out_labels = [labelsA[i] for i in oa] + [labelsB[i] for i in ob]
    out_dims = [dimsA[i] for i in oa] + [dimsB[i] for i in ob]
    flops = _prod(dimsA[i] for i in oa) * _prod(dimsA[i] for i in ca) * _prod(dimsB[i] for i in ob)
    return dict(oindA=oa, cindA=ca, oindB=ob, cindB=cb, labels=out_labels, dims=out_dims,
                flops=flops, peak_elements=_prod(out_dims))


def pair_workload_profile(step, dimsA, dimsB):
    """(m, k, n) of one dense pair step — Backend.jl:377-404."""
    m = int(_prod(dimsA[i] for i in step.oindA))
    k = int(_prod(dimsA[i] for i in step.cindA))
    n = int(_prod(dimsB[i] for i in step.oindB))
    return m, k, n


# ---------------------------------------------------------------------------
# Trees — planner.jl:1-66 and the exact DPs
# ---------------------------------------------------------------------------


def _heuristic_tree(spec):
    slots = spec.preferred_slots
    if spec.dynamic_slot is None:
        tree = slots[0]
        rest = slots[1:]
    else:
        tree = spec.dynamic_slot
        rest = slots
    for s in rest:
        tree = (tree, s)
    return tree


def _shares_positive(a, b):
    sb = set(b)
    return any(l > 0 and l in sb for l in a)


def _optimal_tree(spec, dims):
    """Exact minimum-FLOP tree by DP over subsets (the objective of optimaltree)."""
    n = len(spec.labels)
    full = (1 << n) - 1
    best = {}
    for i in range(n):
        best[1 << i] = (0.0, i, list(spec.labels[i]), list(dims[i]))
    for mask in range(1, full + 1):
        if bin(mask).count("1") < 2:
            continue
        cand = None
        sub = (mask - 1) & mask
        while sub:
            other = mask ^ sub
            if sub < other and sub in best and other in best:
                fa, ta, la, da = best[sub]
                fb, tb, lb, db = best[other]
                if _shares_positive(la, lb):
                    m = dense_cost(la, da, lb, db)
                    cost = fa + fb + m["flops"]
                    if cand is None or cost < cand[0]:
                        cand = (cost, (ta, tb), m["labels"], m["dims"])
            sub = (sub - 1) & mask
        if cand is not None:
            best[mask] = cand
    if full not in best:
        return None
    return best[full][1]


def _greedy_tree(spec, dims):
    """planner.jl:28-66: merge the pair with the smallest (peak, flops)."""
    nodes = [(i, list(spec.labels[i]), list(dims[i])) for i in range(len(spec.labels))]
    while len(nodes) > 1:
        best = None
        connected = False
        for i in range(len(nodes) - 1):
            for j in range(i + 1, len(nodes)):
                if not (set(nodes[i][1]) & set(nodes[j][1])):
                    continue
                connected = True
                m = dense_cost(nodes[i][1], nodes[i][2], nodes[j][1], nodes[j][2])
                score = (m["peak_elements"], m["flops"])
                if best is None or score < best[0]:
                    best = (score, i, j, m)
        if not connected:
            i, j = len(nodes) - 2, len(nodes) - 1
            m = dense_cost(nodes[i][1], nodes[i][2], nodes[j][1], nodes[j][2])
            best = (None, i, j, m)
        _, i, j, m = best
        merged = ((nodes[i][0], nodes[j][0]), m["labels"], m["dims"])
        nodes[i] = merged
        del nodes[j]
    return nodes[0][0]


def _canonical_order(labels):
    """planner.jl:250-258: contracted labels ascending, then open -1, -2, ..."""
    return sorted(range(len(labels)), key=lambda i: (0, labels[i]) if labels[i] > 0 else (1, -labels[i]))


def _sector_dp_trees(spec, dims, legs):
    """planner.jl:285-330: Pareto frontier DP over connected merges."""
    n = len(spec.labels)
    if n > EXACT_TENSOR_LIMIT:
        return []
    full = (1 << n) - 1
    frontiers = {}
    for i in range(n):
        L = conj_legs(legs[i], spec.conjs[i])
        el = stored_elements(L)
        frontiers[1 << i] = [dict(tree=i, labels=list(spec.labels[i]), dims=list(dims[i]), legs=L,
                                  flops=0.0, peak=_prod(dims[i]), sflops=0.0,
                                  speak=el, el=el, dev=0.0)]
    for mask in range(1, full + 1):
        if bin(mask).count("1") < 2:
            continue
        front = []
        left = (mask - 1) & mask
        while left:
            right = mask ^ left
            if right and left < right and frontiers.get(left) and frontiers.get(right):
                for a in frontiers[left]:
                    for b in frontiers[right]:
                        if not _shares_positive(a["labels"], b["labels"]):
                            continue
                        m = dense_cost(a["labels"], a["dims"], b["labels"], b["dims"])
                        sfl, olegs, oel, _, _ = sector_pair_profile(a["legs"], m["oindA"], m["cindA"],
                                                                    b["legs"], m["oindB"])
                        order = _canonical_order(m["labels"])
                        st = dict(tree=(a["tree"], b["tree"]),
                                  labels=[m["labels"][i] for i in order],
                                  dims=[m["dims"][i] for i in order],
                                  legs=[olegs[i] for i in order],
                                  flops=a["flops"] + b["flops"] + m["flops"],
                                  peak=max(a["peak"], b["peak"], m["peak_elements"]),
                                  sflops=a["sflops"] + b["sflops"] + sfl,
                                  speak=max(a["speak"], b["speak"], oel), el=oel,
                                  dev=a["dev"] + b["dev"] + step_device_time(sfl, a["el"] + b["el"] + oel))
                        _insert_pareto(front, st)
            left = (left - 1) & mask
        if front:
            frontiers[mask] = front
    return [s["tree"] for s in frontiers.get(full, [])]


def _insert_pareto(front, cand):
    key = (tuple(cand["labels"]), tuple(cand["dims"]), tuple(cand["legs"]))

    def no_worse(a, b):
        return (a["peak"] <= b["peak"] and a["speak"] <= b["speak"] and a["sflops"] <= b["sflops"]
                and a["dev"] <= b["dev"])

    same = []
    for i, cur in enumerate(front):
        if (tuple(cur["labels"]), tuple(cur["dims"]), tuple(cur["legs"])) != key:
            continue
        if no_worse(cur, cand):
            return
        same.append(i)
    for i in reversed(same):
        if no_worse(cand, front[i]):
            del front[i]
    front.append(cand)


# ---------------------------------------------------------------------------
# Compilation — planner.jl:362-433 (_compile_tree!), :455-512 (live memory)
# ---------------------------------------------------------------------------


def _output_partition(labels, spec):
    wanted = [-(i + 1) for i in range(spec.nopen)]
    if len(labels) != len(wanted):
        raise ValueError(f"planned root has {len(labels)} open legs; expected {len(wanted)}")
    pos = [labels.index(l) for l in wanted]
    nout, nin = spec.out_partition
    return (tuple(pos[:nout]), tuple(pos[nout:nout + nin]))


def _is_identity(order, n):
    return len(order) == n and all(i == j for i, j in enumerate(order))


def compile_plan(tree, spec, dims, legs=None, strategy="env_first", canonical=False,
                 scalar_bytes=16, fused_permutes=False):
    """Walk a binary tree into PairSteps with dense and sector metrics.

    ``fused_permutes=True`` models an executor that folds every permutation
    and conjugation into its GEMM loads/stores (the device engine): no
    transform temporaries are charged to the live-memory model.  False is the
    reference's TensorOperations model (cost.jl:82-122)."""
    n = len(dims)
    structural = legs is not None
    st = dict(nextslot=n - 1, steps=[], flops=0.0, peak=max(_prod(d) for d in dims),
              sflops=0.0, speak=0.0, sbpeak=0.0, dev=0.0, out_dense=[], out_sector=[],
              tmp_dense=[], perm_dense=[], tmp_sector=[], perm_sector=[])
    slot_legs = [conj_legs(legs[i], spec.conjs[i]) for i in range(n)] if structural else None
    if structural:
        st["speak"] = max(stored_elements(L) for L in slot_legs)
        st["sbpeak"] = st["speak"]
    else:
        st["speak"] = st["peak"]
        st["sbpeak"] = st["peak"]

    def rec(t, root=False):
        if isinstance(t, int):
            d = list(dims[t])
            L = slot_legs[t] if structural else None
            dense_payload = _prod(d)
            sector_payload = stored_elements(L) if structural else dense_payload
            return t, list(spec.labels[t]), d, L, spec.conjs[t], dense_payload, sector_payload
        a, la, da, La, ca, dna, sca = rec(t[0])
        b, lb, db, Lb, cb, dnb, scb = rec(t[1])
        m = dense_cost(la, da, lb, db)
        st["nextslot"] += 1
        dst = st["nextslot"]
        if root:
            out = _output_partition(m["labels"], spec)
        elif fused_permutes:
            # Device layout (for sector trees too; their canonical order only
            # keys the search): the larger operand's open legs first, so the
            # new intermediate keeps that operand's unit-stride leg and both
            # the GEMM / streaming stores and the next step's loads stay
            # coalesced.
            na, nb = len(m["oindA"]), len(m["oindB"])
            if _prod(db[i] for i in m["oindB"]) > _prod(da[i] for i in m["oindA"]):
                out = (tuple(range(na, na + nb)) + tuple(range(na)), ())
            else:
                out = (tuple(range(na + nb)), ())
        elif canonical:
            out = (tuple(_canonical_order(m["labels"])), ())
        else:
            out = (tuple(range(len(m["labels"]))), ())
        flat = list(out[0]) + list(out[1])
        out_labels = m["labels"] if root else [m["labels"][i] for i in flat]
        out_dims = m["dims"] if root else [m["dims"][i] for i in flat]
        dense_out = m["peak_elements"]
        if structural:
            sfl, olegs, oel, largest, _ = sector_pair_profile(La, m["oindA"], m["cindA"], Lb, m["oindB"])
            out_legs = [olegs[i] for i in flat]
            sector_out = oel
        else:
            out_legs = None
            sector_out = dense_out
        pA = list(m["oindA"]) + list(m["cindA"])
        pB = list(m["cindB"]) + list(m["oindB"])
        perm_d = perm_s = 0.0
        if not _is_identity(pA, len(da)):
            perm_d += dna
            perm_s += sca
        if not _is_identity(pB, len(db)):
            perm_d += dnb
            perm_s += scb
        if not _is_identity(flat, len(m["labels"])):
            perm_d += dense_out
            perm_s += sector_out
        if fused_permutes:
            perm_d = perm_s = 0.0
            st["tmp_dense"].append(0.0)
            st["tmp_sector"].append(0.0)
        else:
            st["tmp_dense"].append(perm_d + (dna if ca else 0.0) + (dnb if cb else 0.0))
            st["tmp_sector"].append(perm_s + (sca if ca else 0.0) + (scb if cb else 0.0))
        st["perm_dense"].append(perm_d)
        st["perm_sector"].append(perm_s)
        st["steps"].append(PairStep(a, b, dst, tuple(m["oindA"]), tuple(m["cindA"]),
                                    tuple(m["oindB"]), tuple(m["cindB"]), bool(ca), bool(cb), out))
        st["out_dense"].append(dense_out)
        st["out_sector"].append(sector_out)
        st["flops"] += m["flops"]
        st["peak"] = max(st["peak"], m["peak_elements"])
        if structural:
            st["sflops"] += sfl
            st["speak"] = max(st["speak"], sector_out)
            st["sbpeak"] = max(st["sbpeak"], largest)
            st["dev"] += step_device_time(sfl, sca + scb + sector_out, scalar_bytes)
        else:
            st["dev"] += step_device_time(2 * m["flops"], dna + dnb + dense_out, scalar_bytes)
            st["sflops"] += 2 * m["flops"]
            st["speak"] = max(st["speak"], m["peak_elements"])
            st["sbpeak"] = max(st["sbpeak"], m["peak_elements"])
        return dst, out_labels, out_dims, out_legs, False, dense_out, sector_out

    output = rec(tree, root=True)[0]
    # live-memory model
    dense_ops = sum(_prod(d) for d in dims)
    sector_ops = sum(stored_elements(L) for L in slot_legs) if structural else dense_ops
    live_d, live_s = {}, {}
    peak_d, peak_s = dense_ops, sector_ops
    tmp_pk = perm_pk = stmp_pk = sperm_pk = 0.0
    for i, s in enumerate(st["steps"]):
        peak_d = max(peak_d, dense_ops + sum(live_d.values()) + st["out_dense"][i] + st["tmp_dense"][i])
        peak_s = max(peak_s, sector_ops + sum(live_s.values()) + st["out_sector"][i] + st["tmp_sector"][i])
        tmp_pk = max(tmp_pk, st["tmp_dense"][i])
        perm_pk = max(perm_pk, st["perm_dense"][i])
        stmp_pk = max(stmp_pk, st["tmp_sector"][i])
        sperm_pk = max(sperm_pk, st["perm_sector"][i])
        for k in (s.a, s.b):
            live_d.pop(k, None)
            live_s.pop(k, None)
        live_d[s.dst] = st["out_dense"][i]
        live_s[s.dst] = st["out_sector"][i]
    sb = float(scalar_bytes)
    plan = ContractionPlan(
        nslots=st["nextslot"] + 1, output_slot=output, steps=st["steps"], strategy=strategy,
        flops=st["flops"], peak_elements=st["peak"], sector_flops=st["sflops"],
        sector_peak_elements=st["speak"], sector_peak_block_elements=st["sbpeak"],
        scalar_bytes=scalar_bytes, operand_bytes=dense_ops * sb, live_peak_bytes=peak_d * sb,
        known_temporary_peak_bytes=tmp_pk * sb, known_permutation_peak_bytes=perm_pk * sb,
        sector_operand_bytes=sector_ops * sb, sector_live_peak_bytes=peak_s * sb,
        sector_known_temporary_peak_bytes=stmp_pk * sb,
        sector_known_permutation_peak_bytes=sperm_pk * sb,
        scalar_output=spec.nopen == 0)
    plan.tree = tree
    plan.device_time = st["dev"]
    return plan


def _score(plan, mw):
    return plan.flops + mw * plan.live_peak_bytes


def _sector_score(plan, mw):
    return plan.sector_flops + mw * plan.sector_live_peak_bytes


def _device_score(plan, mw):
    """Modelled time in sector-FLOP units (DEVICE_FLOPS / 4 per second), so
    ``memory_weight`` keeps its meaning of one sector FLOP per live byte."""
    return plan.device_time * DEVICE_FLOPS / 4.0 + mw * plan.sector_live_peak_bytes


def _within_floors(c, h):
    return (c.peak_elements <= h.peak_elements and c.sector_peak_elements <= h.sector_peak_elements
            and c.live_peak_bytes <= h.live_peak_bytes
            and c.sector_live_peak_bytes <= h.sector_live_peak_bytes)


def _within_cap(p, cap):
    return p.live_peak_bytes <= cap and p.sector_live_peak_bytes <= cap


def _sector_nontrivial(legs_list):
    for legs in legs_list:
        if not legs:
            continue
        if any(len(l.sectors) > 1 for l in legs):
            dense = _prod(l.dim for l in legs)
            if stored_elements(legs) != dense or block_count(legs) > 1:
                return True
    return False


def plan_contraction(spec, dims, legs=None, optimize=True, memory_weight=1.0, sector_aware=True,
                     scalar_bytes=16, memory_cap_bytes=None, fused_permutes=False, order_cache=None):
    """plan_contraction — planner.jl:584-700.

    ``dims[i]`` are the flat leg dimensions of slot i; ``legs[i]`` (optional)
    its abelian SectorLeg list.  ``fused_permutes`` selects the device
    executor's live-memory model and layouts (see ``compile_plan``) and, for
    sector networks, ranks the admissible candidates by the device time model
    (``step_device_time``) plus ``memory_weight`` times live bytes.  ``order_cache``
    (a dict) reuses the pairwise order chosen for a sector network of the same
    spec and leg dimensions within a factor of two: sweeps meet a new sector
    structure on almost every bond, and the sector search costs far more than
    compiling a known order for the new structure."""
    if not (math.isfinite(memory_weight) and memory_weight >= 0):
        raise ValueError("memory_weight must be finite and nonnegative")
    cap = math.inf if memory_cap_bytes is None else float(memory_cap_bytes)
    for i, (d, l) in enumerate(zip(dims, spec.labels)):
        if len(d) != len(l):
            raise ValueError(f"prototype {i} has {len(d)} legs; spec has {len(l)}")
    label_dims = {}
    for d, ls in zip(dims, spec.labels):
        for x, l in zip(d, ls):
            if label_dims.setdefault(l, x) != x:
                raise ValueError(f"label {l} has incompatible dimensions {label_dims[l]} and {x}")
    structural = optimize and legs is not None and _sector_nontrivial(legs)
    slegs = legs if structural else None
    okey = None
    if order_cache is not None and structural:
        okey = (spec.key(), tuple(tuple(int(round(math.log2(max(x, 1)))) for x in d) for d in dims),
                memory_weight, sector_aware, cap, fused_permutes, scalar_bytes)
        hit = order_cache.get(okey)
        if hit is not None:
            tree, strategy = hit
            p = compile_plan(tree, spec, dims, slegs, strategy, canonical=strategy == "sector_exact",
                             scalar_bytes=scalar_bytes, fused_permutes=fused_permutes)
            if _within_cap(p, cap):
                return p
    heuristic = compile_plan(_heuristic_tree(spec), spec, dims, slegs, "env_first",
                             scalar_bytes=scalar_bytes, fused_permutes=fused_permutes)
    if not optimize:
        if not _within_cap(heuristic, cap):
            raise ValueError("env-first contraction plan requires at least "
                             f"{max(heuristic.live_peak_bytes, heuristic.sector_live_peak_bytes)} "
                             f"bytes, above memory_cap_bytes={cap}")
        return heuristic
    candidates = [heuristic]
    use_sector = sector_aware and structural
    if use_sector:
        for t in _sector_dp_trees(spec, dims, legs):
            candidates.append(compile_plan(t, spec, dims, slegs, "sector_exact", canonical=True,
                                           scalar_bytes=scalar_bytes, fused_permutes=fused_permutes))
    if len(spec.labels) <= EXACT_TENSOR_LIMIT:
        t = _optimal_tree(spec, dims)
        if t is not None:
            candidates.append(compile_plan(t, spec, dims, slegs, "dense_optimal",
                                           scalar_bytes=scalar_bytes, fused_permutes=fused_permutes))
        candidates.append(compile_plan(_greedy_tree(spec, dims), spec, dims, slegs, "memory_greedy",
                                       scalar_bytes=scalar_bytes, fused_permutes=fused_permutes))
    else:
        candidates.append(compile_plan(_greedy_tree(spec, dims), spec, dims, slegs, "dense_greedy",
                                       scalar_bytes=scalar_bytes, fused_permutes=fused_permutes))
    best, best_score = None, math.inf
    for c in candidates:
        if not _within_floors(c, heuristic) or not _within_cap(c, cap):
            continue
        if use_sector:
            s = (_device_score if fused_permutes else _sector_score)(c, memory_weight)
        else:
            s = _score(c, memory_weight)
        if s < best_score:
            best, best_score = c, s
    if best is None:
        raise ValueError(f"no contraction plan fits memory_cap_bytes={cap}; env-first requires at "
                         f"least {max(heuristic.live_peak_bytes, heuristic.sector_live_peak_bytes)} bytes")
    if okey is not None:
        if len(order_cache) > 4096:
            order_cache.clear()
        order_cache[okey] = (best.tree, best.strategy)
    return best


def slot_dims(plan, dims):
    """Flat leg dims of every slot of a compiled plan."""
    out = list(dims) + [None] * len(plan.steps)
    for s in plan.steps:
        da, db = out[s.a], out[s.b]
        inter = [da[i] for i in s.oindA] + [db[i] for i in s.oindB]
        out[s.dst] = [inter[i] for i in list(s.out[0]) + list(s.out[1])]
    return out


def plan_gemm_profile(plan, dims):
    """Per-step (m, k, n) list — the reference's pair_workload_profile histogram."""
    sd = slot_dims(plan, dims)
    return [pair_workload_profile(s, sd[s.a], sd[s.b]) for s in plan.steps]
