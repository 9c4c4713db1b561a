"""SU(2)-symmetric block tensors: the non-abelian sector blocks of the
effective-Hamiltonian path.

A TensorMap with SU(2) symmetry (TensorKit's ``Vect[SU2Irrep]`` spaces,
re-exported by the reference in lib/GraftFoundation/src/Backend/Backend.jl:36)
stores one dense block over the legs' degeneracies for every tuple of leg
irreps that can fuse to the trivial irrep and every fusion tree of that
tuple.  Here a tensor on legs V_0 .. V_{n-1} is

    T = sum_t  T_t (x) X_t,

t running over (irrep tuple, tree) pairs: T_t is the degeneracy block (what
the device stores), X_t the structural tensor of tree t — the
Clebsch-Gordan couplings ((j_0 j_1) k_1 j_2) k_2 ... j_{n-1} -> 0, left to
right, with the map eps (eps U eps^-1 = conj(U)) applied on every dual leg.
The X_t are real and orthonormal, so the inner product of two tensors is
the sum of their blocks' inner products and the Krylov algebra runs on the
flat block data unchanged.

A pairwise contraction is, block pair by block pair,

    C_tc += c(ta, tb, tc) * contract(op(A_ta), op(B_tb)),
    c(ta, tb, tc) = < X_tc , contract(X_ta, X_tb) >,

because contract(X_ta, X_tb) is an invariant tensor on C's legs and the X_tc
are an orthonormal basis of those.  TensorKit reaches the same result by
fusion-tree transformations (F-moves, R-moves) followed by one GEMM per
coupled sector inside Backend.contract_pair! (Backend.jl:209-222); here the
coefficients come straight from the Clebsch-Gordan tensors, so no 6j or
phase convention is involved, and ``SU2Structure.to_dense`` (the embedding
into the plain legs) checks every result.  The device runs every step of a
contraction plan as one grouped GEMM launch over its block pairs, each pair
scaled by its coefficient as the A tile is staged
(``graft_plan_create_coupled``).  Planning uses the dense leg dimensions,
as the reference's planner does for non-unique fusion
(lib/GraftPlanning/src/Planning/planner.jl:566-583).
"""
from __future__ import annotations

import ctypes
import functools
import itertools
import math
from fractions import Fraction

import numpy as np

from . import _native as nat

# ---------------------------------------------------------------------------
# Spin algebra (arguments are twice the spin: tj = 2j, tm = 2m)
# ---------------------------------------------------------------------------


def fusion(tj1, tj2):
    """Irreps in j1 (x) j2: |j1 - j2| .. j1 + j2 in integer steps."""
    return range(abs(tj1 - tj2), tj1 + tj2 + 1, 2)


@functools.lru_cache(maxsize=None)
def clebsch_gordan(tj1, tm1, tj2, tm2, tJ, tM):
    """<j1 m1; j2 m2 | J M> (Condon-Shortley phases) by Racah's closed form,
    evaluated in exact rational arithmetic and rounded once."""
    if tm1 + tm2 != tM or (tj1 + tj2 + tJ) % 2 or not abs(tj1 - tj2) <= tJ <= tj1 + tj2:
        return 0.0
    for tj, tm in ((tj1, tm1), (tj2, tm2), (tJ, tM)):
        if abs(tm) > tj or (tj + tm) % 2:
            return 0.0
    f = math.factorial
    a = (tj1 + tj2 - tJ) // 2
    b = (tj1 - tj2 + tJ) // 2
    c = (-tj1 + tj2 + tJ) // 2
    j1m, j1p = (tj1 - tm1) // 2, (tj1 + tm1) // 2
    j2m, j2p = (tj2 - tm2) // 2, (tj2 + tm2) // 2
    Jm, Jp = (tJ - tM) // 2, (tJ + tM) // 2
    x1 = (tJ - tj2 + tm1) // 2
    x2 = (tJ - tj1 - tm2) // 2
    pre = Fraction((tJ + 1) * f(a) * f(b) * f(c), f((tj1 + tj2 + tJ) // 2 + 1))
    pre *= f(Jp) * f(Jm) * f(j1m) * f(j1p) * f(j2m) * f(j2p)
    s = Fraction(0)
    for k in range(max(0, -x1, -x2), min(a, j1m, j2p) + 1):
        s += Fraction((-1) ** k, f(k) * f(a - k) * f(j1m - k) * f(j2p - k) * f(x1 + k) * f(x2 + k))
    if s == 0:
        return 0.0
    return math.copysign(math.sqrt(pre * s * s), s)


@functools.lru_cache(maxsize=None)
def cg_tensor(tj1, tj2, tJ):
    """Clebsch-Gordan tensor [m1, m2, M]; index i of a spin-j axis is m = j - i."""
    out = np.zeros((tj1 + 1, tj2 + 1, tJ + 1))
    for i1 in range(tj1 + 1):
        for i2 in range(tj2 + 1):
            tM = (tj1 - 2 * i1) + (tj2 - 2 * i2)
            if abs(tM) <= tJ:
                out[i1, i2, (tJ - tM) // 2] = clebsch_gordan(tj1, tj1 - 2 * i1, tj2, tj2 - 2 * i2,
                                                            tJ, tM)
    out.setflags(write=False)
    return out


@functools.lru_cache(maxsize=None)
def epsilon(tj):
    """eps[m, m'] = (-1)^(j - m) delta(m, -m'): eps D(g) eps^-1 = conj(D(g))."""
    e = np.zeros((tj + 1, tj + 1))
    for i in range(tj + 1):
        e[i, tj - i] = -1.0 if i % 2 else 1.0
    e.setflags(write=False)
    return e


def spin_matrices(tj):
    """(Jx, Jy, Jz) of spin j in the basis m = j, j-1, .., -j."""
    m = np.array([(tj - 2 * i) / 2 for i in range(tj + 1)])
    j = tj / 2
    jp = np.zeros((tj + 1, tj + 1))
    for i in range(1, tj + 1):
        jp[i - 1, i] = math.sqrt((j - m[i]) * (j + m[i] + 1))
    jm = jp.T
    return (jp + jm) / 2, (jp - jm) / 2j, np.diag(m)


@functools.lru_cache(maxsize=None)
def fusion_trees(tjs):
    """Internal labels of every left-to-right fusion tree of irreps ``tjs`` to
    the trivial irrep: (k_1, .., k_{n-2}), k_l the coupled spin of legs 0..l."""
    n = len(tjs)
    if n == 0:
        return ((),)
    if n == 1:
        return ((),) if tjs[0] == 0 else ()
    # reach[l]: total spins legs l.. can fuse to
    reach = [None] * (n + 1)
    reach[n] = {0}
    for l in range(n - 1, 0, -1):
        reach[l] = {s for r in reach[l + 1] for s in fusion(tjs[l], r)}
    out = []

    def walk(l, k, labels):
        if l == n - 1:
            if k == tjs[n - 1]:
                out.append(tuple(labels))
            return
        for kn in fusion(k, tjs[l]):
            if kn in reach[l + 1]:
                walk(l + 1, kn, labels + [kn])

    if tjs[0] in reach[1]:
        walk(1, tjs[0], [])
    return tuple(out)


# Structural tensors hold prod(2j + 1) entries: the caches are bounded by
# count (large spins on many legs make single entries large).
@functools.lru_cache(maxsize=4096)
def tree_tensor(tjs, labels, duals):
    """Structural tensor X_t (real, unit Frobenius norm) of one fusion tree."""
    n = len(tjs)
    if n == 0:
        return np.ones(())
    X = np.eye(tjs[0] + 1)
    k = tjs[0]
    for l in range(1, n):
        kn = labels[l - 1] if l < n - 1 else 0
        X = np.tensordot(X, cg_tensor(k, tjs[l], kn), axes=([X.ndim - 1], [0]))
        k = kn
    X = X[..., 0]
    for l in range(n):
        if duals[l]:
            X = np.moveaxis(np.tensordot(epsilon(tjs[l]), X, axes=([1], [l])), 0, l)
    X = np.ascontiguousarray(X)
    X.setflags(write=False)
    return X


@functools.lru_cache(maxsize=1024)
def tree_basis(tjs, duals):
    """(trees, matrix [ntrees, prod(2j+1)]) of all trees of one irrep tuple."""
    trees = fusion_trees(tjs)
    size = int(np.prod([t + 1 for t in tjs])) if tjs else 1
    M = np.zeros((len(trees), size))
    for i, tr in enumerate(trees):
        M[i] = tree_tensor(tjs, tr, duals).reshape(-1)
    M.setflags(write=False)
    return trees, M


# ---------------------------------------------------------------------------
# Legs and block structures
# ---------------------------------------------------------------------------


class SU2Leg:
    """One flat leg: irreps (twice the spin) with degeneracies; ``isdual``
    marks a dual (domain / bra-side) leg, which transforms with conj(U)."""

    __slots__ = ("irreps", "isdual", "_hash")

    def __init__(self, irreps, isdual=False):
        irr = tuple(sorted((int(tj), int(d)) for tj, d in irreps))
        if not irr or any(d <= 0 or tj < 0 for tj, d in irr) or len({tj for tj, _ in irr}) != len(irr):
            raise ValueError(f"SU2Leg: irreps must be distinct spins with positive degeneracy: {irreps}")
        self.irreps = irr
        self.isdual = bool(isdual)
        self._hash = hash((self.irreps, self.isdual))

    @property
    def dim(self):
        return sum(d * (tj + 1) for tj, d in self.irreps)

    def dual(self):
        return SU2Leg(self.irreps, not self.isdual)

    def degs(self):
        return [d for _, d in self.irreps]

    def spins(self):
        return [tj for tj, _ in self.irreps]

    def offsets(self):
        """Start of every irrep's (degeneracy x multiplet) range on the dense leg."""
        out, o = [], 0
        for tj, d in self.irreps:
            out.append(o)
            o += d * (tj + 1)
        return out

    def __eq__(self, other):
        return isinstance(other, SU2Leg) and self.irreps == other.irreps and self.isdual == other.isdual

    def __hash__(self):
        return self._hash

    def __repr__(self):
        s = ", ".join(f"{tj}/2:{d}" if tj % 2 else f"{tj // 2}:{d}" for tj, d in self.irreps)
        return f"SU2Leg({s}{', dual' if self.isdual else ''})"


class SU2Structure:
    """Block layout of an SU(2) tensor: one degeneracy block per (irrep index
    tuple, fusion tree), packed in list order, each first leg fastest.

    ``dims`` are the dense (embedded) leg dimensions; the device descriptor
    carries the degeneracy extents, with one entry per block (blocks of one
    irrep tuple and different trees repeat its sector tuple)."""

    def __init__(self, legs, blocks):
        self.legs = list(legs)
        self.blocks = [(tuple(int(x) for x in ir), tuple(int(x) for x in tr)) for ir, tr in blocks]
        self.index = {b: i for i, b in enumerate(self.blocks)}
        if len(self.index) != len(self.blocks):
            raise ValueError("SU2Structure: duplicate block")
        degs = [l.degs() for l in self.legs]
        self.block_dims = [tuple(degs[l][s] for l, s in enumerate(ir)) for ir, _ in self.blocks]
        self.sizes = [int(np.prod(d)) if d else 1 for d in self.block_dims]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64) \
            if self.sizes else np.zeros(0, np.int64)
        self.total = int(sum(self.sizes))

    @property
    def dims(self):
        return tuple(l.dim for l in self.legs)

    @property
    def duals(self):
        return tuple(l.isdual for l in self.legs)

    def spins(self, b):
        ir = self.blocks[b][0]
        return tuple(self.legs[l].irreps[s][0] for l, s in enumerate(ir))

    @staticmethod
    def invariant(legs):
        """Every (irrep tuple, tree) with a fusion channel to the trivial irrep."""
        legs = list(legs)
        blocks = []
        for ir in itertools.product(*[range(len(l.irreps)) for l in legs]):
            tjs = tuple(legs[l].irreps[s][0] for l, s in enumerate(ir))
            for tr in fusion_trees(tjs):
                blocks.append((ir, tr))
        return SU2Structure(legs, sorted(blocks))

    def key(self):
        k = getattr(self, "_key", None)
        if k is None:
            k = self._key = ("su2", tuple(self.legs), tuple(self.blocks))
        return k

    def desc(self):
        d = getattr(self, "_desc", None)
        if d is None and not self.legs:
            d = self._desc = nat.DescKeepalive([])       # a scalar: one block of one element
        if d is None:
            d = self._desc = nat.DescKeepalive(
                [sum(l.degs()) for l in self.legs],
                ([l.degs() for l in self.legs],
                 np.asarray([ir for ir, _ in self.blocks], dtype=np.int32).reshape(len(self.blocks), -1),
                 self.offsets, None))
        return d

    # -- dense embedding ------------------------------------------------------
    def _slices(self, ir):
        return tuple(slice(self.legs[l].offsets()[s],
                           self.legs[l].offsets()[s] + self.legs[l].irreps[s][1] * (self.legs[l].irreps[s][0] + 1))
                     for l, s in enumerate(ir))

    def to_dense(self, flat):
        """Dense tensor on the embedded legs (index d * (2j + 1) + i inside an irrep)."""
        flat = np.asarray(flat)
        out = np.zeros(self.dims, dtype=np.complex128)
        n = len(self.legs)
        for b, (ir, tr) in enumerate(self.blocks):
            tjs = self.spins(b)
            o = int(self.offsets[b])
            T = flat[o:o + self.sizes[b]].reshape(self.block_dims[b], order="F")
            X = tree_tensor(tjs, tr, self.duals)
            full = np.einsum(T, list(range(0, 2 * n, 2)), X, list(range(1, 2 * n, 2)),
                             list(range(2 * n)))
            out[self._slices(ir)] += full.reshape([d * (tj + 1) for d, tj in zip(self.block_dims[b], tjs)])
        return out

    def from_dense(self, arr):
        """Degeneracy blocks of the projection of a dense tensor onto this
        structure's invariant subspace (exact for an invariant tensor)."""
        arr = np.asarray(arr, dtype=np.complex128)
        if arr.shape != self.dims:
            raise ValueError(f"from_dense: expected shape {self.dims}, got {arr.shape}")
        n = len(self.legs)
        flat = np.zeros(self.total, dtype=np.complex128)
        for b, (ir, tr) in enumerate(self.blocks):
            tjs = self.spins(b)
            shape = []
            for d, tj in zip(self.block_dims[b], tjs):
                shape += [d, tj + 1]
            sub = np.ascontiguousarray(arr[self._slices(ir)]).reshape(shape)
            X = tree_tensor(tjs, tr, self.duals)
            T = np.einsum(sub, list(range(2 * n)), X, list(range(1, 2 * n, 2)),
                          list(range(0, 2 * n, 2)))
            o = int(self.offsets[b])
            flat[o:o + self.sizes[b]] = np.asarray(T).reshape(-1, order="F")
        return flat

    def dense_from_flat(self, host):
        return self.to_dense(host)

    def __repr__(self):
        return f"SU2Structure({len(self.legs)} legs, {len(self.blocks)} blocks, {self.total} elements)"


def is_su2(structure):
    return isinstance(structure, SU2Structure)


def random_flat(rng, structure):
    """Random complex degeneracy blocks (unit-variance entries)."""
    n = structure.total
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / math.sqrt(2.0)


def to_device(structure, flat, ncod=None):
    """A DTensor holding SU(2) degeneracy blocks (dims: the dense leg dims)."""
    import torch

    from .tensors import DTensor, _device
    data = torch.from_numpy(np.ascontiguousarray(np.asarray(flat, dtype=np.complex128))).to(_device())
    return DTensor(structure.dims, data, ncod, structure)


# ---------------------------------------------------------------------------
# Coupled pairwise steps
# ---------------------------------------------------------------------------

_COEF_CACHE = {}
_COEF_CACHE_MAX = 1 << 16
COEF_TOL = 1e-13


def _pair_coefficients(tjsA, dualsA, tjsB, dualsB, oA, cA, oB, cB, perm, dualsC):
    """c[ta, tb, tc] for every tree of A's irreps (ta), of B's (tb) and of the
    output's (tc): <X_tc, contract(X_ta, X_tb)>.  Cached on the spins,
    directions and step geometry."""
    key = (tjsA, dualsA, tjsB, dualsB, oA, cA, oB, cB, perm, dualsC)
    hit = _COEF_CACHE.get(key)
    if hit is not None:
        return hit
    treesA, MA = tree_basis(tjsA, dualsA)
    treesB, MB = tree_basis(tjsB, dualsB)
    nA, nB = len(treesA), len(treesB)
    inter = tuple(tjsA[i] for i in oA) + tuple(tjsB[i] for i in oB)
    tjsC = tuple(inter[p] for p in perm)
    treesC, MC = tree_basis(tjsC, dualsC)
    if not (nA and nB and treesC):
        hit = _COEF_CACHE[key] = (treesA, treesB, tjsC, treesC, np.zeros((nA, nB, len(treesC))))
        return hit
    shapeA = tuple(t + 1 for t in tjsA)
    shapeB = tuple(t + 1 for t in tjsB)
    XA = MA.reshape((nA,) + shapeA).transpose((0,) + tuple(1 + i for i in oA + cA))
    XB = MB.reshape((nB,) + shapeB).transpose((0,) + tuple(1 + i for i in cB + oB))
    kdim = int(np.prod([tjsA[i] + 1 for i in cA])) if cA else 1
    ma = int(np.prod([tjsA[i] + 1 for i in oA])) if oA else 1
    nb = int(np.prod([tjsB[i] + 1 for i in oB])) if oB else 1
    Y = np.einsum("aik,bkj->abij", XA.reshape(nA, ma, kdim), XB.reshape(nB, kdim, nb))
    Y = Y.reshape((nA, nB) + tuple(t + 1 for t in inter)).transpose((0, 1) + tuple(2 + p for p in perm))
    coef = Y.reshape(nA, nB, -1) @ MC.T            # [nA, nB, nC]
    coef[np.abs(coef) < COEF_TOL] = 0.0
    if len(_COEF_CACHE) >= _COEF_CACHE_MAX:
        _COEF_CACHE.clear()
    hit = _COEF_CACHE[key] = (treesA, treesB, tjsC, treesC, coef)
    return hit


def compile_pair(SA, legsA, SB, legsB, oA, cA, oB, cB, perm, fixed_out=None):
    """One pairwise step: the result's structure and its block pairs.

    ``legsA`` / ``legsB`` are the operands' legs as the step sees them (dual
    flipped where the operand is conjugated); ``SA`` / ``SB`` their stored
    structures (the structural tensors are real, so conjugation changes the
    legs' directions and not the coefficients).  Returns (C structure,
    int32 array of (C block, A block, B block) triples grouped by C block,
    float64 coefficients)."""
    oA, cA, oB, cB, perm = tuple(oA), tuple(cA), tuple(oB), tuple(cB), tuple(perm)
    for i, j in zip(cA, cB):
        if legsA[i] != legsB[j].dual():
            raise ValueError(f"coupled step: contracted legs {legsA[i]} and {legsB[j]} are not dual")
    inter = [legsA[i] for i in oA] + [legsB[i] for i in oB]
    out_legs = [inter[p] for p in perm]
    dualsC = tuple(l.isdual for l in out_legs)
    if fixed_out is not None and list(fixed_out.legs) != out_legs:
        raise ValueError("coupled step: the given output structure has other legs")
    groupsB = {}
    for bi, (ir, _) in enumerate(SB.blocks):
        groupsB.setdefault(tuple(ir[j] for j in cB), {}).setdefault(ir, []).append(bi)
    groupsA = {}
    for ai, (ir, _) in enumerate(SA.blocks):
        groupsA.setdefault(ir, []).append(ai)
    dA, dB = SA.duals, SB.duals
    chunks = []          # per irrep pattern pair: (C irreps, C trees, tc, A block, B block, coefficient)
    for irA, alist in groupsA.items():
        ck = tuple(irA[i] for i in cA)
        gB = groupsB.get(ck)
        if not gB:
            continue
        tjsA = tuple(SA.legs[l].irreps[s][0] for l, s in enumerate(irA))
        posA = {SA.blocks[ai][1]: ai for ai in alist}
        for irB, blist in gB.items():
            tjsB = tuple(SB.legs[l].irreps[s][0] for l, s in enumerate(irB))
            treesA, treesB, tjsC, treesC, coef = _pair_coefficients(
                tjsA, dA, tjsB, dB, oA, cA, oB, cB, perm, dualsC)
            if not treesC:
                continue
            x, y, z = np.nonzero(coef)
            if not len(x):
                continue
            posB = {SB.blocks[bi][1]: bi for bi in blist}
            pa = np.asarray([posA.get(t, -1) for t in treesA], dtype=np.int64)[x]
            pb = np.asarray([posB.get(t, -1) for t in treesB], dtype=np.int64)[y]
            ok = (pa >= 0) & (pb >= 0)
            if not ok.all():
                x, y, z, pa, pb = x[ok], y[ok], z[ok], pa[ok], pb[ok]
            interir = tuple(irA[i] for i in oA) + tuple(irB[i] for i in oB)
            chunks.append((tuple(interir[p] for p in perm), treesC, z, pa, pb, coef[x, y, z]))
    if fixed_out is not None:
        C = fixed_out
    else:
        keys = sorted({(irC, treesC[int(t)]) for irC, treesC, z, _, _, _ in chunks
                       for t in np.unique(z)})
        if not keys:
            raise ValueError("coupled step: no block pair contributes")
        C = SU2Structure(out_legs, keys)
    cidx = C.index
    parts, cfs = [], []
    for irC, treesC, z, pa, pb, cf in chunks:
        ci = np.asarray([cidx.get((irC, t), -1) for t in treesC], dtype=np.int64)[z]
        if (ci < 0).any():
            bad = treesC[int(z[np.argmax(ci < 0)])]
            raise ValueError(f"coupled step: block {(irC, bad)} lies outside the given output structure")
        parts.append(np.stack([ci, pa, pb], axis=1))
        cfs.append(cf)
    if not parts:
        return C, np.zeros((0, 3), np.int32), np.zeros(0)
    trip = np.concatenate(parts).astype(np.int32)
    cf = np.concatenate(cfs).astype(np.float64)
    order = np.argsort(trip[:, 0], kind="stable")
    return C, np.ascontiguousarray(trip[order]), np.ascontiguousarray(cf[order])


def host_execute_pair(C, trip, coef, SA, flatA, conjA, SB, flatB, conjB, oA, cA, oB, cB, perm):
    """The numpy form of one coupled step (tests and small host checks)."""
    out = np.zeros(C.total, dtype=np.complex128)
    nOA = len(oA)
    for (ci, ai, bi), c in zip(trip, coef):
        a = flatA[SA.offsets[ai]:SA.offsets[ai] + SA.sizes[ai]].reshape(SA.block_dims[ai], order="F")
        b = flatB[SB.offsets[bi]:SB.offsets[bi] + SB.sizes[bi]].reshape(SB.block_dims[bi], order="F")
        a = np.conj(a) if conjA else a
        b = np.conj(b) if conjB else b
        r = np.tensordot(a.transpose(tuple(oA) + tuple(cA)), b.transpose(tuple(cB) + tuple(oB)),
                         axes=(list(range(nOA, nOA + len(cA))), list(range(len(cB)))))
        r = np.asarray(r).transpose(tuple(perm))
        o = int(C.offsets[ci])
        out[o:o + C.sizes[ci]] += c * r.reshape(-1, order="F")
    return out


# ---------------------------------------------------------------------------
# Device plans
# ---------------------------------------------------------------------------


class CoupledPlan:
    """A ContractionPlan over SU(2) tensors compiled into libgraft_heff.so
    (graft_plan_create_coupled): every step one grouped GEMM launch over its
    block pairs.  Quacks like contractions.NativePlan for EnvCache and
    EffectiveMap."""

    def __init__(self, plan, operands, out_structure=None, out_dims=None):
        self.plan = plan
        self.ninputs = plan.ninputs
        structs = [op.blocks for op in operands]
        if not all(is_su2(s) for s in structs):
            raise ValueError("CoupledPlan: every operand must be an SU(2) block tensor")
        slot_struct = list(structs) + [None] * len(plan.steps)
        self.step_structs, self.step_pairs, self.step_coefs = [], [], []
        for s in plan.steps:
            SA, SB = slot_struct[s.a], slot_struct[s.b]
            legsA = [l.dual() for l in SA.legs] if s.conjA else list(SA.legs)
            legsB = [l.dual() for l in SB.legs] if s.conjB else list(SB.legs)
            perm = tuple(s.out[0]) + tuple(s.out[1])
            fixed = out_structure if s.dst == plan.output_slot else None
            C, trip, cf = compile_pair(SA, legsA, SB, legsB, s.oindA, s.cindA, s.oindB, s.cindB,
                                       perm, fixed)
            slot_struct[s.dst] = C
            self.step_structs.append(C)
            self.step_pairs.append(trip)
            self.step_coefs.append(cf)
        self.out_structure = slot_struct[plan.output_slot]
        self.out_dims = out_dims if out_dims is not None else self.out_structure.dims
        ctx = nat.default_context()
        descs = [st.desc() for st in structs]
        arr = (nat.TensorDescC * len(descs))(*[d.c for d in descs])
        odescs = [C.desc() for C in self.step_structs]
        oarr = (nat.TensorDescC * len(odescs))(*[d.c for d in odescs])
        steps = (nat.PairStepC * len(plan.steps))(*[nat.make_step(s) for s in plan.steps])
        npairs = np.asarray([len(c) for c in self.step_coefs], dtype=np.int64)
        trip = np.ascontiguousarray(np.concatenate(self.step_pairs).reshape(-1).astype(np.int32)) \
            if npairs.sum() else np.zeros(3, np.int32)
        coefs = np.ascontiguousarray(np.concatenate(self.step_coefs)) if npairs.sum() \
            else np.zeros(1)
        h = ctypes.c_void_p()
        nat.check(nat.lib().graft_plan_create_coupled(
            ctx.handle, steps, len(plan.steps), plan.nslots, plan.output_slot, arr, len(descs),
            oarr, npairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
            trip.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            coefs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(h)))
        self.handle = h
        ws, macs, outn, db = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        nat.check(nat.lib().graft_plan_info(h, ctypes.byref(ws), ctypes.byref(macs), ctypes.byref(outn)))
        nat.check(nat.lib().graft_plan_device_bytes(h, ctypes.byref(db)))
        self.workspace_bytes = ws.value
        self.macs = macs.value
        self.output_elements = outn.value
        self.device_bytes = db.value
        self.npairs = int(npairs.sum())

    # the rest of NativePlan's interface
    from .contractions import NativePlan as _NP
    step_info = _NP.step_info
    step_gemms = _NP.step_gemms
    execute = _NP.execute
    execute_timed = _NP.execute_timed
    set_tile_config = _NP.set_tile_config
    set_orientation = _NP.set_orientation
    close = _NP.close
    __del__ = _NP.__del__
    del _NP

    def autotune(self, operands, out, precision=nat.PREC_FP64, repeats=2):
        from .contractions import NativePlan
        NativePlan.autotune(self, operands, out, precision, repeats)


def root_cap(legs):
    """Cap closing the root's parent legs (bra, operator channel, ket): the
    identity between the bra and ket legs on the operator's trivial channel.
    With a trivial root leg this is the ones tensor of the reference's
    _root_cap! (envcache.jl:626-655); a root leg carrying the target total
    spin S is closed by the identity on its multiplet, so every map acts as
    H on each of the 2S + 1 states."""
    st = SU2Structure.invariant(legs)
    dims = st.dims
    if len(dims) != 3 or dims[1] != 1 or dims[0] != dims[2]:
        raise ValueError("SU(2) root cap: expected (S, trivial, S*) root legs")
    dense = np.zeros(dims, dtype=np.complex128)
    for i in range(dims[0]):
        dense[i, 0, i] = 1.0
    flat = st.from_dense(dense)
    if np.abs(st.to_dense(flat) - dense).max() > 1e-12:
        raise ValueError("SU(2) root cap: the root legs do not close invariantly")
    return to_device(st, flat, len(legs))


# ---------------------------------------------------------------------------
# Single contractions, gauge and model builders
# ---------------------------------------------------------------------------

_PAIR_PLANS = {}
_PAIR_PLANS_MAX = 64


class _Step:
    """The fields of one planner PairStep (nat.make_step reads these)."""

    def __init__(self, a, b, dst, oindA, cindA, oindB, cindB, conjA, conjB, out):
        self.a, self.b, self.dst = a, b, dst
        self.oindA, self.cindA, self.oindB, self.cindB = oindA, cindA, oindB, cindB
        self.conjA, self.conjB, self.out = conjA, conjB, out


class _OnePlan:
    def __init__(self, step):
        self.steps = [step]
        self.nslots = 3
        self.output_slot = 2
        self.ninputs = 2


def contract(A, cA, B, cB, conjA=False, conjB=False, perm=None, ncod=None, out_structure=None):
    """contract_pair (Backend.jl:166-170) on SU(2) device tensors: A's legs
    ``cA`` with B's legs ``cB``; the result's legs are A's open legs then B's
    (in order), permuted by ``perm``.  Plans are cached per structure."""
    cA, cB = tuple(cA), tuple(cB)
    oA = tuple(i for i in range(A.rank) if i not in cA)
    oB = tuple(i for i in range(B.rank) if i not in cB)
    nout = len(oA) + len(oB)
    perm = tuple(range(nout)) if perm is None else tuple(perm)
    ncod = nout if ncod is None else int(ncod)
    key = (A.blocks.key(), B.blocks.key(), cA, cB, bool(conjA), bool(conjB), perm, ncod,
           None if out_structure is None else out_structure.key())
    p = _PAIR_PLANS.get(key)
    if p is None:
        step = _Step(0, 1, 2, oA, cA, oB, cB, bool(conjA), bool(conjB), (perm[:ncod], perm[ncod:]))
        p = CoupledPlan(_OnePlan(step), [A, B], out_structure)
        if len(_PAIR_PLANS) >= _PAIR_PLANS_MAX:
            _PAIR_PLANS.pop(next(iter(_PAIR_PLANS)))
        _PAIR_PLANS[key] = p
    from .tensors import DTensor
    out = DTensor.empty(p.out_structure.dims, ncod, p.out_structure)
    p.execute([A, B], out)
    return out


def fuse_multiplicities(m1, m2, maxspin=None):
    """{spin: multiplicity} of the product of two {spin: multiplicity} spaces."""
    out = {}
    for a, x in m1.items():
        for b, y in m2.items():
            for k in fusion(a, b):
                if maxspin is None or k <= maxspin:
                    out[k] = out.get(k, 0) + x * y
    return out


def bond_legs(topo, tphys=1, maxdeg=None, maxspin=None):
    """Bond leg (the child's space, not dual) of every edge child -> parent:
    the irreps of the child's subtree (children bonds (x) physical spin),
    each degeneracy capped at ``maxdeg`` and spins above ``maxspin`` dropped."""
    from .trees import nnodes, postorder
    legs = [None] * nnodes(topo)
    for n in postorder(topo):
        m = {tphys: 1}
        for c in topo.children[n]:
            m = fuse_multiplicities(m, dict(legs[c].irreps), maxspin)
        if maxdeg is not None:
            m = {k: min(v, maxdeg) for k, v in m.items()}
        legs[n] = SU2Leg(sorted(m.items()))
    return legs


def site_legs(topo, n, bonds, tphys=1, target=0):
    """Legs of TTNS tensor n: children bonds, physical spin, parent bond
    (dual); the root's parent leg carries the target total spin (twice it)."""
    legs = [bonds[c] for c in topo.children[n]] + [SU2Leg([(tphys, 1)])]
    parent = bonds[n] if topo.parent[n] >= 0 else SU2Leg([(target, 1)])
    return legs + [parent.dual()]


def left_orthonormalize(structure, flat):
    """Make a site tensor an isometry from its parent leg (the last leg, dual)
    onto the others: per parent irrep J, the degeneracy blocks stacked over
    (irrep tuple, tree, degeneracies) get orthonormal columns scaled by
    sqrt(2J + 1), the norm the structural tensors give each column (their
    partial trace over every other leg is 1 / (2J + 1) on the parent's
    multiplet).  The thin QR of the stack (TensorKit's per-sector
    leftorth, ttns.jl:181-203) keeps the tensor's span."""
    flat = np.array(flat, dtype=np.complex128)
    last = len(structure.legs) - 1
    byJ = {}
    for b, (ir, _) in enumerate(structure.blocks):
        byJ.setdefault(ir[last], []).append(b)
    for s, blist in byJ.items():
        tJ, dJ = structure.legs[last].irreps[s]
        rows = [structure.sizes[b] // dJ for b in blist]
        M = np.concatenate([flat[structure.offsets[b]:structure.offsets[b] + structure.sizes[b]]
                            .reshape(rows[i], dJ, order="F") for i, b in enumerate(blist)])
        if M.shape[0] < dJ:
            raise ValueError("left_orthonormalize: a parent irrep has more degeneracy than its span")
        Q, R = np.linalg.qr(M)
        Q = Q * np.sign(np.real(np.diag(R)) + (np.real(np.diag(R)) == 0))[None, :]
        Q *= math.sqrt(tJ + 1)
        o = 0
        for i, b in enumerate(blist):
            flat[structure.offsets[b]:structure.offsets[b] + structure.sizes[b]] = \
                Q[o:o + rows[i]].reshape(-1, order="F")
            o += rows[i]
    return flat


def random_ttns(rng, topo, tphys=1, maxdeg=None, maxspin=None, center=None, target=0,
                root_phys=True):
    """Random SU(2) tree state of total spin target / 2: every site spin
    tphys/2 (the root without a site when ``root_phys`` is false), bonds from
    ``bond_legs``; every tensor but the centre's (the root by default) an
    isometry towards its parent, the centre normalized."""
    from .networks import TTNS
    from .trees import nnodes
    bonds = bond_legs(topo, tphys, maxdeg, maxspin)
    root = next(i for i in range(nnodes(topo)) if topo.parent[i] < 0)
    center = root if center is None else center
    if center != root:
        raise ValueError("random_ttns: the centre must be the root")
    tensors = []
    hasphys = [root_phys or n != root for n in range(nnodes(topo))]
    for n in range(nnodes(topo)):
        legs = site_legs(topo, n, bonds, tphys, target)
        if not hasphys[n]:
            del legs[len(topo.children[n])]
        st = SU2Structure.invariant(legs)
        if st.total == 0:
            raise ValueError(f"random_ttns: node {topo.ids[n]} admits no invariant block")
        flat = random_flat(rng, st)
        flat = flat / np.linalg.norm(flat) if n == center else left_orthonormalize(st, flat)
        tensors.append(to_device(st, flat, len(legs) - 1))
    return TTNS(topo, tensors, center, hasphys)


CHANNEL = SU2Leg([(0, 2), (2, 1)])     # idle, done (scalars), one open S vector (spin 1)


def heisenberg_dense_site(K, isroot, J=1.0, tphys=1, phys=True):
    """Dense TTNO tensor of J sum_edges S_i . S_j (site spin tphys / 2) at a
    node with K children, channels in CHANNEL's embedded basis: 0 idle,
    1 done, 2..4 the spherical components q = +1, 0, -1 of an open term; legs
    (children channels..., P_out, P_in, parent channel).  The root's parent
    channel is 'done' only.  The node emits T_q (its spin's spherical
    components: T_+1 = -S+/sqrt2, T_0 = Sz, T_-1 = S-/sqrt2) towards the
    parent and closes a child's T_q with T_q^dagger:
    S_i . S_j = sum_q T_q(i) T_q(j)^dagger.  A root without a physical leg
    (``phys`` false; legs: children channels..., parent channel) couples the
    open terms of every pair of its children, S_k . S_l = sum_q (-1)^q
    T_q(k) T_-q(l) (T_q^dagger = (-1)^q T_-q), and passes a child's 'done'."""
    if not phys:
        if not isroot:
            raise ValueError("heisenberg_dense_site: only the root may lack a physical leg")
        W = np.zeros((5,) * K + (1,), dtype=np.complex128)
        for k in range(K):
            ch = [0] * K
            ch[k] = 1
            W[tuple(ch) + (0,)] = 1.0
            for l in range(k + 1, K):
                for i, q in enumerate((1, 0, -1)):
                    ch = [0] * K
                    ch[k], ch[l] = 2 + i, 2 + (2 - i)
                    W[tuple(ch) + (0,)] = J * (-1) ** abs(q)
        return W
    jx, jy, jz = spin_matrices(tphys)
    sp, sm = jx + 1j * jy, jx - 1j * jy
    T = {1: -sp / math.sqrt(2.0), 0: jz.astype(np.complex128), -1: sm / math.sqrt(2.0)}
    d = tphys + 1
    I = np.eye(d, dtype=np.complex128)
    pdim = 1 if isroot else 5
    W = np.zeros((5,) * K + (d, d, pdim), dtype=np.complex128)
    idle = (0,) * K
    done = 0 if isroot else 1
    if not isroot:
        W[idle + (slice(None), slice(None), 0)] = I
        for i, q in enumerate((1, 0, -1)):
            W[idle + (slice(None), slice(None), 2 + i)] = T[q]
    for k in range(K):
        ch = list(idle)
        ch[k] = 1
        W[tuple(ch) + (slice(None), slice(None), done)] += I
        for i, q in enumerate((1, 0, -1)):
            ch[k] = 2 + i
            W[tuple(ch) + (slice(None), slice(None), done)] += J * T[q].conj().T
    return W


def heisenberg_ttno(topo, J=1.0, tphys=1, root_phys=True):
    """SU(2) TTNO of the spin-(tphys/2) Heisenberg model J sum_<ij> S_i . S_j on the
    tree's edges (the dense tensors of ``heisenberg_dense_site`` projected onto
    their invariant blocks; the projection is exact, which ``from_dense``
    followed by ``to_dense`` confirms).  With ``root_phys`` false the root has
    no site: its children's spins are coupled pairwise through it instead."""
    from .networks import TTNO
    from .trees import nnodes
    tensors = []
    for n in range(nnodes(topo)):
        K = len(topo.children[n])
        isroot = topo.parent[n] < 0
        phys = root_phys or not isroot
        site = [SU2Leg([(tphys, 1)]), SU2Leg([(tphys, 1)], True)] if phys else []
        legs = [CHANNEL] * K + site + [(SU2Leg([(0, 1)]) if isroot else CHANNEL).dual()]
        st = SU2Structure.invariant(legs)
        Wd = heisenberg_dense_site(K, isroot, J, tphys, phys)
        flat = st.from_dense(Wd)
        if np.abs(st.to_dense(flat) - Wd).max() > 1e-12:
            raise RuntimeError("Heisenberg TTNO tensor is not SU(2) invariant")
        tensors.append(to_device(st, flat, K + len(site) // 2))
    return TTNO(topo, tensors, ishermitian=True)


def random_ttno(rng, topo, tphys=1, channel=CHANNEL):
    """Random invariant TTNO tensors on the given channel leg (tests)."""
    from .networks import TTNO
    from .trees import nnodes
    tensors = []
    for n in range(nnodes(topo)):
        K = len(topo.children[n])
        isroot = topo.parent[n] < 0
        legs = [channel] * K + [SU2Leg([(tphys, 1)]), SU2Leg([(tphys, 1)], True),
                                (SU2Leg([(0, 1)]) if isroot else channel).dual()]
        st = SU2Structure.invariant(legs)
        tensors.append(to_device(st, random_flat(rng, st), K + 1))
    return TTNO(topo, tensors)


def two_site_tensor(psi, n, m):
    """Theta = A_n contracted with A_m over their edge (effective.jl:827-843),
    on every invariant block of its legs: the two-site map reaches blocks
    whose n-side coupled spin the bond between n and m does not carry."""
    from .trees import childslot
    A, Bm = psi.tensors[n], psi.tensors[m]
    k = childslot(psi.topo, m, n)
    legs = A.blocks.legs[:-1] + [l for j, l in enumerate(Bm.blocks.legs) if j != k]
    return contract(A, (A.rank - 1,), Bm, (k,), ncod=len(legs),
                    out_structure=SU2Structure.invariant(legs))


# ---------------------------------------------------------------------------
# Two-site split and gauge moves (per coupled sector)
# ---------------------------------------------------------------------------


def _scale_axis(structure, flat, axis, scales):
    """Multiply every block along ``axis`` (a degeneracy index of that leg's
    irrep s) by scales[s][index]."""
    out = np.array(flat, dtype=np.complex128)
    for b, (ir, _) in enumerate(structure.blocks):
        sc = scales.get(ir[axis])
        if sc is None:
            continue
        o = int(structure.offsets[b])
        blk = out[o:o + structure.sizes[b]].reshape(structure.block_dims[b], order="F")
        shape = [1] * blk.ndim
        shape[axis] = len(sc)
        blk *= np.asarray(sc).reshape(shape)
        out[o:o + structure.sizes[b]] = blk.reshape(-1, order="F")
    return out


def _cut_key(st, b, pn):
    """Block b of ``st`` seen across the cut after its first pn legs (whose
    fusion trees pass through the coupled spin J of those legs): (J, row key =
    (their irreps, their internal tree labels), column key, rows, columns)."""
    ir, tr = st.blocks[b]
    J = tr[pn - 2] if pn >= 2 else st.spins(b)[0]
    rk = (ir[:pn], tr[:max(pn - 2, 0)])
    ck = (ir[pn:], tr[pn - 1:] if pn >= 2 else tr)
    nr = int(np.prod(st.block_dims[b][:pn]))
    nc = int(np.prod(st.block_dims[b][pn:])) if len(ir) > pn else 1
    return J, rk, ck, nr, nc


def _cut_pieces(st, host, pn):
    """Every block of a tensor as its matrix across the cut after pn legs:
    {(J, row key, column key): matrix}, with the row and column sizes per J."""
    rows, cols, pieces = {}, {}, {}
    for b in range(len(st.blocks)):
        J, rk, ck, nr, nc = _cut_key(st, b, pn)
        rows.setdefault(J, {}).setdefault(rk, nr)
        cols.setdefault(J, {}).setdefault(ck, nc)
        o = int(st.offsets[b])
        pieces[(J, rk, ck)] = host[o:o + st.sizes[b]].reshape(nr, nc, order="F")
    return rows, cols, pieces


def split_two_site_(psi, theta, n, m, trunc_rank, center_on="n"):
    """split_two_site! (effective.jl:852-876) for SU(2) states.

    Theta's fusion trees pass through the coupled spin J of the n-side legs
    (its label after the first pn legs), so Theta is block diagonal in J:
    M_J rows = (n-side irreps, n-side tree, degeneracies), columns = the
    m-side.  Its dense embedding has the singular values sigma(M_J) /
    sqrt(2J + 1), each (2J + 1) times: the spectrum the truncation rules see
    (``trunc_rank`` of the descending dense spectrum); whole multiplets are
    kept, as many as fit.  A_n = sqrt(2J + 1) U_J per J (a dense isometry
    onto the new bond); A_m = conj(A_n) contracted with Theta over the
    n-side legs (the coupled contraction recouples it into A_m's own trees),
    with the bond at child slot k.  With the centre on n, A_n takes the dense
    singular values and A_m = conj(A_n) contracted with Theta's polar factor
    (every kept singular value one)."""
    from .trees import childslot
    t = psi.topo
    k = childslot(t, m, n)
    An, Am = psi.tensors[n], psi.tensors[m]
    pn = An.ncod
    st = theta.blocks
    host = theta.data.detach().cpu().numpy()
    rows, cols, pieces = _cut_pieces(st, host, pn)
    svd = {}
    spectrum = []
    for J in rows:
        rkeys, ckeys = list(rows[J]), list(cols[J])
        roff = np.concatenate([[0], np.cumsum([rows[J][r] for r in rkeys])])
        coff = np.concatenate([[0], np.cumsum([cols[J][c] for c in ckeys])])
        M = np.zeros((roff[-1], coff[-1]), dtype=np.complex128)
        for i, r in enumerate(rkeys):
            for j, c in enumerate(ckeys):
                p = pieces.get((J, r, c))
                if p is not None:
                    M[roff[i]:roff[i + 1], coff[j]:coff[j + 1]] = p
        U, s, Vh = np.linalg.svd(M, full_matrices=False)
        sd = s / math.sqrt(J + 1)
        svd[J] = (rkeys, roff, U, sd, ckeys, coff, Vh)
        spectrum += [(x, J) for x in sd]
    spectrum.sort(key=lambda v: -v[0])
    dense = np.repeat([v[0] for v in spectrum], [v[1] + 1 for v in spectrum]) if spectrum \
        else np.zeros(0)
    r = int(trunc_rank(dense))
    keep, used = {}, 0
    for x, J in spectrum:
        if used + J + 1 > r:
            break
        if x <= 0.0:
            break
        keep[J] = keep.get(J, 0) + 1
        used += J + 1
    if not keep:                         # keep the leading multiplet, whatever its size
        x, J = spectrum[0]
        keep[J] = 1
    kept_sq = sum(float(np.sum(svd[J][3][:d] ** 2)) * (J + 1) for J, d in keep.items())
    disc = math.sqrt(max(0.0, float(np.sum(dense ** 2)) - kept_sq))
    bond = SU2Leg(sorted(keep.items()))
    bidx = {tj: i for i, (tj, _) in enumerate(bond.irreps)}
    # A_n on every invariant block of (n-side legs, bond*)
    nlegs = list(An.blocks.legs[:pn]) + [bond.dual()]
    Sn = SU2Structure.invariant(nlegs)
    fn = np.zeros(Sn.total, dtype=np.complex128)
    for b, (ir, tr) in enumerate(Sn.blocks):
        J = bond.irreps[ir[-1]][0]
        rkeys, roff, U = svd[J][:3]
        rk = (ir[:pn], tr[:max(pn - 2, 0)])
        if rk not in rkeys:
            continue
        i = rkeys.index(rk)
        d = keep[J]
        blk = U[roff[i]:roff[i + 1], :d] * math.sqrt(J + 1)
        o = int(Sn.offsets[b])
        fn[o:o + Sn.sizes[b]] = blk.reshape(-1, order="F")
    An_iso = to_device(Sn, fn, pn)
    # A_m = conj(A_n) . Theta, the bond moved to child slot k
    rank_m = Am.rank
    perm = tuple(0 if j == k else 1 + (j if j < k else j - 1) for j in range(rank_m))
    mlegs = [bond] + list(theta.blocks.legs[pn:])
    mlegs = [mlegs[p] for p in perm]
    sig = {bidx[J]: svd[J][3][:d] for J, d in keep.items()}
    if center_on == "n":
        # the isometric side from Theta's polar factor sqrt(2J+1) U_J Vh_J
        # (every kept dense singular value set to one): conj(A_n) . Theta' =
        # Vh exactly, with no division by small singular values
        fp = np.zeros(st.total, dtype=np.complex128)
        for b in range(len(st.blocks)):
            J, rk, ck, _, _ = _cut_key(st, b, pn)
            d = keep.get(J)
            if not d:
                continue
            rkeys, roff, U, _, ckeys, coff, Vh = svd[J]
            i = rkeys.index(rk)
            j = ckeys.index(ck)
            blk = math.sqrt(J + 1) * (U[roff[i]:roff[i + 1], :d] @ Vh[:d, coff[j]:coff[j + 1]])
            o = int(st.offsets[b])
            fp[o:o + st.sizes[b]] = blk.reshape(-1, order="F")
        polar = to_device(st, fp, theta.ncod)
        psi.tensors[m] = contract(An_iso, tuple(range(pn)), polar, tuple(range(pn)), conjA=True,
                                  perm=perm, ncod=Am.ncod, out_structure=SU2Structure.invariant(mlegs))
        psi.tensors[n] = to_device(Sn, _scale_axis(Sn, fn, pn, sig), pn)
    else:
        Amn = contract(An_iso, tuple(range(pn)), theta, tuple(range(pn)), conjA=True, perm=perm,
                       ncod=Am.ncod, out_structure=SU2Structure.invariant(mlegs))
        psi.tensors[n] = An_iso
        psi.tensors[m] = Amn
    psi.center = n if center_on == "n" else m
    return psi, dict(discarded=disc, spectrum=dense, kept=bond.dim)


def move_center_edge_(psi, m, cache=None):
    """One gauge move of the centre to the neighbour m (ttns.jl:181-203): the
    two-site tensor of the edge split again with every nonzero multiplet kept
    and the centre on m (an SVD where the reference takes a QR; exact)."""
    t = psi.topo
    n = psi.center
    child, parent = (n, m) if t.parent[n] == m else (m, n)
    theta = two_site_tensor(psi, child, parent)

    def keep_nonzero(s):
        if len(s) == 0:
            return 0
        return int(np.sum(s > 1e-14 * s[0]))

    split_two_site_(psi, theta, child, parent, keep_nonzero, "n" if m == child else "m")
    if cache is not None:
        cache.invalidate_edge_(n, m)
    return psi


# ---------------------------------------------------------------------------
# Link splits of the one-site TDVP sweep
# ---------------------------------------------------------------------------


def _identity_block_scale(tj, duals):
    """X of the two-leg tree (j, j) is c * identity; returns c."""
    X = tree_tensor((tj, tj), (), duals)
    c = X[0, 0]
    if np.abs(X - c * np.eye(tj + 1)).max() > 1e-14:
        raise RuntimeError("two-leg fusion tree is not proportional to the identity")
    return c


def identity(leg):
    """The identity on ``leg`` as a two-leg tensor on (leg*, leg)."""
    st = SU2Structure.invariant([leg.dual(), leg])
    flat = np.zeros(st.total, dtype=np.complex128)
    for b, (ir, _) in enumerate(st.blocks):
        tj, d = leg.irreps[ir[1]]
        o = int(st.offsets[b])
        flat[o:o + st.sizes[b]] = (np.eye(d) / _identity_block_scale(tj, st.duals)).reshape(-1, order="F")
    return to_device(st, flat, 1)


def _move_leg(A, src, dst):
    """A with leg ``src`` moved to position ``dst`` (a contraction with the identity)."""
    if src == dst:
        return A
    r = A.rank
    I = identity(A.blocks.legs[src])
    order = [j for j in range(r) if j != src]
    order.insert(dst, None)
    # intermediate legs: A's legs but src in order, then the identity's leg
    perm = tuple(r - 1 if j is None else (j if j < src else j - 1) for j in order)
    legs = [A.blocks.legs[j] for j in range(r) if j != src] + [A.blocks.legs[src]]
    legs = [legs[p] for p in perm]
    return contract(A, (src,), I, (0,), perm=perm, ncod=A.ncod, out_structure=SU2Structure.invariant(legs))


def _qr_last(A):
    """A = Q . (R (x) 1) on its last leg: per irrep J of that leg the blocks
    stacked over (irrep tuple, tree, degeneracies) are QR-factored; Q takes
    sqrt(2J + 1) times the orthonormal factor (a dense isometry onto the new
    bond, which keeps the old leg's irreps and direction) and R the triangular
    factor over sqrt(2J + 1), so that densely A = Q (R (x) 1).  Returns (Q, R
    by irrep of the new bond, new bond leg)."""
    st = A.blocks
    host = A.data.detach().cpu().numpy()
    last = len(st.legs) - 1
    byJ = {}
    for b, (ir, _) in enumerate(st.blocks):
        byJ.setdefault(ir[last], []).append(b)
    Qs, Rs, newirr = {}, {}, []
    for s in sorted(byJ):
        tj, d = st.legs[last].irreps[s]
        blist = byJ[s]
        rows = [st.sizes[b] // d for b in blist]
        M = np.concatenate([host[st.offsets[b]:st.offsets[b] + st.sizes[b]].reshape(rows[i], d, order="F")
                            for i, b in enumerate(blist)])
        Q, R = np.linalg.qr(M)
        Qs[tj] = (blist, rows, Q * math.sqrt(tj + 1))
        Rs[tj] = R / math.sqrt(tj + 1)           # the blocks' sqrt(2J + 1) moved into Q
        newirr.append((tj, Q.shape[1]))
    old = st.legs[last]
    new = SU2Leg(newirr, old.isdual)
    legs = list(st.legs[:last]) + [new]
    Sq = SU2Structure.invariant(legs)
    fq = np.zeros(Sq.total, dtype=np.complex128)
    for b, (ir, tr) in enumerate(Sq.blocks):
        tj = new.irreps[ir[last]][0]
        blist, rows, Q = Qs[tj]
        key = (ir[:last], tr)
        src = [i for i, bb in enumerate(blist) if (st.blocks[bb][0][:last], st.blocks[bb][1]) == key]
        if not src:
            continue
        i = src[0]
        o = int(Sq.offsets[b])
        r0 = sum(rows[:i])
        fq[o:o + Sq.sizes[b]] = Q[r0:r0 + rows[i]].reshape(-1, order="F")
    return to_device(Sq, fq, A.ncod), Rs, new


def _link_tensor(structure, mats):
    """Two-leg tensor on ``structure`` whose block of irrep j is mats[j] / c_j
    (dense: mats[j] (x) identity on the multiplet)."""
    flat = np.zeros(structure.total, dtype=np.complex128)
    for b, (ir, _) in enumerate(structure.blocks):
        tj = structure.legs[0].irreps[ir[0]][0]
        m = mats.get(tj)
        if m is None:
            continue
        o = int(structure.offsets[b])
        blk = m / _identity_block_scale(tj, structure.duals)
        if blk.shape != structure.block_dims[b]:
            raise ValueError("link tensor block does not match the map's structure")
        flat[o:o + structure.sizes[b]] = blk.reshape(-1, order="F")
    return to_device(structure, flat, 1)


def _link_matrices(C):
    """The per-irrep matrices of a two-leg tensor (inverse of _link_tensor)."""
    st = C.blocks
    host = C.data.detach().cpu().numpy()
    out = {}
    for b, (ir, _) in enumerate(st.blocks):
        tj = st.legs[0].irreps[ir[0]][0]
        o = int(st.offsets[b])
        out[tj] = host[o:o + st.sizes[b]].reshape(st.block_dims[b], order="F") * \
            _identity_block_scale(tj, st.duals)
    return out


def _absorb_on_leg(P, k, mats, irreps=None):
    """Leg k of P multiplied by mats[j] on the degeneracy index of irrep j
    (new index first): P'[.., i, ..] = sum_d mats[j][i, d] P[.., d, ..].
    ``irreps`` (spin, degeneracy) of the new leg, when it carries irreps the
    old one lacks (their blocks are zero); by default those of ``mats``."""
    st = P.blocks
    host = P.data.detach().cpu().numpy()
    leg = st.legs[k]
    if irreps is None:
        irreps = [(tj, mats[tj].shape[0]) for tj, _ in leg.irreps if tj in mats]
    nleg = SU2Leg(irreps, leg.isdual)
    old_of = {tj: s for s, (tj, _) in enumerate(leg.irreps)}
    legs = list(st.legs)
    legs[k] = nleg
    Sn = SU2Structure.invariant(legs)
    fn = np.zeros(Sn.total, dtype=np.complex128)
    for b, (ir, tr) in enumerate(Sn.blocks):
        tj = nleg.irreps[ir[k]][0]
        old_s = old_of.get(tj)
        if old_s is None or tj not in mats:
            continue
        src = st.index.get((ir[:k] + (old_s,) + ir[k + 1:], tr))
        if src is None:
            continue
        o = int(st.offsets[src])
        blk = host[o:o + st.sizes[src]].reshape(st.block_dims[src], order="F")
        blk = np.moveaxis(np.tensordot(mats[tj], blk, axes=([1], [k])), 0, k)
        on = int(Sn.offsets[b])
        fn[on:on + Sn.sizes[b]] = blk.reshape(-1, order="F")
    return to_device(Sn, fn, P.ncod)


def _permute(T, perm):
    """T with its legs reordered (leg perm[i] becomes leg i): a contraction of
    leg 0 with the identity, which recouples the trees."""
    r = T.rank
    p = tuple(r - 1 if j == 0 else j - 1 for j in perm)
    legs = [T.blocks.legs[j] for j in perm]
    return contract(T, (0,), identity(T.blocks.legs[0]), (0,), perm=p, ncod=T.ncod,
                    out_structure=SU2Structure.invariant(legs))


def cbe_link_split(ev, psi, H, u, v, dz):
    """_cbe_split for SU(2) states: the predictor selector's exact sector-aware
    fallback for graded maps (cbe.jl:536-560, 389-399) with the state-preserving
    splice (cbe.jl:575-584).

    The edge's leg of the centre tensor A_u is moved last (Al).  The projected
    core Z = (1 - Q Q^H) H_eff2 Theta (1 - conj(P) P^T) is materialized: the
    two-site map applied to Theta = A_u A_v, the partner's projection as two
    coupled contractions, and the active side's per coupled spin J of A_u's
    external legs, where the core is block diagonal (rows = those legs' irreps
    and trees, the same row basis as Al's blocks).  Its dense spectrum is
    sigma(Z_J) / sqrt(2J + 1), each (2J + 1) times; whole multiplets are
    spawned in descending order while they fit the budget min(room, max_add,
    neigs) in dense dimensions and reach spawn_threshold * ||A_u|| / |dz|.
    Per J, U_J = [Qh_J, E_J] (Qh_J: the orthonormal QR factor of Al's rows,
    E_J the selected left vectors, projected against Qh_J again and
    orthonormalized), stored as sqrt(2J + 1) U_J (a dense isometry); R_J =
    [R_qr_J; 0] / sqrt(2J + 1), so densely A = U (R (x) 1) and the state is
    unchanged.  Spins new to the bond have no R block (zero weight).

    Returns (U in Al's leg order, {spin: R_J}, the new bond leg)."""
    from .contractions import eff_h2
    from .trees import childslot
    t = psi.topo
    A = psi.tensors[u]
    up = t.parent[u] == v
    k = A.rank - 1 if up else childslot(t, u, v)
    Al = A if up else _move_leg(A, k, A.rank - 1)
    pu = Al.rank - 1
    ast = Al.blocks
    arows, _, apieces = _cut_pieces(ast, Al.data.detach().cpu().numpy(), pu)
    old = ast.legs[-1]
    strat = ev.cbe
    qdim = sum((tj + 1) * min(sum(arows[tj].values()), d) for tj, d in old.irreps if tj in arows)
    room = max(0, ev.trunc.maxdim - qdim)
    budget = min(room, strat.max_add, strat.neigs)
    info = {"selected": 0, "budget": budget, "solver": "exact_svd_sector"}
    ev.last_cbe_info = info
    # per spin: Al's rows and the QR of its (rows x degeneracy) matrix
    qr = {}
    for s, (tj, d) in enumerate(old.irreps):
        if tj not in arows:
            continue
        rk = list(arows[tj])
        M = np.concatenate([apieces[(tj, r, ((s,), ()))] for r in rk])
        Qh, R = np.linalg.qr(M)
        qr[tj] = (rk, [arows[tj][r] for r in rk], Qh, R)
    zrows, zsvd = {}, {}
    if budget > 0 and dz != 0:
        n, m = (u, v) if up else (v, u)
        theta = two_site_tensor(psi, n, m)
        Y = eff_h2(ev.cache, psi, H, n, m, out_structure=theta.blocks)(theta)
        P = psi.tensors[v]
        kv = childslot(t, v, u) if up else P.rank - 1
        pv = P.rank - 1
        if not up:                                   # Theta's legs: v's first
            Y = _permute(Y, tuple(range(pv, pv + pu)) + tuple(range(pv)))
        pext = tuple(j for j in range(P.rank) if j != kv)
        T1 = contract(Y, tuple(range(pu, pu + pv)), P, pext, conjB=True, ncod=pu)
        T2 = contract(T1, (pu,), P, (kv,), ncod=Y.ncod, out_structure=Y.blocks)
        zh = (Y.data - T2.data).detach().cpu().numpy()
        zr, zc, zp = _cut_pieces(Y.blocks, zh, pu)
        cand = []
        for J in zr:
            # rows: Al's keys first (in QR order), then the core's others
            size = dict(zip(qr[J][0], qr[J][1])) if J in qr else {}
            size.update(zr[J])
            rk = list(qr[J][0]) if J in qr else []
            rk += [r for r in zr[J] if r not in rk]
            rsz = [size[r] for r in rk]
            ck = list(zc[J])
            roff = np.concatenate([[0], np.cumsum(rsz)]).astype(int)
            coff = np.concatenate([[0], np.cumsum([zc[J][c] for c in ck])]).astype(int)
            Z = np.zeros((roff[-1], coff[-1]), dtype=np.complex128)
            for i, r in enumerate(rk):
                for j, c in enumerate(ck):
                    pc = zp.get((J, r, c))
                    if pc is not None:
                        Z[roff[i]:roff[i + 1], coff[j]:coff[j + 1]] = pc
            if J in qr:                              # the active side's projection, per J
                Qh = np.zeros((roff[-1], qr[J][2].shape[1]), dtype=np.complex128)
                Qh[:qr[J][2].shape[0]] = qr[J][2]
                Z = Z - Qh @ (Qh.conj().T @ Z)
            U, sv, _ = np.linalg.svd(Z, full_matrices=False)
            zrows[J] = (rk, rsz)
            zsvd[J] = U
            cand += [(-float(x) / math.sqrt(J + 1), J, i) for i, x in enumerate(sv)]
        norm_a = float(np.linalg.norm(A.data.detach().cpu().numpy()))
        cutoff = strat.spawn_threshold * max(norm_a, np.finfo(float).eps) / abs(complex(dz))
        cand.sort()
        info["sigma_max"] = -cand[0][0] if cand else 0.0
        sel, used = {}, 0
        for negx, J, i in cand:
            if -negx < cutoff or used + J + 1 > budget:
                break
            sel.setdefault(J, []).append(i)
            used += J + 1
        info["selected"] = used
    else:
        sel = {}
    # U_J = [Qh_J, E_J] on the rows of both; R_J = [R_qr_J; 0]
    cols, Rs = {}, {}
    for J in sorted(set(qr) | set(sel)):
        rk, rsz = zrows[J] if J in zrows else (qr[J][0], qr[J][1])
        nrow = int(np.sum(rsz))
        Qh = np.zeros((nrow, qr[J][2].shape[1] if J in qr else 0), dtype=np.complex128)
        if J in qr:
            Qh[:qr[J][2].shape[0]] = qr[J][2]
        Uj = Qh
        if J in sel:
            E = zsvd[J][:, sel[J]]
            E = E - Qh @ (Qh.conj().T @ E)
            E, _ = np.linalg.qr(E)
            Uj = np.concatenate([Qh, E], axis=1)
        cols[J] = (rk, np.concatenate([[0], np.cumsum(rsz)]).astype(int), Uj)
        if J in qr:
            R = qr[J][3]
            Rs[J] = np.concatenate([R, np.zeros((Uj.shape[1] - R.shape[0], R.shape[1]))]) / math.sqrt(J + 1)
    U, new = _isometry_from_columns(ast.legs[:-1], old, cols, Al.ncod)
    return U, Rs, new


def _isometry_from_columns(ext, old, cols, ncod):
    """The tensor on (ext legs, new bond) whose blocks of bond spin J are
    sqrt(2J + 1) times the rows of cols[J] = (row keys, row offsets, U_J)
    (orthonormal columns: a dense isometry); the new bond has U_J's column
    count for every J and old's direction."""
    pu = len(ext)
    new = SU2Leg([(J, cols[J][2].shape[1]) for J in sorted(cols) if cols[J][2].shape[1] > 0],
                 old.isdual)
    Su = SU2Structure.invariant(list(ext) + [new])
    fu = np.zeros(Su.total, dtype=np.complex128)
    for b in range(len(Su.blocks)):
        J, rk, _, nr, _ = _cut_key(Su, b, pu)
        rkeys, roff, Uj = cols[J]
        if rk not in rkeys:
            continue
        i = rkeys.index(rk)
        o = int(Su.offsets[b])
        fu[o:o + Su.sizes[b]] = (math.sqrt(J + 1) * Uj[roff[i]:roff[i + 1]]).reshape(-1, order="F")
    return to_device(Su, fu, ncod), new


def null_enrich_split(A, maxdim):
    """_null_enrich_split (expansion.jl:560-575) for SU(2) states: A (bond
    last) == U (R (x) 1) with U a dense isometry of dense bond dimension up to
    ``maxdim``: per spin J of the external legs' fusion, the QR factor of A's
    rows completed by orthonormal directions of its complement (zero weight:
    the state is unchanged; spins the bond lacks enter with complement
    columns only).  Whole multiplets fill the room, smallest spins first (the
    reference's choice among these equal-weight directions is arbitrary as
    well).  Returns (U, {spin: R_J}, new bond leg), or (A, None, None) when
    there is no room."""
    st = A.blocks
    pu = A.rank - 1
    old = st.legs[-1]
    ext = st.legs[:-1]
    mult = {}
    for l in ext:
        mult = fuse_multiplicities(mult, dict(l.irreps)) if mult else dict(l.irreps)
    room = min(maxdim, int(np.prod(A.dims[:-1]))) - old.dim
    if room <= 0:
        return A, None, None
    # every row key of every spin the external legs couple to
    full = SU2Structure.invariant(list(ext) + [SU2Leg([(J, 1) for J in mult], old.isdual)])
    rows = {}
    for b in range(len(full.blocks)):
        J, rk, _, nr, _ = _cut_key(full, b, pu)
        rows.setdefault(J, {}).setdefault(rk, nr)
    _, _, pieces = _cut_pieces(st, A.data.detach().cpu().numpy(), pu)
    cols, Rs, left = {}, {}, room
    for J in sorted(rows):
        rk = list(rows[J])
        rsz = [rows[J][r] for r in rk]
        roff = np.concatenate([[0], np.cumsum(rsz)]).astype(int)
        s = [i for i, (tj, _) in enumerate(old.irreps) if tj == J]
        if s:
            d = old.irreps[s[0]][1]
            M = np.zeros((roff[-1], d), dtype=np.complex128)
            for i, r in enumerate(rk):
                pc = pieces.get((J, r, ((s[0],), ())))
                if pc is not None:
                    M[roff[i]:roff[i + 1]] = pc
            Qc, R = np.linalg.qr(M, mode="complete")
            q = min(roff[-1], d)
            Q, comp = Qc[:, :q], Qc[:, q:]
            Rs[J] = R[:q]
        else:
            Q, comp = np.zeros((roff[-1], 0), dtype=np.complex128), np.eye(roff[-1], dtype=np.complex128)
        e = min(comp.shape[1], left // (J + 1))
        left -= e * (J + 1)
        Uj = np.concatenate([Q, comp[:, :e]], axis=1)
        if Uj.shape[1]:
            cols[J] = (rk, roff, Uj)
        if J in Rs:
            Rs[J] = np.concatenate([Rs[J], np.zeros((e, Rs[J].shape[1]))]) / math.sqrt(J + 1)
    if left == room:
        return A, None, None
    U, new = _isometry_from_columns(ext, old, cols, A.ncod)
    return U, Rs, new


def evolve_link_and_move_(ev, psi, H, u, v, dz, herm, exponentiate, cbe=False):
    """The link step of the one-site TDVP sweep (tdvp.jl:257-301) on SU(2)
    states: the centre tensor is QR-split on the edge's leg per irrep (moved to
    the last position by a contraction with the identity when it is a child
    slot), the link tensor evolves backwards under the zero-site map and
    multiplies into the neighbour.  With bond expansion (TDVP1_CBE,
    cbe.jl:240-261) the split is cbe_link_split's: the neighbour's leg takes
    the new bond, zero on the spawned directions."""
    from .contractions import eff_h0
    from .trees import childslot
    t = psi.topo
    cache = ev.cache
    kw = dict(ishermitian=herm, krylovdim=ev.krylovdim, tol=ev.tol, eager=ev.eager)
    A = psi.tensors[u]
    if t.parent[u] == v:
        if cbe:                                            # A = U . (R (x) 1), U = [Q, E]
            Q, R, new = cbe_link_split(ev, psi, H, u, v, dz)
        else:
            Q, R, new = _qr_last(A)                        # A = Q . (R (x) 1) on the parent leg
        psi.tensors[u] = Q
        cache.invalidate_node_(u)
        k0 = eff_h0(cache, psi, H, u, v, precision=ev.precision)
        C = _link_tensor(k0.blocks, R)                     # (new, old*)
        C, _ = exponentiate(k0, -dz, C, **kw)
        psi.tensors[v] = _absorb_on_leg(psi.tensors[v], childslot(t, v, u), _link_matrices(C),
                                        new.irreps)
    else:
        k = childslot(t, u, v)
        if cbe:
            Ql, R, new = cbe_link_split(ev, psi, H, u, v, dz)
        else:
            Ql, R, new = _qr_last(_move_leg(A, k, A.rank - 1))
        psi.tensors[u] = _move_leg(Ql, Ql.rank - 1, k)
        cache.invalidate_node_(u)
        psi.center = u
        k0 = eff_h0(cache, psi, H, v, u, precision=ev.precision)
        C = _link_tensor(k0.blocks, {tj: r.T for tj, r in R.items()})     # (old, new*)
        C, _ = exponentiate(k0, -dz, C, **kw)
        P = psi.tensors[v]
        psi.tensors[v] = _absorb_on_leg(P, P.rank - 1, {tj: m.T for tj, m in _link_matrices(C).items()},
                                        new.irreps)
    psi.center = v
    cache.invalidate_node_(v)
    return psi
