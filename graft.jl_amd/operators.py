"""Closed-form nearest-neighbour TTNOs for the workloads of this engine.

The reference builds TTNOs from an OpSum through a state-diagram compiler
(lib/GraftTTNOBuild/src/TTNOBuild/statediagram.jl:713, out of scope here).
For Hamiltonians made of on-site terms and two-body terms on tree edges the
state diagram is the finite automaton below, so it is written down directly.
The virtual leg on edge (child c -> parent p) carries the state of the
subtree below c:
    channel 0  identity      (no operator placed below)
    channel 1  done          (every term below is complete)
    channel 2+a  open term a (operator A_a placed on c, completed at p by B_a)
Bond dimension w = 2 + number of two-body operator types: 3 for the
transverse-field Ising model (identity, done, Z) and 5 for Heisenberg
(identity, done, S+, S-, Z), as the reference's builder produces.
Leg order (ttno.jl:1-9): children channels..., P_out, P_in, parent channel.
"""
from __future__ import annotations


import numpy as np

from .networks import TTNO
from .tensors import DTensor
from .trees import edges, nchildren, nnodes, postorder

SX = np.array([[0, 1], [1, 0]], dtype=complex)
SZ = np.array([[1, 0], [0, -1]], dtype=complex)
SP = np.array([[0, 1], [0, 0]], dtype=complex)
SM = np.array([[0, 0], [1, 0]], dtype=complex)


def nn_ttno_host(topo, d, onsite, pair_ops, pair_coef):
    """Host arrays of the automaton TTNO.

    onsite[n]: d x d matrix (coefficient included) or None.
    pair_ops: list of (A_a, B_a); the term on edge (c, p) is
              pair_coef[c][a] * A_a(c) B_a(p).
    """
    nt = len(pair_ops)
    w = 2 + nt
    I = np.eye(d, dtype=complex)
    W = []
    for n in range(nnodes(topo)):
        K = nchildren(topo, n)
        root = topo.parent[n] < 0
        up_dim = 1 if root else w
        T = np.zeros([w] * K + [d, d, up_dim], dtype=complex)

        def put(children, up, op):
            if root:
                if up != 1:
                    return
                up = 0
            T[tuple(children) + (slice(None), slice(None), up)] += op

        idle = [0] * K
        put(idle, 0, I)                         # identity passes up
        if onsite[n] is not None:
            put(idle, 1, onsite[n])             # on-site term completes here
        for k, c in enumerate(topo.children[n]):
            ch = list(idle)
            ch[k] = 1
            put(ch, 1, I)                       # a completed subtree passes up
            for a, (_, Bop) in enumerate(pair_ops):
                ch = list(idle)
                ch[k] = 2 + a
                put(ch, 1, pair_coef[c][a] * Bop)   # complete the edge term (c, n)
        for a, (Aop, _) in enumerate(pair_ops):
            put(idle, 2 + a, Aop)               # open a term towards the parent
        W.append(T)
    return W


def terms_ttno_host(topo, phys, terms):
    """Host arrays of the automaton TTNO for a sum of one- and two-site terms on
    any sites of the tree; nodes with ``phys[n] == 0`` carry no physical leg
    (W :: children... <- parent, as ttno.jl:1-9 allows).

    ``terms``: list of (coef, [(site, op), ...]) with one or two distinct sites.
    The virtual leg on edge (c -> parent) carries: 0 identity, 1 done, and one
    channel per two-site term with exactly one site below c (its operator is
    placed, the partner is still to come).  A term completes, with its
    coefficient, at the lowest node whose subtree holds all its sites."""
    n = nnodes(topo)
    below = [set() for _ in range(n)]
    for v in postorder(topo):
        below[v].add(v)
        for c in topo.children[v]:
            below[v] |= below[c]
    for coef, ops in terms:
        sites = [s for s, _ in ops]
        if not 1 <= len(sites) <= 2 or len(set(sites)) != len(sites):
            raise ValueError("terms_ttno_host: every term acts on one or two distinct sites")
        if any(not phys[s] for s in sites):
            raise ValueError("terms_ttno_host: a term acts on a node without a physical leg")
    chan = {}
    for c in range(n):
        if topo.parent[c] < 0:
            continue
        ch = {"id": 0, "done": 1}
        for i, (_, ops) in enumerate(terms):
            if len(ops) == 2 and sum(s in below[c] for s, _ in ops) == 1:
                ch[i] = len(ch)
        chan[c] = ch
    W = []
    for v in range(n):
        kids = topo.children[v]
        K = len(kids)
        d = phys[v]
        root = topo.parent[v] < 0
        up_dim = 1 if root else len(chan[v])
        T = np.zeros([len(chan[c]) for c in kids] + ([d, d] if d else []) + [up_dim], dtype=complex)
        I = np.eye(d, dtype=complex) if d else None

        def put(children, up, op, coef=1.0):
            if root:
                if up != "done":
                    return
                u = 0
            else:
                u = chan[v][up]
            idx = tuple(children) + ((slice(None), slice(None)) if d else ()) + (u,)
            T[idx] += coef * (op if d else 1.0)

        idle = [0] * K
        put(idle, "id", I)                                  # identity passes up
        for k in range(K):
            ch = list(idle)
            ch[k] = 1
            put(ch, "done", I)                              # a completed subtree passes up
        for i, (coef, ops) in enumerate(terms):
            inside = [(s, o) for s, o in ops if s in below[v]]
            if len(inside) < len(ops):
                if len(inside) == 1 and not root:           # carry the open term upwards
                    s, o = inside[0]
                    if s == v:
                        put(idle, i, o)
                    else:
                        k = next(k for k, c in enumerate(kids) if s in below[c])
                        ch = list(idle)
                        ch[k] = chan[kids[k]][i]
                        put(ch, i, I)
                continue
            if any(all(s in below[c] for s, _ in ops) for c in kids):
                continue                                    # completes further down
            ch = list(idle)
            op = I
            for s, o in ops:
                if s == v:
                    op = o
                else:
                    k = next(k for k, c in enumerate(kids) if s in below[c])
                    ch[k] = chan[kids[k]][i]
            put(ch, "done", op, coef)
        W.append(T)
    return W


def _upload(topo, W, hermitian=True):
    tens = [DTensor.from_numpy(a, nchildren(topo, n) + (1 if a.ndim == nchildren(topo, n) + 3 else 0))
            for n, a in enumerate(W)]
    return TTNO(topo, tens, ishermitian=hermitian)


def terms_ttno(topo, phys, terms, hermitian=True):
    return _upload(topo, terms_ttno_host(topo, phys, terms), hermitian)


def tfi_host(topo, J=1.0, g=0.9):
    """-J sum_edges Z Z - g sum X (test/runtests.jl:19-29)."""
    n = nnodes(topo)
    coef = {c: [-J] for c, _ in edges(topo)}
    return nn_ttno_host(topo, 2, [-g * SX] * n, [(SZ, SZ)], coef)


def tfi_random_host(topo, Js, hs):
    """Edge couplings Js[child] and fields hs[node]."""
    coef = {c: [-Js[c]] for c, _ in edges(topo)}
    return nn_ttno_host(topo, 2, [-hs[i] * SX for i in range(nnodes(topo))], [(SZ, SZ)], coef)


def heisenberg_host(topo, J=1.0):
    """J/2 (S+S- + S-S+) + J/4 ZZ on every edge (test/runtests.jl:32-41)."""
    n = nnodes(topo)
    coef = {c: [J / 2, J / 2, J / 4] for c, _ in edges(topo)}
    return nn_ttno_host(topo, 2, [None] * n, [(SP, SM), (SM, SP), (SZ, SZ)], coef)


def tfi_ttno(topo, J=1.0, g=0.9):
    return _upload(topo, tfi_host(topo, J, g))


def tfi_random_ttno(topo, Js, hs):
    return _upload(topo, tfi_random_host(topo, Js, hs))


def heisenberg_ttno(topo, J=1.0):
    return _upload(topo, heisenberg_host(topo, J))


# ---------------------------------------------------------------------------
# Spinful Hubbard model on the tree graph with U(1) x U(1) (N, 2 S_z) sectors.
#
# Fermions are mapped to d = 4 sites (|0>, |up>, |dn>, |up dn>) with the
# Jordan-Wigner order = depth-first preorder of the tree (a parent precedes
# its subtree, children in leg order), up before down on a site.  For a hop
# between a parent i and its child j the string covers the subtrees of the
# children of i that precede j, which the automaton carries in a parity
# channel P = prod (-1)^n:
#     c^dag_i c_j = (c^dag_i P_i) (prod_{k in earlier sibling subtrees} P_k) c_j
# Channels: 0 identity, 1 done, 2 parity, 3.. open operators placed below
# (c_up, c_dn, c^dag_up, c^dag_dn), completed at the parent.
# ---------------------------------------------------------------------------

_A_UP = np.zeros((4, 4), dtype=complex)
_A_UP[0, 1] = 1.0
_A_UP[2, 3] = 1.0
_A_DN = np.zeros((4, 4), dtype=complex)
_A_DN[0, 2] = 1.0
_A_DN[1, 3] = 1.0
FUP = np.diag([1.0, -1.0, 1.0, -1.0]).astype(complex)
C_UP = _A_UP                       # annihilates the up fermion
C_DN = FUP @ _A_DN                 # intra-site string (-1)^{n_up}
CD_UP = C_UP.conj().T
CD_DN = C_DN.conj().T
PAR = np.diag([1.0, -1.0, -1.0, 1.0]).astype(complex)   # (-1)^(n_up + n_dn)
N_UP = CD_UP @ C_UP
N_DN = CD_DN @ C_DN
ID4 = np.eye(4, dtype=complex)

HUBBARD_OPEN = [C_UP, C_DN, CD_UP, CD_DN]
# channel charges (N, 2 S_z) of the open operators placed below
HUBBARD_OPEN_CHARGES = [(-1, -1), (-1, 1), (1, 1), (1, -1)]
HUBBARD_PHYS_CHARGES = [(0, 0), (1, 1), (1, -1), (2, 0)]


def hubbard_host(topo, t=1.0, U=4.0):
    """Dense TTNO arrays of H = -t sum_{tree edges, s} (c^dag c + h.c.) + U sum n_up n_dn."""
    w = 3 + len(HUBBARD_OPEN)
    W = []
    for n in range(nnodes(topo)):
        K = nchildren(topo, n)
        root = topo.parent[n] < 0
        T = np.zeros([w] * K + [4, 4, 1 if root else w], dtype=complex)

        def put(children, up, op):
            if root:
                if up != 1:
                    return
                up = 0
            T[tuple(children) + (slice(None), slice(None), up)] += op

        idle = [0] * K
        put(idle, 0, ID4)
        put([2] * K, 2, PAR)
        put(idle, 1, U * (N_UP @ N_DN))
        for k in range(K):
            ch = list(idle)
            ch[k] = 1
            put(ch, 1, ID4)
            for a, op in enumerate(HUBBARD_OPEN):
                ch = [2] * k + [3 + a] + [0] * (K - k - 1)
                if a < 2:      # child carries c_s: complete with -t c^dag_s P
                    put(ch, 1, -t * (HUBBARD_OPEN[a + 2] @ PAR))
                else:          # child carries c^dag_s: complete with -t P c_s
                    put(ch, 1, -t * (PAR @ HUBBARD_OPEN[a - 2]))
        for a, op in enumerate(HUBBARD_OPEN):
            put(idle, 3 + a, op)
        W.append(T)
    return W


def hubbard_channel_leg():
    """Channel space of the Hubbard TTNO as (charge, degeneracy) in channel order."""
    from .planning import SectorLeg
    return SectorLeg([((0, 0), 3)] + [(q, 1) for q in HUBBARD_OPEN_CHARGES])


def hubbard_phys_leg():
    from .planning import SectorLeg
    return SectorLeg([(q, 1) for q in HUBBARD_PHYS_CHARGES])


def blockify(arr, legs, ncod):
    """Split a dense array into the invariant blocks of ``legs`` (flat-leg charges).

    Raises if a non-zero entry lies outside every invariant block."""
    from .sectors import HostBlocks, invariant_blocks
    starts = [np.concatenate([[0], np.cumsum(l.degs())]) for l in legs]
    blocks = {}
    covered = np.zeros(arr.shape, dtype=bool)
    for key in invariant_blocks(legs):
        sl = tuple(slice(int(starts[l][s]), int(starts[l][s + 1])) for l, s in enumerate(key))
        blocks[key] = np.array(arr[sl])
        covered[sl] = True
    if np.abs(arr[~covered]).max(initial=0.0) > 0:
        raise ValueError("tensor has entries outside the charge-conserving blocks")
    return HostBlocks(legs, blocks, ncod)


def hubbard_sector_ttno(topo, t=1.0, U=4.0):
    """The Hubbard TTNO as U(1) x U(1) sector tensors on the device."""
    from .planning import SectorLeg
    W = hubbard_host(topo, t, U)
    ch = hubbard_channel_leg()
    ph = hubbard_phys_leg()
    tens = []
    for n, a in enumerate(W):
        K = nchildren(topo, n)
        up = SectorLeg([((0, 0), 1)]) if topo.parent[n] < 0 else ch
        legs = [ch] * K + [ph, ph.dual(), up.dual()]
        tens.append(blockify(a, legs, K + 1).to_device())
    return TTNO(topo, tens, ishermitian=True)
