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

import itertools

import numpy as np

from .networks import TTNO
from .tensors import DTensor
from .trees import edges, nchildren, nnodes

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


def _upload(topo, W, hermitian=True):
    tens = [DTensor.from_numpy(a, nchildren(topo, n) + 1) for n, a in enumerate(W)]
    return TTNO(topo, tens, ishermitian=hermitian)


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
