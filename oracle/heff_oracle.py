"""CPU oracle for the effective-Hamiltonian (H_eff) path — TEST INFRASTRUCTURE ONLY.

This module is a plain numpy restatement of the reference algorithm (Graft.jl)
for the one-, two- and zero-site effective Hamiltonians, the environment
update, the dense test recipes that pin them, and the pairwise cost model.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the timed CPU
baseline.  The product path (``graft.jl_amd``) never imports it.

Arrays hold one TensorKit ``TensorMap`` with trivial sectors each: a numpy
array whose axes are the flat legs (codomain legs, then domain legs) in the
reference order.  Flattening to a device buffer uses Fortran (first leg
fastest) order, which is TensorKit's block layout for ``ComplexSpace``.

How each function is pinned (so that a plausible mistake fails a check):

* ``dense_hamiltonian`` — closed forms: the open transverse-field Ising chain
  equals its free-fermion spectrum; two-site Heisenberg singlet/triplet
  energies; Kronecker ordering checked against single-site expectation values.
* ``ttno_to_dense`` — equal to ``dense_hamiltonian`` for the TTNOs the
  product builds (any wrong channel, coefficient or leg order fails).
* ``ttns_to_dense`` / ``canonicalize`` — norm and gauge invariance, isometry
  of every non-centre tensor, overlap with a brute-force reconstruction.
* ``build_env`` / ``env_closure`` — the full closure equals the dense
  expectation value <psi|H|psi>.
* ``h1_apply`` / ``h2_apply`` / ``h0_apply`` — the tangent-space identity
  y = P^dagger H_dense P x with P the isometric embedding of the site (link)
  tensor into the full Hilbert space, built from ``ttns_to_dense`` and
  ``dense_hamiltonian`` only; hermiticity <a, H b> = <H a, b>.
* ``dense_cost`` / ``u1_chain_plan_costs`` — the reference's known answers
  (flops 60 / 63, sector_flops 30 / 28 in
  lib/GraftPlanning/test/runtests.jl:167-204).
* ``exact_groundstate`` / ``exact_evolve`` — eigendecomposition identities.
* ``execute_pair_steps`` — equals ``ncon`` on the same network.
* ``tdvp1_step`` — with every bond at full rank it equals exact propagation
  exp(dz H) (each site step and the link step after it act on one subspace
  and cancel, the last site step is the whole propagator); with truncated
  bonds it conserves the norm and the energy, and the order-2 step is
  time-reversible (a step of -dz undoes a step of dz).

Nothing in this module is "parity unpinned".
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------
# Trees — lib/GraftFoundation/src/Trees/topology.jl:22-122, geometries.jl:51-92
# Node indices are 0-based; parent -1 marks the root (index 0).
# ---------------------------------------------------------------------------


@dataclass
class Tree:
    ids: list
    parent: list
    children: list = field(default_factory=list)

    @property
    def n(self):
        return len(self.parent)


def tree_from_edges(root, edges):
    """TreeTopology(root, edges) — topology.jl:30-47 (edges given top-down)."""
    ids = [root]
    index = {root: 0}
    parent = [-1]
    children = [[]]
    for p, c in edges:
        ids.append(c)
        index[c] = len(ids) - 1
        parent.append(index[p])
        children.append([])
        children[index[p]].append(index[c])
    return Tree(ids, parent, children)


def star_tree(nbranches, branchlength):
    """star_topology — geometries.jl:64-75."""
    edges = []
    for b in range(1, nbranches + 1):
        prev = "center"
        for l in range(1, branchlength + 1):
            node = f"b{b}_{l}"
            edges.append((prev, node))
            prev = node
    return tree_from_edges("center", edges)


def binary_tree(depth):
    """binary_topology — geometries.jl:82-92 (breadth-first node order)."""
    edges = []
    for d in range(1, depth + 1):
        for k in range(1, 2 ** d + 1):
            edges.append((f"n{d - 1}_{(k + 1) // 2}", f"n{d}_{k}"))
    return tree_from_edges("n0_1", edges)


def chain_tree(n):
    """mps_topology — geometries.jl:16-20 (rooted at the last site)."""
    ids = [f"site{i}" for i in range(1, n + 1)]
    edges = [(ids[i], ids[i - 1]) for i in range(n - 1, 0, -1)]
    return tree_from_edges(ids[-1], edges)


def neighbors(t, i):
    """Children in leg order, then the parent — topology.jl:66-70."""
    ns = list(t.children[i])
    if t.parent[i] >= 0:
        ns.append(t.parent[i])
    return ns


def childslot(t, p, c):
    """0-based slot of child c among p's children — topology.jl:73-77."""
    return t.children[p].index(c)


def postorder(t):
    """topology.jl:88-100."""
    out = []

    def rec(i):
        for c in t.children[i]:
            rec(c)
        out.append(i)

    rec(0)
    return out


def path_between(t, a, b):
    """topology.jl:132-141."""
    def to_root(x):
        p = [x]
        while t.parent[p[-1]] >= 0:
            p.append(t.parent[p[-1]])
        return p

    pa, pb = to_root(a), to_root(b)
    inb = set(pb)
    pos = next(i for i, x in enumerate(pa) if x in inb)
    lca = pa[pos]
    up = pa[: pos + 1]
    down = list(reversed(pb[: pb.index(lca)]))
    return up + down


def stateleg(t, u, w):
    """Flat (0-based) TTNS leg of u towards neighbour w — envcache.jl:852-853."""
    if t.parent[u] == w:
        return len(t.children[u]) + 1
    return childslot(t, u, w)


def opleg(t, u, w):
    """Flat (0-based) TTNO leg of u towards w — envcache.jl:855-857."""
    if t.parent[u] == w:
        return len(t.children[u]) + 2
    return childslot(t, u, w)


# ---------------------------------------------------------------------------
# Dense Hamiltonians — GraftTestUtils.jl:433-453 (kron ordering: the first
# physical site is the fastest index), spin matrices Symbolic.jl:177-186,
# models test/runtests.jl:19-41.
# ---------------------------------------------------------------------------

SX = np.array([[0, 1], [1, 0]], dtype=complex)
SZ = np.array([[1, 0], [0, -1]], dtype=complex)
SP = np.array([[0, 1], [0, 0]], dtype=complex)
SM = np.array([[0, 0], [1, 0]], dtype=complex)
ID2 = np.eye(2, dtype=complex)


def tree_edges(t):
    """(child, parent) pairs in node order — topology.jl:85."""
    return [(i, t.parent[i]) for i in range(t.n) if t.parent[i] >= 0]


def tfi_terms(t, J=1.0, g=0.9):
    """H = -J sum_edges Z_c Z_p - g sum_i X_i — runtests.jl:19-29."""
    terms = [(-J, [(c, SZ), (p, SZ)]) for c, p in tree_edges(t)]
    terms += [(-g, [(i, SX)]) for i in range(t.n)]
    return terms


def tfi_terms_random(t, Js, hs):
    """Per-edge couplings J_e (edge = child node) and per-site fields h_i."""
    terms = [(-Js[c], [(c, SZ), (p, SZ)]) for c, p in tree_edges(t)]
    terms += [(-hs[i], [(i, SX)]) for i in range(t.n)]
    return terms


def heisenberg_terms(t, J=1.0):
    """J/2 (S+S- + S-S+) + J/4 ZZ on every tree edge — runtests.jl:32-41."""
    terms = []
    for c, p in tree_edges(t):
        terms.append((J / 2, [(c, SP), (p, SM)]))
        terms.append((J / 2, [(c, SM), (p, SP)]))
        terms.append((J / 4, [(c, SZ), (p, SZ)]))
    return terms


def dense_hamiltonian(nsites, terms, d=2):
    """Sum of coefficient * kron(reverse(local matrices)) — GraftTestUtils.jl:438-453."""
    D = d ** nsites
    H = np.zeros((D, D), dtype=complex)
    for coeff, ops in terms:
        mats = [np.eye(d, dtype=complex) for _ in range(nsites)]
        for site, m in ops:
            mats[site] = m
        full = mats[nsites - 1]
        for s in range(nsites - 2, -1, -1):
            full = np.kron(full, mats[s])
        H += coeff * full
    return H


def exact_groundstate(H):
    """GraftTestUtils.jl:466-469."""
    w, v = np.linalg.eigh((H + H.conj().T) / 2)
    return w[0], v[:, 0]


def exact_evolve(H, v, dz):
    """exp(dz H) v by eigendecomposition of the hermitian H — GraftTestUtils.jl:472-473."""
    w, U = np.linalg.eigh((H + H.conj().T) / 2)
    return U @ (np.exp(dz * w) * (U.conj().T @ v))


def tfi_chain_exact_energy(n, J, g):
    """Ground energy of the open TFI chain -J sum Z_i Z_{i+1} - g sum X_i.

    Jordan–Wigner free fermions: E0 = -1/2 * sum_k eps_k where eps_k are the
    singular values of the (n x n) matrix A - B with A = -2 g I... here we use
    the standard BdG form: h_{ij} blocks built from J and g, E0 = -sum_k eps_k/2.
    """
    # BdG for H = -J sum Z_i Z_{i+1} - g sum X_i (a rotation of the usual
    # -J sum X X - g sum Z form): eps_k = singular values of M with
    # M_{ii} = 2 g, M_{i,i+1} = -2 J (Lieb–Schultz–Mattis).
    M = np.zeros((n, n))
    for i in range(n):
        M[i, i] = 2 * g
        if i + 1 < n:
            M[i, i + 1] = -2 * J
    eps = np.linalg.svd(M, compute_uv=False)
    return -0.5 * eps.sum()


# ---------------------------------------------------------------------------
# ncon — the label convention of TensorOperations' ncon used throughout the
# reference (labels > 0 contracted, < 0 open, output ordered -1, -2, ...).
# ---------------------------------------------------------------------------

_LETTERS = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"


_PATHS = {}      # contraction order per network and operand shapes


def ncon(tensors, labels, conjs=None):
    """Contract a labelled network pairwise with numpy.tensordot along the
    order numpy.einsum_path picks (every label joins at most two tensors, so
    each pair is one tensordot); open labels -1, -2, ... give the output order."""
    conjs = conjs or [False] * len(tensors)
    alllab = sorted({l for ls in labels for l in ls}, key=lambda x: (x < 0, abs(x)))
    sym = {l: _LETTERS[i] for i, l in enumerate(alllab)}
    opens = sorted([l for l in alllab if l < 0], key=lambda x: -x)
    subs = ",".join("".join(sym[l] for l in ls) for ls in labels)
    out = "".join(sym[l] for l in opens)
    ops = [np.conj(t) if c else np.asarray(t) for t, c in zip(tensors, conjs)]
    labs = [list(l) for l in labels]
    if len(ops) > 1:
        pkey = (subs, out, tuple(o.shape for o in ops))
        path = _PATHS.get(pkey)
        if path is None:
            path = _PATHS[pkey] = np.einsum_path(f"{subs}->{out}", *ops, optimize="greedy")[0][1:]
        for group in path:
            idx = sorted(group, reverse=True)
            picked = [(ops.pop(i), labs.pop(i)) for i in idx][::-1]
            A, la = picked[0]
            for Bt, lb in picked[1:]:             # a group of more than two: left to right
                shared = [x for x in la if x in lb]
                A = np.tensordot(A, Bt, axes=([la.index(x) for x in shared],
                                              [lb.index(x) for x in shared]))
                la = [x for x in la if x not in shared] + [x for x in lb if x not in shared]
            ops.append(A)
            labs.append(la)
    res, lr = ops[0], labs[0]
    if len(lr) != len(set(lr)):
        raise ValueError("ncon: a label is repeated within one tensor (trace) — not supported")
    return np.transpose(res, [lr.index(l) for l in opens]) if opens else res.reshape(())


# ---------------------------------------------------------------------------
# TTNS / TTNO dense forms — GraftTestUtils.jl:245-304
# ---------------------------------------------------------------------------


def ttns_to_dense(t, tensors, phys):
    """Full state vector, first physical site fastest — GraftTestUtils.jl:245-270."""
    sites = [i for i in range(t.n) if phys[i]]
    openlabel = {s: -(k + 1) for k, s in enumerate(sites)}
    rootlabel = -(len(sites) + 1)
    labels = []
    for i in range(t.n):
        A = tensors[i]
        idx = [0] * A.ndim
        for k, c in enumerate(t.children[i]):
            idx[k] = c + 1
        if phys[i]:
            idx[len(t.children[i])] = openlabel[i]
        idx[-1] = rootlabel if t.parent[i] < 0 else i + 1
        labels.append(idx)
    res = ncon(list(tensors), labels)
    return res.reshape(-1, order="F")


def ttno_to_dense(t, W, phys=None):
    """Dense operator of a TTNO closed by a ones cap — GraftTestUtils.jl:272-304.

    ``phys[i]`` false marks a node without physical legs (W :: children <- parent);
    the dense operator then acts on the physical nodes in node order."""
    phys = [True] * t.n if phys is None else list(phys)
    sites = [i for i in range(t.n) if phys[i]]
    pos = {s: k for k, s in enumerate(sites)}
    ns = len(sites)
    labels = []
    tensors = []
    for i in range(t.n):
        idx = [0] * W[i].ndim
        K = len(t.children[i])
        for k, c in enumerate(t.children[i]):
            idx[k] = c + 1
        if phys[i]:
            idx[K] = -(pos[i] + 1)
            idx[K + 1] = -(ns + pos[i] + 1)
        if t.parent[i] < 0:
            lbl = t.n + 1
            idx[-1] = lbl
            tensors.append(np.ones((W[i].shape[-1],), dtype=complex))
            labels.append([lbl])
        else:
            idx[-1] = i + 1
        tensors.append(W[i])
        labels.append(idx)
    arr = ncon(tensors, labels)
    D = int(round(math.sqrt(arr.size)))
    return arr.reshape(D, D, order="F")


def random_complex(rng, shape):
    """i.i.d. standard complex normal entries (E|z|^2 = 1), like Julia randn(ComplexF64)."""
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / math.sqrt(2)


def left_orth(A):
    """QR across the (codomain | last leg) split: A = Q . C with Q isometric."""
    shp = A.shape
    M = A.reshape(-1, shp[-1], order="F")
    Q, R = np.linalg.qr(M, mode="reduced")
    return Q.reshape(shp[:-1] + (Q.shape[1],), order="F"), R


def absorb_on_leg(A, C, k):
    """A' = C applied to codomain leg k of A — Backend.jl:533-543."""
    out = np.tensordot(C, A, axes=([1], [k]))
    return np.moveaxis(out, 0, k)


def canonicalize(t, tensors):
    """Leaf-to-root QR sweep, centre at the root — GraftTestUtils.jl:48-63."""
    tensors = list(tensors)
    for n in postorder(t):
        if t.parent[n] < 0:
            continue
        Q, C = left_orth(tensors[n])
        tensors[n] = Q
        p = t.parent[n]
        tensors[p] = absorb_on_leg(tensors[p], C, childslot(t, p, n))
    return tensors


def move_center(t, tensors, center, target):
    """QR gauge moves along the path — ttns.jl:126-203."""
    tensors = list(tensors)
    path = path_between(t, center, target)
    for a, b in zip(path[:-1], path[1:]):
        A = tensors[a]
        if t.parent[a] == b:
            Q, C = left_orth(A)
            tensors[a] = Q
            tensors[b] = absorb_on_leg(tensors[b], C, childslot(t, b, a))
        else:
            k = childslot(t, a, b)
            At = np.moveaxis(A, k, -1)
            Q, C = left_orth(At)                   # At = Q . C ; C :: Y <- V_e
            tensors[a] = np.moveaxis(Q, -1, k)
            B = tensors[b]
            tensors[b] = np.tensordot(B, C.T, axes=([B.ndim - 1], [0]))
    return tensors


def random_ttns(rng, t, d, D, center=0):
    """random_ttns(rng, ComplexF64, topo, phys, C^D) — GraftTestUtils.jl:73-91."""
    tensors = []
    for n in range(t.n):
        shape = tuple([D] * len(t.children[n]) + [d] + [1 if t.parent[n] < 0 else D])
        tensors.append(random_complex(rng, shape))
    tensors = canonicalize(t, tensors)
    if center != 0:
        tensors = move_center(t, tensors, 0, center)
    nrm = np.linalg.norm(tensors[center])
    tensors[center] = tensors[center] / nrm
    return tensors


# ---------------------------------------------------------------------------
# Environments — envcache.jl:1010-1101 (_build_env_spec) and :919-999
# (_build_env_ncon_reference): operands (ket, envs..., cap?, W, conj(bra)),
# open legs (ket, op, bra).
# ---------------------------------------------------------------------------


def build_env(t, ket, W, bra, u, v, envs):
    """env(u -> v), or the closed scalar when v is None."""
    A = ket[u]
    B = bra[u]
    K = len(t.children[u])
    aidx = [0] * A.ndim
    bidx = [0] * B.ndim
    widx = [0] * W[u].ndim
    tensors = [A]
    labels = [aidx]
    conjs = [False]
    nxt = [0]

    def fresh():
        nxt[0] += 1
        return nxt[0]

    pin, pout = fresh(), fresh()
    aidx[K] = pin
    widx[K + 1] = pin
    widx[K] = pout
    bidx[K] = pout
    for w in neighbors(t, u):
        la = stateleg(t, u, w)
        if w == v:
            aidx[la] = -1
            widx[opleg(t, u, w)] = -2
            bidx[la] = -3
        else:
            E = envs[(w, u)]
            e = [fresh(), fresh(), fresh()]
            aidx[la] = e[0]
            widx[opleg(t, u, w)] = e[1]
            bidx[la] = e[2]
            tensors.append(E)
            labels.append(e)
            conjs.append(False)
    if t.parent[u] < 0:
        ka, ko, kb = fresh(), fresh(), fresh()
        aidx[-1] = ka
        bidx[-1] = kb
        widx[-1] = ko
        tensors.append(np.ones((B.shape[-1], W[u].shape[-1], A.shape[-1]), dtype=complex))
        labels.append([kb, ko, ka])
        conjs.append(False)
    tensors.append(W[u])
    labels.append(widx)
    conjs.append(False)
    tensors.append(B)
    labels.append(bidx)
    conjs.append(True)
    res = ncon(tensors, labels, conjs)
    return res if v is not None else complex(res)


def all_envs(t, psi, W):
    """Every directed environment env(u -> v), built leaves-first."""
    envs = {}

    def get(u, v):
        if (u, v) in envs:
            return envs[(u, v)]
        for w in neighbors(t, u):
            if w != v:
                get(w, u)
        envs[(u, v)] = build_env(t, psi, W, psi, u, v, envs)
        return envs[(u, v)]

    for u in range(t.n):
        for v in neighbors(t, u):
            get(u, v)
    return envs


def env_closure(t, psi, W, envs, u):
    """<psi|H|psi> closed at node u — expectation.jl:32-41."""
    return build_env(t, psi, W, psi, u, None, envs)


# ---------------------------------------------------------------------------
# Effective maps — effective.jl:627-686 (_h1_spec), :739-749 (_h0_spec),
# :808-843 (two_site_space / two_site_tensor), :878-976 (_h2_spec).
# ---------------------------------------------------------------------------


def h1_network(t, W, envs, n, xshape):
    """Labels and static operands of the one-site map (effective.jl:627-686)."""
    K = len(t.children[n])
    Nx = len(xshape)
    xidx = [0] * Nx
    widx = [0] * W[n].ndim
    labels = [xidx, widx]
    statics = [W[n]]
    nxt = [0]

    def fresh():
        nxt[0] += 1
        return nxt[0]

    pin = fresh()
    xidx[K] = pin
    widx[K + 1] = pin
    widx[K] = -(K + 1)
    for w in neighbors(t, n):
        lx = stateleg(t, n, w)
        kk, oo = fresh(), fresh()
        xidx[lx] = kk
        widx[opleg(t, n, w)] = oo
        labels.append([kk, oo, -(lx + 1)])
        statics.append(envs[(w, n)])
    if t.parent[n] < 0:
        ka, ko = fresh(), fresh()
        xidx[-1] = ka
        widx[-1] = ko
        labels.append([ka, ko, -Nx])
        statics.append(np.ones((xshape[-1], W[n].shape[-1], xshape[-1]), dtype=complex))
    return labels, statics


def h1_apply(t, W, envs, n, x):
    labels, statics = h1_network(t, W, envs, n, x.shape)
    return ncon([x] + statics, labels)


def two_site_tensor(t, psi, n, m):
    """Theta = A_n contracted with A_m over their edge, all-codomain legs — effective.jl:827-843."""
    A, B = psi[n], psi[m]
    k = childslot(t, m, n)
    out = np.tensordot(A, B, axes=([A.ndim - 1], [k]))
    return out


def h2_network(t, W, envs, n, m, xshape):
    """Labels and statics of the two-site map (effective.jl:878-976)."""
    Kn, Km = len(t.children[n]), len(t.children[m])
    k = childslot(t, m, n)
    pn = Kn + 1
    Nx = len(xshape)
    xidx = [0] * Nx
    wn = [0] * W[n].ndim
    wm = [0] * W[m].ndim
    labels = [xidx, wn, wm]
    statics = [W[n], W[m]]
    nxt = [0]

    def fresh():
        nxt[0] += 1
        return nxt[0]

    def mpos(j):  # 0-based flat leg j of A_m (j != k) -> 0-based Theta leg
        return pn + (j if j < k else j - 1)

    ob = fresh()
    wn[-1] = ob
    wm[k] = ob
    pin = fresh()
    xidx[Kn] = pin
    wn[Kn + 1] = pin
    wn[Kn] = -(Kn + 1)
    pin = fresh()
    xidx[mpos(Km)] = pin
    wm[Km + 1] = pin
    wm[Km] = -(mpos(Km) + 1)
    for w in neighbors(t, n):
        if w == m:
            continue
        lx = childslot(t, n, w)
        kk, oo = fresh(), fresh()
        xidx[lx] = kk
        wn[lx] = oo
        labels.append([kk, oo, -(lx + 1)])
        statics.append(envs[(w, n)])
    for w in neighbors(t, m):
        if w == n:
            continue
        if t.parent[m] == w:
            lw, lx = W[m].ndim - 1, Nx - 1
        else:
            lw = childslot(t, m, w)
            lx = mpos(lw)
        kk, oo = fresh(), fresh()
        xidx[lx] = kk
        wm[lw] = oo
        labels.append([kk, oo, -(lx + 1)])
        statics.append(envs[(w, m)])
    if t.parent[m] < 0:
        ka, ko = fresh(), fresh()
        xidx[Nx - 1] = ka
        wm[-1] = ko
        labels.append([ka, ko, -Nx])
        statics.append(np.ones((xshape[-1], W[m].shape[-1], xshape[-1]), dtype=complex))
    return labels, statics


def h2_apply(t, W, envs, n, m, theta):
    labels, statics = h2_network(t, W, envs, n, m, theta.shape)
    return ncon([theta] + statics, labels)


def h0_apply(envs, n, m, C):
    """y = env(n->m) . C . env(m->n) — effective.jl:739-749."""
    return ncon([C, envs[(n, m)], envs[(m, n)]], [[1, 2], [1, 3, -1], [2, 3, -2]])


# ---------------------------------------------------------------------------
# Tangent-space embeddings: P maps a site/two-site/link tensor to the full
# Hilbert space with every other tensor fixed (GraftTestUtils.jl to_dense
# recipe).  Used to pin the effective maps against dense H.
# ---------------------------------------------------------------------------


def embedding_h1(t, psi, phys, n):
    shape = psi[n].shape
    cols = []
    for j in range(int(np.prod(shape))):
        e = np.zeros(int(np.prod(shape)), dtype=complex)
        e[j] = 1
        tens = list(psi)
        tens[n] = e.reshape(shape, order="F")
        cols.append(ttns_to_dense(t, tens, phys))
    return np.stack(cols, axis=1)


def embedding_h2(t, psi, phys, n, m):
    """Theta legs -> state, with A_n = identity-reshaped and A_m = Theta."""
    k = childslot(t, m, n)
    theta_shape = two_site_tensor(t, psi, n, m).shape
    An = psi[n]
    pn = An.ndim - 1
    size = int(np.prod(theta_shape))
    cols = []
    # Represent Theta exactly: A_n' = identity on its codomain legs (new bond
    # of dim prod(n codomain)), A_m' = Theta with that bond at slot k.
    dn = int(np.prod(An.shape[:-1]))
    ident = np.eye(dn, dtype=complex).reshape(An.shape[:-1] + (dn,), order="F")
    for j in range(size):
        e = np.zeros(size, dtype=complex)
        e[j] = 1
        th = e.reshape(theta_shape, order="F")
        thm = th.reshape((dn,) + theta_shape[pn:], order="F")
        Am = np.moveaxis(thm, 0, k)
        tens = list(psi)
        tens[n] = ident
        tens[m] = Am
        cols.append(ttns_to_dense(t, tens, phys))
    return np.stack(cols, axis=1)


def embedding_h0(t, psi_split, phys, n, m):
    """Link C (V_n-side x V_m-side) inserted on edge (n child, m parent)."""
    An = psi_split[n]
    Dn = An.shape[-1]
    Dm = psi_split[m].shape[childslot(t, m, n)]
    cols = []
    for j in range(Dn * Dm):
        e = np.zeros(Dn * Dm, dtype=complex)
        e[j] = 1
        C = e.reshape((Dn, Dm), order="F")
        tens = list(psi_split)
        tens[n] = np.tensordot(An, C, axes=([An.ndim - 1], [0]))
        cols.append(ttns_to_dense(t, tens, phys))
    return np.stack(cols, axis=1)


# ---------------------------------------------------------------------------
# Cost model — Backend.jl:432-441 (dense pair_cost = m*k*n complex MACs),
# cost.jl:148-174 (dense_cost), Backend.jl:328-368 (sector loop 2*m_q k_q n_q).
# ---------------------------------------------------------------------------


def dense_cost(labelsA, dimsA, labelsB, dimsB):
    shared = set(labelsA) & set(labelsB)
    oa = [i for i, l in enumerate(labelsA) if l not in shared]
    ca = [i for i, l in enumerate(labelsA) if l in shared]
    ob = [i for i, l in enumerate(labelsB) if l not in shared]
    cb = [labelsB.index(labelsA[i]) for i in ca]
    out_labels = [labelsA[i] for i in oa] + [labelsB[i] for i in ob]
    out_dims = [dimsA[i] for i in oa] + [dimsB[i] for i in ob]
    flops = float(np.prod([dimsA[i] for i in oa])) * float(np.prod([dimsA[i] for i in ca])) \
        * float(np.prod([dimsB[i] for i in ob]))
    return dict(oindA=oa, cindA=ca, oindB=ob, cindB=cb, labels=out_labels,
                dims=out_dims, flops=flops, peak_elements=float(np.prod(out_dims)))


def u1_chain_plan_costs(maps, order):
    """Dense and sector FLOPs of a product of U(1)-graded rank-2 maps.

    ``maps[i] = (cod, dom)`` with ``cod``/``dom`` dicts charge -> degeneracy;
    ``order`` is ``"left"`` ((M0 M1) M2 ...) or ``"right"`` (M0 (M1 M2)).
    For a rank-2 map cod <- dom the block of coupled charge q is
    cod[q] x dom[q]; one GEMM per charge costs 2 m_q k_q n_q
    (Backend.jl:344-353).
    """
    def dims(sp):
        return sum(sp.values())

    def product(a, b):
        (ca, da), (cb, db) = a, b
        dense = dims(ca) * dims(da) * dims(db)
        sec = sum(2.0 * ca.get(q, 0) * da.get(q, 0) * db.get(q, 0) for q in set(ca) | set(db))
        return (ca, db), dense, sec

    tot_dense = tot_sec = 0.0
    if order == "left":
        cur = maps[0]
        for nxt in maps[1:]:
            cur, d, s = product(cur, nxt)
            tot_dense += d
            tot_sec += s
    else:
        cur = maps[-1]
        for prv in reversed(maps[:-1]):
            cur, d, s = product(prv, cur)
            tot_dense += d
            tot_sec += s
    return tot_dense, tot_sec


# ---------------------------------------------------------------------------
# Pairwise plan execution with TensorOperations' structure (explicit permute,
# then one GEMM, then the output permute) — executor.jl:16-45,
# Backend.jl:166-170.  This is also the timed CPU baseline.
# ---------------------------------------------------------------------------


def contract_pair(A, oindA, cindA, conjA, B, cindB, oindB, conjB, out):
    """out is the flattened output permutation over (A open..., B open...)."""
    a = np.conj(A) if conjA else A
    b = np.conj(B) if conjB else B
    Ap = np.transpose(a, list(oindA) + list(cindA))
    Bp = np.transpose(b, list(cindB) + list(oindB))
    ma = int(np.prod([A.shape[i] for i in oindA])) if oindA else 1
    kk = int(np.prod([A.shape[i] for i in cindA])) if cindA else 1
    nb = int(np.prod([B.shape[i] for i in oindB])) if oindB else 1
    C = Ap.reshape(ma, kk) @ Bp.reshape(kk, nb)
    C = C.reshape([A.shape[i] for i in oindA] + [B.shape[i] for i in oindB])
    return np.ascontiguousarray(np.transpose(C, list(out)))


def execute_pair_steps(steps, operands, nslots, output_slot):
    """Walk a list of (a, b, dst, oindA, cindA, oindB, cindB, conjA, conjB, out) steps (0-based)."""
    slots = [None] * nslots
    for i, op in enumerate(operands):
        slots[i] = op
    for (a, b, dst, oA, cA, oB, cB, conjA, conjB, out) in steps:
        slots[dst] = contract_pair(slots[a], oA, cA, conjA, slots[b], cB, oB, conjB, out)
        slots[a] = None
        slots[b] = None
    return slots[output_slot]


# ---------------------------------------------------------------------------
# Spinful fermions (d = 4 per site: |0>, |up>, |dn>, |up dn>), Jordan-Wigner
# order = depth-first preorder of the tree (examples/hubbard_2x2_common.jl
# uses the same explicit JW encoding, :80-121).  Independent dense
# construction used to pin the Hubbard TTNO.
# ---------------------------------------------------------------------------


def preorder(t):
    out = []

    def rec(i):
        out.append(i)
        for c in t.children[i]:
            rec(c)

    rec(0)
    return out


def fermion_operators(t):
    """Dense global c_{i,s} (s = 0 up, 1 dn) for every site, JW in preorder."""
    n = t.n
    a = np.zeros((2, 2))
    a[0, 1] = 1.0                        # single-mode annihilator
    z = np.diag([1.0, -1.0])
    i2 = np.eye(2)
    # each site = two modes (up = mode 0 = fast bit, dn = mode 1); global mode
    # order: preorder of sites, up before dn
    order = preorder(t)
    mode_rank = {}
    for r, site in enumerate(order):
        mode_rank[(site, 0)] = 2 * r
        mode_rank[(site, 1)] = 2 * r + 1
    nm = 2 * n
    ops = {}
    for (site, s), rank in mode_rank.items():
        mats = []
        # kron position of mode (site, s): site-major in node index, up fastest
        for k in range(n):
            for ss in range(2):
                rk = mode_rank[(k, ss)]
                if (k, ss) == (site, s):
                    mats.append(a)
                elif rk < rank:
                    mats.append(z)
                else:
                    mats.append(i2)
        full = mats[-1]
        for m in reversed(mats[:-1]):
            full = np.kron(full, m)
        ops[(site, s)] = full.astype(complex)
    return ops


def hubbard_dense(t, hop=1.0, U=4.0):
    """-hop sum_{tree edges, s} (c^dag c + h.c.) + U sum n_up n_dn, dense."""
    c = fermion_operators(t)
    D = 4 ** t.n
    H = np.zeros((D, D), dtype=complex)
    for ch, p in tree_edges(t):
        for s in range(2):
            h = c[(p, s)].conj().T @ c[(ch, s)]
            H += -hop * (h + h.conj().T)
    for i in range(t.n):
        nu = c[(i, 0)].conj().T @ c[(i, 0)]
        nd = c[(i, 1)].conj().T @ c[(i, 1)]
        H += U * (nu @ nd)
    return H, c


def number_operators(c, n):
    N = sum(c[(i, s)].conj().T @ c[(i, s)] for i in range(n) for s in range(2))
    Sz2 = sum(c[(i, 0)].conj().T @ c[(i, 0)] - c[(i, 1)].conj().T @ c[(i, 1)] for i in range(n))
    return N, Sz2


def sector_ground_energy(H, N, Sz2, n_tot, sz2_tot):
    """Lowest eigenvalue of H in the (N, 2 S_z) sector."""
    d = np.real(np.diag(N))
    s = np.real(np.diag(Sz2))
    idx = np.where((np.abs(d - n_tot) < 1e-9) & (np.abs(s - sz2_tot) < 1e-9))[0]
    return np.linalg.eigvalsh(H[np.ix_(idx, idx)])[0]


# ---------------------------------------------------------------------------
# Gradients of the energy functional E = <psi|H|psi> / <psi|psi>
# ---------------------------------------------------------------------------


def identity_ttno(t, d=2):
    """TTNO of the identity operator: channel dimension 1 on every bond."""
    W = []
    for n in range(t.n):
        K = len(t.children[n])
        T = np.zeros([1] * K + [d, d, 1], dtype=complex)
        T[(0,) * K + (slice(None), slice(None), 0)] = np.eye(d)
        W.append(T)
    return W


def expectation_gradient(t, psi, W, d=2):
    """E, <psi|psi>, dE/d conj(A_n) for every site and d<psi|H|psi>/dW_n.

    Site gradients: (H_n A_n - E N_n A_n) / <psi|psi> with H_n, N_n the
    one-site maps of H and of the identity built from environments of the
    given (not necessarily canonical) tensors.  TTNO derivatives: the closure
    at n is linear in W_n, so entry idx is the closure with W_n replaced by the
    unit tensor at idx (no derivative network is used)."""
    Wi = identity_ttno(t, d)
    eH = all_envs(t, psi, W)
    eI = all_envs(t, psi, Wi)
    num = env_closure(t, psi, W, eH, 0)
    Z = env_closure(t, psi, Wi, eI, 0)
    E = (num / Z).real
    grads = [(h1_apply(t, W, eH, n, psi[n]) - E * h1_apply(t, Wi, eI, n, psi[n])) / Z
             for n in range(t.n)]
    Rs = []
    for n in range(t.n):
        R = np.zeros(W[n].shape, dtype=complex)
        for idx in np.ndindex(*W[n].shape):
            unit = np.zeros(W[n].shape, dtype=complex)
            unit[idx] = 1.0
            W2 = list(W)
            W2[n] = unit
            R[idx] = env_closure(t, psi, W2, eH, n)
        Rs.append(R)
    return E, Z, grads, Rs


def dense_energy(t, psi, terms):
    """<psi|H|psi> / <psi|psi> of the dense state vector (any gauge)."""
    v = ttns_to_dense(t, psi, [True] * t.n)
    H = dense_hamiltonian(t.n, terms)
    return float((np.vdot(v, H @ v) / np.vdot(v, v)).real)


# ---------------------------------------------------------------------------
# TDVP1 — tdvp.jl:171-301: the tree projector-splitting sweep on dense
# tensors.  Every local map is P^dagger H P with P the isometric embedding of
# the site (link) tensor, exponentiated exactly (no Krylov iteration).
# ---------------------------------------------------------------------------


def _expm_apply(h, dz, x):
    """exp(dz h) x for a hermitian matrix h, by eigendecomposition."""
    w, U = np.linalg.eigh((h + h.conj().T) / 2)
    return U @ (np.exp(dz * w) * (U.conj().T @ x))


def _local_exp(P, H, dz, X):
    x = X.reshape(-1, order="F")
    return _expm_apply(P.conj().T @ H @ P, dz, x).reshape(X.shape, order="F")


def _evolve_link_and_move(t, psi, H, u, v, dz, phys):
    """_evolve_link_and_move! with the QR splits (tdvp.jl:257-301): split the
    centre u towards v, evolve the link tensor with exp(-dz h0), absorb it into v."""
    psi = list(psi)
    if t.parent[u] == v:
        Q, C = left_orth(psi[u])                       # C :: V_new <- V_e
        psi[u] = Q
        C = _local_exp(embedding_h0(t, psi, phys, u, v), H, -dz, C)
        psi[v] = absorb_on_leg(psi[v], C, childslot(t, v, u))
    else:
        k = childslot(t, u, v)
        Q, Cd = left_orth(np.moveaxis(psi[u], k, -1))  # Cd :: Y <- V_e
        psi[u] = np.moveaxis(Q, -1, k)
        C = _local_exp(embedding_h0(t, psi, phys, v, u), H, -dz, Cd.T)   # :: V_e <- Y
        B = psi[v]
        psi[v] = np.tensordot(B, C, axes=([B.ndim - 1], [0]))
    return psi


def _tdvp1_sweep(t, psi, center, H, dz, phys, rev, terminal_dz=None, skip_initial_site=False):
    """_tdvp1_sweep! (tdvp.jl:197-251): post-order (reversed when rev) site
    updates; a link crossed child -> parent (parent -> child when rev) is
    evolved backward, any other crossing is a plain QR gauge move."""
    order = postorder(t)
    if rev:
        order = order[::-1]
    psi = move_center(t, psi, center, order[0])
    for i, n in enumerate(order):
        if not (skip_initial_site and i == 0):
            site_dz = terminal_dz if (terminal_dz is not None and i == len(order) - 1) else dz
            psi[n] = _local_exp(embedding_h1(t, psi, phys, n), H, site_dz, psi[n])
        if i == len(order) - 1:
            break
        seg = path_between(t, n, order[i + 1])
        for u, v in zip(seg[:-1], seg[1:]):
            if (t.parent[u] == v) != rev:
                psi = _evolve_link_and_move(t, psi, H, u, v, dz, phys)
            else:
                psi = move_center(t, psi, u, v)
    return psi, order[-1]


def tdvp1_step(t, psi, center, H, dz, order=2, phys=None):
    """step! of TDVP1 (tdvp.jl:171-195) on dense tensors with the dense
    Hamiltonian H: order 1 is one forward sweep; order 2 a forward half-step
    sweep whose last site takes the full step, then a reverse half-step sweep
    that skips its first site.  Returns (tensors, centre)."""
    phys = [True] * t.n if phys is None else list(phys)
    psi = list(psi)
    if order == 1:
        return _tdvp1_sweep(t, psi, center, H, dz, phys, rev=False)
    if order == 2:
        psi, center = _tdvp1_sweep(t, psi, center, H, dz / 2, phys, rev=False, terminal_dz=dz)
        return _tdvp1_sweep(t, psi, center, H, dz / 2, phys, rev=True, skip_initial_site=True)
    raise ValueError("order must be 1 or 2")


# ---------------------------------------------------------------------------
# Two-site DMRG on the host — GroundState.jl:102-143 (the CPU sweep baseline).
# Local problems by restarted Lanczos on the oracle's h2 map (krylovdim 20,
# absolute residual 1e-12); environments cached per directed edge and dropped when a
# tensor on their side changes (envcache.jl:812-844); truncated SVD split.
# ---------------------------------------------------------------------------


def _side(t, u, v):
    """Nodes on u's side of edge (u, v)."""
    acc, stack = set(), [(u, v)]
    while stack:
        i, frm = stack.pop()
        acc.add(i)
        for w in neighbors(t, i):
            if w != frm:
                stack.append((w, i))
    return acc


def _env(t, psi, W, envs, u, v):
    if (u, v) not in envs:
        for w in neighbors(t, u):
            if w != v:
                _env(t, psi, W, envs, w, u)
        envs[(u, v)] = build_env(t, psi, W, psi, u, v, envs)
    return envs[(u, v)]


def lanczos_lowest(mv, x0, krylovdim=20, tol=1e-12, maxiter=100):
    """Lowest eigenpair of a hermitian map by restarted Lanczos: full
    reorthogonalization, restart from the Ritz vector, until the Ritz residual
    |beta * last component| <= tol (absolute, as KrylovKit compares it)."""
    v = x0 / np.linalg.norm(x0)
    theta = 0.0
    for _ in range(maxiter):
        V = [v]
        T = np.zeros((krylovdim + 1, krylovdim), dtype=complex)
        k = 0
        for j in range(krylovdim):
            w = mv(V[j])
            for _ in range(2):
                c = np.array([np.vdot(u, w) for u in V])
                w = w - sum(ci * u for ci, u in zip(c, V))
                T[: len(V), j] += c
            beta = np.linalg.norm(w)
            T[j + 1, j] = beta
            k = j + 1
            if beta < 1e-14 * max(1.0, np.abs(T[: k, : k]).max()):
                break
            V.append(w / beta)
        Tk = T[:k, :k]
        lam, U = np.linalg.eigh((Tk + Tk.conj().T) / 2)
        theta, y = float(lam[0]), U[:, 0]
        v = sum(yi * u for yi, u in zip(y, V[:k]))
        v = v / np.linalg.norm(v)
        if abs(T[k, k - 1] * y[-1]) <= tol or k < krylovdim:
            break
    return theta, v


def dmrg2_sweeps(t, psi, W, center, maxdim, nsweeps=1, krylovdim=20, tol=1e-12):
    """dmrg2! sweeps (post-order bonds, then reversed) with maxdim truncation.
    Returns (tensors, centre, energies per sweep)."""
    psi = list(psi)
    sides = {}
    envs = {}

    def drop(nodes):
        for key in [k for k in envs if nodes & sides.setdefault(k, _side(t, *k))]:
            del envs[key]

    bonds = [n for n in postorder(t) if t.parent[n] >= 0]
    schedule = [(n, "m") for n in bonds] + [(n, "n") for n in reversed(bonds)]
    energies = []
    for _ in range(nsweeps):
        E = float("nan")
        for n, center_on in schedule:
            m = t.parent[n]
            moved = set(path_between(t, center, n))
            psi = move_center(t, psi, center, n)
            center = n
            drop(moved)
            for w in neighbors(t, n):
                if w != m:
                    _env(t, psi, W, envs, w, n)
            for w in neighbors(t, m):
                if w != n:
                    _env(t, psi, W, envs, w, m)
            theta = two_site_tensor(t, psi, n, m)
            shape = theta.shape
            size = theta.size

            def mv(x):
                y = h2_apply(t, W, envs, n, m, np.asarray(x).reshape(shape, order="F"))
                return y.reshape(-1, order="F")

            if size <= krylovdim + 1:
                Hm = np.stack([mv(e) for e in np.eye(size, dtype=complex)], axis=1)
                w_, v_ = np.linalg.eigh((Hm + Hm.conj().T) / 2)
                E, vec = float(w_[0]), v_[:, 0]
            else:
                E, vec = lanczos_lowest(mv, theta.reshape(-1, order="F"), krylovdim, tol)
            theta = vec.reshape(shape, order="F")
            pn = psi[n].ndim - 1
            M = theta.reshape(int(np.prod(shape[:pn])), -1, order="F")
            U, S, Vh = np.linalg.svd(M, full_matrices=False)
            r = max(1, min(maxdim, int(np.sum(S > 0))))
            U, S, Vh = U[:, :r], S[:r], Vh[:r]
            if center_on == "n":
                U = U * S[None, :]
            else:
                Vh = S[:, None] * Vh
            psi[n] = U.reshape(shape[:pn] + (r,), order="F")
            k = childslot(t, m, n)
            Bm = Vh.reshape((r,) + shape[pn:], order="F")
            psi[m] = np.moveaxis(Bm, 0, k)
            center = n if center_on == "n" else m
            drop({n, m})
        energies.append(E)
    return psi, center, energies
