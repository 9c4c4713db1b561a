"""Tree topologies (host side) — the reference's Trees module.

Mirrors lib/GraftFoundation/src/Trees/topology.jl:22-233 and
geometries.jl:16-92: an immutable tree whose child order defines the leg
order of node tensors.  Node indices are 0-based here (the reference is
1-based); the root is node 0 and ``parent[root] == -1``.
"""
from __future__ import annotations


class TreeTopology:
    """TreeTopology(root, edges) with edges given top-down as (parent, child)."""

    def __init__(self, root, edges):
        self.ids = [root]
        self.index = {root: 0}
        self.parent = [-1]
        self.children = [[]]
        for p, c in edges:
            if p not in self.index:
                raise ValueError(f"parent {p} unknown when adding edge {p} => {c} (edges must be given top-down)")
            if c in self.index:
                raise ValueError(f"node {c} added twice")
            self.ids.append(c)
            self.index[c] = len(self.ids) - 1
            self.parent.append(self.index[p])
            self.children.append([])
            self.children[self.index[p]].append(self.index[c])
        self.root = 0
        self.depth = [0] * len(self.ids)
        for i in range(1, len(self.ids)):
            self.depth[i] = self.depth[self.parent[i]] + 1

    def __eq__(self, other):
        return (isinstance(other, TreeTopology) and self.ids == other.ids
                and self.parent == other.parent and self.children == other.children)

    def __hash__(self):
        return hash((tuple(self.ids), tuple(self.parent)))

    def __repr__(self):
        return f"TreeTopology({len(self.ids)} nodes, root={self.ids[0]})"


def nnodes(t):
    return len(t.ids)


def nchildren(t, i):
    return len(t.children[i])


def isroot(t, i):
    return t.parent[i] < 0


def neighbors(t, i):
    """Children in leg order, then the parent — topology.jl:66-70."""
    ns = list(t.children[i])
    if t.parent[i] >= 0:
        ns.append(t.parent[i])
    return ns


def childslot(t, p, c):
    """0-based leg slot of child c on parent p — topology.jl:73-77."""
    try:
        return t.children[p].index(c)
    except ValueError:
        raise ValueError(f"{t.ids[c]} is not a child of {t.ids[p]}") from None


def edges(t):
    """(child, parent) pairs — topology.jl:85."""
    return [(i, t.parent[i]) for i in range(nnodes(t)) if t.parent[i] >= 0]


def postorder(t):
    """Children before parents, root last — topology.jl:88-100."""
    out = []
    stack = [(t.root, False)]
    while stack:
        i, done = stack.pop()
        if done:
            out.append(i)
            continue
        stack.append((i, True))
        for c in reversed(t.children[i]):
            stack.append((c, False))
    return out


def path_to_root(t, a):
    p = [a]
    while t.parent[p[-1]] >= 0:
        p.append(t.parent[p[-1]])
    return p


def path_between(t, a, b):
    """Unique path a -> b inclusive — topology.jl:132-141."""
    pa, pb = path_to_root(t, a), path_to_root(t, b)
    inb = set(pb)
    pos = next(i for i, x in enumerate(pa) if x in inb)
    lca = pa[pos]
    return pa[: pos + 1] + list(reversed(pb[: pb.index(lca)]))


def subtree_nodes(t, n, avoiding):
    """Component containing n once edge (n, avoiding) is cut — topology.jl:150-165."""
    acc = []
    stack = [(n, avoiding)]
    while stack:
        i, frm = stack.pop()
        acc.append(i)
        for m in neighbors(t, i):
            if m != frm:
                stack.append((m, i))
    return acc


def mps_topology(n, prefix="site"):
    """Chain rooted at the last site — geometries.jl:16-20."""
    ids = [f"{prefix}{i}" for i in range(1, n + 1)]
    return TreeTopology(ids[-1], [(ids[i], ids[i - 1]) for i in range(n - 1, 0, -1)])


def star_topology(nbranches, branchlength, center="center", prefix="b"):
    """Centre plus nbranches chains — geometries.jl:64-75."""
    es = []
    for b in range(1, nbranches + 1):
        prev = center
        for l in range(1, branchlength + 1):
            node = f"{prefix}{b}_{l}"
            es.append((prev, node))
            prev = node
    return TreeTopology(center, es)


def binary_topology(depth, prefix="n"):
    """Perfect binary tree, root at depth 0 — geometries.jl:82-92."""
    es = []
    for d in range(1, depth + 1):
        for k in range(1, 2 ** d + 1):
            es.append((f"{prefix}{d - 1}_{(k + 1) // 2}", f"{prefix}{d}_{k}"))
    return TreeTopology(f"{prefix}0_1", es)


def fork_topology(nteeth, toothlength, prefix="spine"):
    """Spine with one chain per spine node — geometries.jl:101-118."""
    es = []
    for s in range(2, nteeth + 1):
        es.append((f"{prefix}{s - 1}", f"{prefix}{s}"))
    for s in range(1, nteeth + 1):
        prev = f"{prefix}{s}"
        for l in range(1, toothlength + 1):
            node = f"tooth{s}_{l}"
            es.append((prev, node))
            prev = node
    return TreeTopology(f"{prefix}1", es)
