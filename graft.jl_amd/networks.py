"""TTNS / TTNO containers and gauge moves — lib/GraftNetworks/src/Networks/{ttns,ttno}.jl.

Leg conventions (Backend.jl:9-21, ttno.jl:1-9):
  TTNS A[n] :: (children V_c...) (x) P  <-  V_parent   flat legs: children, phys, parent
  TTNO W[n] :: (children U_c...) (x) P_out <- P_in (x) U_parent
The root's parent leg has dimension 1.  Tensors live on the device (DTensor).
"""
from __future__ import annotations

import numpy as np

from . import backend as B
from .tensors import DTensor
from .trees import childslot, nchildren, nnodes, path_between, postorder


class TTNS:
    """Tree tensor network state (ttns.jl:22-39)."""

    def __init__(self, topo, tensors, center, hasphys=None):
        if len(tensors) != nnodes(topo):
            raise ValueError("need one tensor per node")
        self.topo = topo
        self.tensors = list(tensors)
        self.hasphys = list(hasphys) if hasphys is not None else [
            t.ncod == nchildren(topo, i) + 1 for i, t in enumerate(self.tensors)]
        self.center = int(center)
        self.check_arrows()

    def __getitem__(self, n):
        return self.tensors[n]

    def copy(self):
        return TTNS(self.topo, [t.clone() for t in self.tensors], self.center, self.hasphys)

    def physleg(self, n):
        if not self.hasphys[n]:
            raise ValueError(f"node {self.topo.ids[n]} has no physical leg")
        return nchildren(self.topo, n)

    def parentleg(self, n):
        return self.tensors[n].rank - 1

    def bond_dims(self):
        return [self.tensors[c].dims[-1] for c in range(nnodes(self.topo)) if self.topo.parent[c] >= 0]

    def check_arrows(self):
        """ttns.jl:79-93."""
        t = self.topo
        for n in range(nnodes(t)):
            A = self.tensors[n]
            if A.rank - A.ncod != 1:
                raise ValueError(f"node {t.ids[n]}: expected exactly one parent (domain) leg")
            if A.ncod != nchildren(t, n) + int(self.hasphys[n]):
                raise ValueError(f"node {t.ids[n]}: codomain legs != children (+ physical)")
            for k, c in enumerate(t.children[n]):
                if A.dims[k] != self.tensors[c].dims[-1]:
                    raise ValueError(f"edge {t.ids[c]} -> {t.ids[n]}: slot dim {A.dims[k]} != child "
                                     f"parent-leg dim {self.tensors[c].dims[-1]}")
        return True


class TTNO:
    """Tree tensor network operator (ttno.jl:17-34)."""

    def __init__(self, topo, tensors, ishermitian=False):
        if len(tensors) != nnodes(topo):
            raise ValueError("need one tensor per node")
        self.topo = topo
        self.tensors = list(tensors)
        self.hasphys = [t.rank - t.ncod == 2 for t in self.tensors]
        self.ishermitian = bool(ishermitian)

    def __getitem__(self, n):
        return self.tensors[n]


def norm(psi):
    """Norm of a canonical state: the norm of its centre tensor."""
    return float(psi.tensors[psi.center].data.abs().pow(2).sum().sqrt().item())


def normalize_(psi):
    A = psi.tensors[psi.center]
    nrm = float(np.sqrt(float((A.data.abs() ** 2).sum().item())))
    psi.tensors[psi.center] = DTensor(A.dims, A.data / nrm, A.ncod, A.blocks)
    return psi


def _move_center_edge_(psi, m, cache):
    """ttns.jl:181-203 (QR up- and down-moves)."""
    if psi.tensors[psi.center].blocks is not None:
        if hasattr(psi.tensors[psi.center].blocks, "dense_from_flat"):
            from . import su2
            return su2.move_center_edge_(psi, m, cache)
        from . import sectors
        return sectors.move_center_edge_(psi, m, cache)
    t = psi.topo
    n = psi.center
    A = psi.tensors[n]
    if t.parent[n] == m:
        Q, C = B.left_orth(A)
        psi.tensors[n] = Q
        psi.tensors[m] = B.absorb_on_leg(psi.tensors[m], C, childslot(t, m, n))
    else:
        k = childslot(t, n, m)
        Q, C = B.orth_factor_leg(A, k)
        psi.tensors[n] = Q
        psi.tensors[m] = B.mul_domain(psi.tensors[m], B.transpose2(C))
    psi.center = m
    if cache is not None:
        cache.invalidate_edge_(n, m)
    return psi


def move_center_(psi, target, cache=None):
    """move_center! — ttns.jl:126-132."""
    path = path_between(psi.topo, psi.center, target)
    for v in path[1:]:
        _move_center_edge_(psi, v, cache)
    return psi


def update_tensor_(psi, n, A, caches=(), gauge=True):
    """update_tensor! — ttns.jl:213-222."""
    if gauge and n != psi.center:
        raise ValueError("updating a non-center tensor breaks the gauge invariant; "
                         "move_center_ first or pass gauge=False")
    psi.tensors[n] = A
    for c in caches:
        c.invalidate_node_(n)
    return psi


# ---------------------------------------------------------------------------
# Host-side construction of initial states (input preparation).
# ---------------------------------------------------------------------------


def _host_left_orth(A):
    shp = A.shape
    Q, R = np.linalg.qr(A.reshape(-1, shp[-1], order="F"), mode="reduced")
    return Q.reshape(shp[:-1] + (Q.shape[1],), order="F"), R


def random_ttns_host(rng, topo, phys, bond, center=0):
    """random_ttns (GraftTestUtils.jl:73-91) on the host: i.i.d. complex normal
    entries, leaf-to-root QR canonicalization, centre at the root, normalized.

    ``phys``: dict node -> physical dim (0/absent = no physical leg) or an int
    for every node; ``bond``: int or dict child-node -> bond dim."""
    n_nodes = nnodes(topo)

    def pdim(n):
        return phys if isinstance(phys, int) else phys.get(n, 0)

    def edim(n):
        return bond if isinstance(bond, int) else bond[n]

    tensors = []
    for n in range(n_nodes):
        shape = [edim(c) for c in topo.children[n]]
        if pdim(n):
            shape.append(pdim(n))
        shape.append(1 if topo.parent[n] < 0 else edim(n))
        a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)
        tensors.append(a)
    for n in postorder(topo):
        if topo.parent[n] < 0:
            continue
        Q, C = _host_left_orth(tensors[n])
        tensors[n] = Q
        p = topo.parent[n]
        k = childslot(topo, p, n)
        tensors[p] = np.moveaxis(np.tensordot(C, tensors[p], axes=([1], [k])), 0, k)
    r = topo.root
    tensors[r] = tensors[r] / np.linalg.norm(tensors[r])
    hasphys = [bool(pdim(n)) for n in range(n_nodes)]
    return tensors, hasphys


def random_ttns(rng, topo, phys, bond, center=0):
    """Random canonical TTNS uploaded to the device, centre at ``center``."""
    host, hasphys = random_ttns_host(rng, topo, phys, bond)
    tensors = [DTensor.from_numpy(a, a.ndim - 1) for a in host]
    psi = TTNS(topo, tensors, topo.root, hasphys)
    if center != topo.root:
        move_center_(psi, center)
    return psi


def ttns_from_host(topo, arrays, center):
    return TTNS(topo, [DTensor.from_numpy(a, a.ndim - 1) for a in arrays], center)
