"""Device-resident block store: TensorKit TensorMap blocks in HBM.

A ``DTensor`` is one TensorMap with trivial sectors (a single dense block,
flat legs = codomain legs then domain legs, first leg fastest, complex128
interleaved), or an abelian sector tensor (a list of packed dense blocks,
one per allowed tuple of sector indices).  The bytes live in a torch
allocation on the current HIP device; torch is used only as the allocator.
Every arithmetic operation on DTensors goes through libgraft_heff.so.
"""
from __future__ import annotations

import ctypes


import numpy as np
import torch

from . import _native as nat


class _StructKey:
    """Hashable identity of a block structure with the hash computed once."""

    __slots__ = ("value", "_hash")

    def __init__(self, value):
        self.value = value
        self._hash = hash(value)

    def __hash__(self):
        return self._hash

    def __eq__(self, other):
        return isinstance(other, _StructKey) and (self is other or (
            self._hash == other._hash and self.value == other.value))


class BlockStructure:
    """Abelian block layout: flat legs (SectorLeg), block sector tuples, offsets."""

    def __init__(self, legs, blocks, _int_tuples=False):
        self.legs = list(legs)
        if not _int_tuples:
            blocks = [tuple(int(x) for x in b) for b in blocks]
        self.blocks = blocks
        self.index = {b: i for i, b in enumerate(blocks)}
        n = len(self.legs)
        if blocks and n:
            arr = np.asarray(blocks, dtype=np.int64).reshape(len(blocks), n)
            dims = np.empty_like(arr)
            for l, leg in enumerate(self.legs):
                dims[:, l] = np.asarray([d for _, d in leg.sectors], dtype=np.int64)[arr[:, l]]
            sizes = dims.prod(axis=1)
        else:
            dims = np.zeros((len(blocks), 0), dtype=np.int64)
            sizes = np.ones(len(blocks), dtype=np.int64)
        self._dims = dims
        self._block_dims = None
        self.offsets = (np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if len(sizes)
                        else np.zeros(0, np.int64))
        self.sizes = sizes.tolist()
        self.total = int(sizes.sum())

    @property
    def block_dims(self):
        """Per block, its extent on every leg (tuples, built on first use)."""
        if self._block_dims is None:
            self._block_dims = [tuple(r) for r in self._dims.tolist()]
        return self._block_dims

    @staticmethod
    def invariant(legs):
        """All sector tuples whose flat charges sum to zero (sorted, first leg slowest)."""
        legs = list(legs)
        if not legs:
            return BlockStructure(legs, [()])
        from .planning import _leg_codes
        n = len(legs)
        codes = [[q for q, _ in _leg_codes(l)] for l in legs]
        # reach[l]: charge codes the legs l.. can produce; a partial tuple over
        # legs < l survives only if its charge can still be cancelled.
        reach = [None] * (n + 1)
        reach[n] = {0}
        for l in range(n - 1, -1, -1):
            reach[l] = {q + r for q in set(codes[l]) for r in reach[l + 1]}
        partial = [((), 0)]
        for l in range(n):
            nxt = []
            rl = reach[l + 1]
            cl = list(enumerate(codes[l]))
            for tup, tot in partial:
                for s, q in cl:
                    if -(tot + q) in rl:
                        nxt.append((tup + (s,), tot + q))
            partial = nxt
        return BlockStructure(legs, sorted(tup for tup, _ in partial), _int_tuples=True)

    @property
    def dims(self):
        return tuple(l.dim for l in self.legs)

    def key(self):
        k = getattr(self, "_key", None)
        if k is None:
            k = self._key = _StructKey((tuple(self.legs), tuple(self.blocks)))
        return k

    def permuted(self, perm):
        legs = [self.legs[p] for p in perm]
        blocks = [tuple(b[p] for p in perm) for b in self.blocks]
        order = sorted(range(len(blocks)), key=lambda i: blocks[i])
        return BlockStructure(legs, [blocks[i] for i in order], _int_tuples=True)

    def desc(self):
        d = getattr(self, "_desc", None)
        if d is None:
            d = self._desc = nat.DescKeepalive(
                self.dims, ([l.degs() for l in self.legs],
                            np.asarray(self.blocks, dtype=np.int32).reshape(len(self.blocks), -1),
                            self.offsets, None))
        return d


class TensorProto:
    """Structure of a tensor without data (planning prototype)."""

    def __init__(self, dims, ncod=None, blocks=None):
        self.dims = tuple(int(d) for d in dims)
        self.ncod = len(self.dims) if ncod is None else int(ncod)
        self.blocks = blocks
        self._desc = None

    @property
    def rank(self):
        return len(self.dims)

    def desc(self):
        if self._desc is None:
            self._desc = self.blocks.desc() if self.blocks is not None else nat.DescKeepalive(self.dims)
        return self._desc

    def structure_key(self):
        return (self.dims, None if self.blocks is None else self.blocks.key())


def _device():
    return torch.device("cuda", torch.cuda.current_device())


class DTensor:
    """One device-resident tensor with flat leg dims ``dims``; ``ncod`` codomain legs."""

    __slots__ = ("dims", "ncod", "data", "blocks", "_desc")

    def __init__(self, dims, data, ncod=None, blocks=None):
        self.dims = tuple(int(d) for d in dims)
        self.ncod = len(self.dims) if ncod is None else int(ncod)
        self.blocks = blocks
        n = blocks.total if blocks is not None else int(np.prod(self.dims)) if self.dims else 1
        if data.dtype != torch.complex128 or data.dim() != 1 or data.numel() != n:
            raise ValueError(f"DTensor data must be a flat complex128 tensor of {n} elements")
        if data.device.type != "cuda":
            raise ValueError("DTensor data must live on the HIP device")
        self.data = data
        self._desc = None

    # -- construction -------------------------------------------------------
    @staticmethod
    def empty(dims, ncod=None, blocks=None):
        n = blocks.total if blocks is not None else int(np.prod(dims)) if len(dims) else 1
        return DTensor(dims, torch.empty(n, dtype=torch.complex128, device=_device()), ncod, blocks)

    @staticmethod
    def zeros(dims, ncod=None, blocks=None):
        n = blocks.total if blocks is not None else int(np.prod(dims)) if len(dims) else 1
        return DTensor(dims, torch.zeros(n, dtype=torch.complex128, device=_device()), ncod, blocks)

    @staticmethod
    def from_numpy(arr, ncod=None):
        """Upload a dense numpy array whose axes are the flat legs."""
        arr = np.asarray(arr, dtype=np.complex128)
        flat = np.ascontiguousarray(arr.reshape(-1, order="F"))
        data = torch.from_numpy(flat).to(_device())
        return DTensor(arr.shape, data, ncod)

    @staticmethod
    def from_blocks(blocks, host_blocks, ncod=None):
        """Upload an abelian tensor from a list of dense numpy blocks (block order)."""
        flat = np.concatenate([np.asarray(b, dtype=np.complex128).reshape(-1, order="F")
                               for b in host_blocks]) if host_blocks else np.zeros(0, np.complex128)
        data = torch.from_numpy(np.ascontiguousarray(flat)).to(_device())
        return DTensor(blocks.dims, data, ncod, blocks)

    def to_numpy(self):
        """Download; dense tensors as one array, sector tensors as the dense embedding."""
        host = self.data.detach().cpu().numpy()
        if self.blocks is None:
            return host.reshape(self.dims, order="F")
        if hasattr(self.blocks, "dense_from_flat"):     # non-abelian structures embed themselves
            return self.blocks.dense_from_flat(host)
        out = np.zeros(self.dims, dtype=np.complex128)
        starts = [np.concatenate([[0], np.cumsum(l.degs())[:-1]]) for l in self.blocks.legs]
        for i, b in enumerate(self.blocks.blocks):
            d = self.blocks.block_dims[i]
            o = int(self.blocks.offsets[i])
            blk = host[o:o + self.blocks.sizes[i]].reshape(d, order="F")
            sl = tuple(slice(int(starts[l][s]), int(starts[l][s]) + d[l]) for l, s in enumerate(b))
            out[sl] = blk
        return out

    def block_numpy(self, i):
        host = self.data.detach().cpu().numpy()
        o = int(self.blocks.offsets[i])
        return host[o:o + self.blocks.sizes[i]].reshape(self.blocks.block_dims[i], order="F")

    # -- properties ---------------------------------------------------------
    @property
    def rank(self):
        return len(self.dims)

    @property
    def numel(self):
        return self.data.numel()

    @property
    def nbytes(self):
        return self.data.numel() * 16

    @property
    def ptr(self):
        return self.data.data_ptr()

    def desc(self):
        if self._desc is None:
            self._desc = self.blocks.desc() if self.blocks is not None else nat.DescKeepalive(self.dims)
        return self._desc

    def structure_key(self):
        return (self.dims, None if self.blocks is None else self.blocks.key())

    def clone(self):
        return DTensor(self.dims, self.data.clone(), self.ncod, self.blocks)

    def like(self):
        return DTensor.empty(self.dims, self.ncod, self.blocks)

    def __repr__(self):
        kind = "dense" if self.blocks is None else f"{len(self.blocks.blocks)} blocks"
        return f"DTensor(dims={self.dims}, ncod={self.ncod}, {kind})"


# ---------------------------------------------------------------------------
# Krylov vector algebra on DTensors (K8 kernels)
# ---------------------------------------------------------------------------


def _ctx():
    return nat.default_context()


def inner(x, y):
    """<x, y> = sum conj(x) y (complex)."""
    out = (ctypes.c_double * 2)()
    nat.check(nat.lib().graft_vec_dotc(_ctx().handle, x.numel, ctypes.c_void_p(x.ptr),
                                       ctypes.c_void_p(y.ptr), out))
    return complex(out[0], out[1])


def norm(x):
    out = ctypes.c_double()
    nat.check(nat.lib().graft_vec_nrm2(_ctx().handle, x.numel, ctypes.c_void_p(x.ptr),
                                       ctypes.byref(out)))
    return out.value


def axpby_(alpha, x, beta, y):
    """y <- alpha x + beta y."""
    a = (ctypes.c_double * 2)(complex(alpha).real, complex(alpha).imag)
    b = (ctypes.c_double * 2)(complex(beta).real, complex(beta).imag)
    nat.check(nat.lib().graft_vec_axpby(_ctx().handle, y.numel, a, ctypes.c_void_p(x.ptr), b,
                                        ctypes.c_void_p(y.ptr)))
    return y


def scale_(y, alpha):
    return axpby_(0.0, y, alpha, y)


def multi_inner(V, y):
    """[<V_j, y>] in one pass over y."""
    k = len(V)
    res = []
    for s in range(0, k, 64):
        chunk = V[s:s + 64]
        out = (ctypes.c_double * (2 * len(chunk)))()
        nat.check(nat.lib().graft_vec_multi_dotc(_ctx().handle, y.numel, len(chunk),
                                                 nat.ptr_array([v.ptr for v in chunk]),
                                                 ctypes.c_void_p(y.ptr), out))
        res += [complex(out[2 * j], out[2 * j + 1]) for j in range(len(chunk))]
    return np.array(res, dtype=np.complex128)


def multi_axpy_(V, c, y):
    """y <- y + sum_j c_j V_j in one pass over y."""
    c = np.asarray(c, dtype=np.complex128)
    for s in range(0, len(V), 64):
        chunk = V[s:s + 64]
        cc = c[s:s + 64]
        arr = (ctypes.c_double * (2 * len(chunk)))(*[x for z in cc for x in (z.real, z.imag)])
        nat.check(nat.lib().graft_vec_multi_axpy(_ctx().handle, y.numel, len(chunk),
                                                 nat.ptr_array([v.ptr for v in chunk]), arr,
                                                 ctypes.c_void_p(y.ptr)))
    return y


def lincomb(V, c, like):
    """sum_j c_j V_j into a fresh tensor."""
    y = DTensor.zeros(like.dims, like.ncod, like.blocks)
    return multi_axpy_(V, c, y)
