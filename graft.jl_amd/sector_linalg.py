"""Per-sector factorizations of U(1) x U(1) block tensors on the device.

TensorKit factorizes a block-sparse TensorMap one coupled sector at a time:
every block that shares the sector of the factorized leg (a QR across one
leg) or the total charge of the row legs (an SVD across a leg bipartition)
belongs to one dense matrix, rows = the tuples of the row legs, columns =
the tuples of the column legs.  The reference calls these through
MatrixAlgebraKit: ``left_orth`` / ``orth_factor_leg`` for the gauge moves
(Backend.jl:595-609, ttns.jl:181-203, tdvp.jl:289-302) and ``svd_trunc``
for the two-site split (Backend.jl:474-505, effective.jl:852-876).

Here every sector matrix is gathered on the device from the blocked tensor
in one indexed copy (a per-element map built once per block structure), the
per-sector factorizations run on the device (Householder or shifted
Cholesky QR; exact SVD, or the deflated Gram spectrum of ``backend`` for
large sectors), and the factors are scattered back into block tensors by
the same kind of map.  Nothing is downloaded except singular values.
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np
import torch

from . import backend as B
from .planning import SectorLeg
from .tensors import BlockStructure, DTensor


# ---------------------------------------------------------------------------
# Sector matrix maps
# ---------------------------------------------------------------------------


class SectorMatrices:
    """Layout of the per-group dense matrices of one block structure.

    Blocks are grouped by ``group(block tuple)``; inside a group, rows are the
    distinct tuples of ``row_legs`` (sorted) and columns those of
    ``col_legs`` (sorted), each spanning the product of its legs' dimensions
    (first leg fastest).  Group g's matrix is R_g x C_g, column-major, at
    element ``base[g]`` of one packed buffer.  ``dst`` maps every element of
    the blocked tensor to its position in that buffer."""

    def __init__(self, st, row_legs, col_legs, group):
        self.row_legs, self.col_legs = list(row_legs), list(col_legs)
        dims = st._dims                              # nblocks x rank (numpy)
        info = {}
        for b, key in enumerate(st.blocks):
            g = group(key)
            rt = tuple(key[l] for l in row_legs)
            ct = tuple(key[l] for l in col_legs)
            e = info.setdefault(g, ({}, {}))
            e[0][rt] = int(np.prod([dims[b, l] for l in row_legs])) if row_legs else 1
            e[1][ct] = int(np.prod([dims[b, l] for l in col_legs])) if col_legs else 1
        self.groups = sorted(info)
        self.gindex = {g: i for i, g in enumerate(self.groups)}
        self.rows, self.cols, self.rowoff, self.coloff = [], [], [], []
        self.R, self.C, self.base = [], [], []
        total = 0
        for g in self.groups:
            rd, cd = info[g]
            rts, cts = sorted(rd), sorted(cd)
            ro = dict(zip(rts, np.concatenate([[0], np.cumsum([rd[t] for t in rts])[:-1]]).tolist()))
            co = dict(zip(cts, np.concatenate([[0], np.cumsum([cd[t] for t in cts])[:-1]]).tolist()))
            R, C = sum(rd.values()), sum(cd.values())
            self.rows.append(rts)
            self.cols.append(cts)
            self.rowoff.append(ro)
            self.coloff.append(co)
            self.R.append(R)
            self.C.append(C)
            self.base.append(total)
            total += R * C
        self.total = total
        self.dst = self._element_map(st, group) if st.total else None

    def _element_map(self, st, group):
        nb, rank = len(st.blocks), len(st.legs)
        dims = st._dims.astype(np.int64).reshape(nb, rank)
        bstride = np.ones((nb, rank), dtype=np.int64)    # block storage: first leg fastest
        for l in range(1, rank):
            bstride[:, l] = bstride[:, l - 1] * dims[:, l - 1]
        mstride = np.zeros((nb, rank), dtype=np.int64)   # contribution of each leg's index
        start = np.zeros(nb, dtype=np.int64)
        for b, key in enumerate(st.blocks):
            gi = self.gindex[group(key)]
            R = self.R[gi]
            ro = self.rowoff[gi][tuple(key[l] for l in self.row_legs)]
            co = self.coloff[gi][tuple(key[l] for l in self.col_legs)]
            start[b] = self.base[gi] + ro + co * R
            s = 1
            for l in self.row_legs:
                mstride[b, l] = s
                s *= dims[b, l]
            s = R
            for l in self.col_legs:
                mstride[b, l] = s
                s *= dims[b, l]
        dev = torch.device("cuda", torch.cuda.current_device())
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
        off = t(st.offsets.astype(np.int64))
        i = torch.arange(st.total, device=dev, dtype=torch.int64)
        b = torch.searchsorted(off, i, right=True) - 1
        loc = i - off[b]
        dst = t(start)[b]
        tb, td, tm = t(bstride), t(dims), t(mstride)
        for l in range(rank):
            idx = torch.div(loc, tb[:, l][b], rounding_mode="floor") % td[:, l][b]
            dst += idx * tm[:, l][b]
        return dst

    def matrix(self, buf, g):
        """Group g's R x C matrix (a column-major view into ``buf``)."""
        gi = self.gindex[g]
        R, C, o = self.R[gi], self.C[gi], self.base[gi]
        return buf[o:o + R * C].view(C, R).T

    def gather(self, t):
        """Packed sector matrices of the block tensor t (entries no block covers are zero)."""
        buf = torch.zeros(self.total, dtype=torch.complex128, device=t.data.device)
        if self.dst is not None:
            buf[self.dst] = t.data
        return buf

    def scatter(self, buf, st, ncod):
        """The block tensor of structure st whose blocks read the packed buffer."""
        data = buf[self.dst] if self.dst is not None else torch.zeros(
            0, dtype=torch.complex128, device=buf.device)
        return DTensor(st.dims, data, ncod, st)


_MAPS = OrderedDict()
MAP_CACHE_BYTES = 2 << 30      # device bytes the cached element maps may hold


def _maps(st, row_legs, col_legs, kind, group):
    key = (st.key(), tuple(row_legs), tuple(col_legs), kind)
    m = _MAPS.get(key)
    if m is None:
        m = SectorMatrices(st, row_legs, col_legs, group)
        _MAPS[key] = m
        used = sum(8 * x.total for x in _MAPS.values() if x.dst is not None)
        while len(_MAPS) > 1 and used > MAP_CACHE_BYTES:
            _, old = _MAPS.popitem(last=False)
            used -= 8 * old.total if old.dst is not None else 0
    else:
        _MAPS.move_to_end(key)
    return m


# ---------------------------------------------------------------------------
# Block QR isolating one leg — MatrixAlgebraKit left_orth per sector
# ---------------------------------------------------------------------------


def qr_isolating(A, k=None):
    """Block QR of a device sector tensor that isolates leg k (the last leg when
    k is None): A = Q . R per sector s of that leg, where the matrix of sector
    s stacks every block with s on leg k (rows = the other legs' tuples, in
    sorted order).  Q keeps A's leg order with leg k replaced by the new bond,
    whose sector s has degeneracy min(rows, columns) of its matrix; R :: new
    <- old is a 2-leg device tensor (flat legs: the new bond's dual, leg k).
    Sectors of leg k that no block carries are dropped from the new bond."""
    st = A.blocks
    N = len(st.legs)
    k = N - 1 if k is None else k
    others = [l for l in range(N) if l != k]
    M = _maps(st, others, [k], "qr", lambda key: key[k])
    buf = M.gather(A)
    leg = st.legs[k]
    qs, rs, new = [], [], []
    for s in M.groups:
        Q, R = B.thin_qr(M.matrix(buf, s))
        qs.append(Q)
        rs.append(R)
        new.append((s, Q.shape[1]))
    new_leg = SectorLeg([(leg.sectors[s][0], r) for s, r in new])
    remap = {s: i for i, (s, _) in enumerate(new)}
    qlegs = list(st.legs)
    qlegs[k] = new_leg
    qst = BlockStructure(qlegs, sorted(key[:k] + (remap[key[k]],) + key[k + 1:] for key in st.blocks
                                       if key[k] in remap), _int_tuples=True)
    MQ = _maps(qst, others, [k], "qr", lambda key: key[k])
    qbuf = torch.cat([Q.T.reshape(-1) for Q in qs])
    Q = MQ.scatter(qbuf, qst, A.ncod)
    rst = BlockStructure([new_leg.dual(), leg], sorted((i, s) for s, i in remap.items()),
                         _int_tuples=True)
    order = {b: i for i, b in enumerate(rst.blocks)}
    pieces = [None] * len(rs)
    for s, R in zip(M.groups, rs):
        pieces[order[(remap[s], s)]] = R.T.reshape(-1)
    Rt = DTensor(rst.dims, torch.cat(pieces), 1, rst)
    return Q, Rt


def permute2(T):
    """Transpose of a 2-leg device sector tensor (legs swapped, one codomain leg)."""
    st = T.blocks
    nst = BlockStructure([st.legs[1], st.legs[0]], sorted((b[1], b[0]) for b in st.blocks),
                         _int_tuples=True)
    pieces = []
    for b in nst.blocks:
        i = st.index[(b[1], b[0])]
        r, c = st.block_dims[i]
        o = int(st.offsets[i])
        # the stored (r x c) block read as (c x r) row-major is its transpose;
        # flattening that transpose's transpose row-major stores (c x r) column-major
        pieces.append(T.data[o:o + r * c].view(c, r).T.reshape(-1))
    data = torch.cat(pieces) if pieces else T.data[:0]
    return DTensor(nst.dims, data, 1, nst)


def to_structure(T, st, ncod):
    """A 2-leg-or-more device sector tensor re-laid into structure ``st`` over
    the same legs: its blocks copied, blocks it lacks zero."""
    if list(T.blocks.legs) != list(st.legs):
        raise ValueError("to_structure: the tensor's legs differ from the structure's")
    src = T.blocks
    extra = set(src.blocks) - set(st.blocks)
    for b in extra:
        i = src.index[b]
        o = int(src.offsets[i])
        if bool(torch.any(T.data[o:o + src.sizes[i]] != 0)):
            raise ValueError(f"to_structure: block {b} lies outside the structure")
    pieces = []
    for i, b in enumerate(st.blocks):
        j = src.index.get(b)
        if j is None:
            pieces.append(torch.zeros(st.sizes[i], dtype=torch.complex128, device=T.data.device))
        else:
            o = int(src.offsets[j])
            pieces.append(T.data[o:o + src.sizes[j]])
    data = torch.cat(pieces) if pieces else T.data[:0]
    return DTensor(st.dims, data, ncod, st)


# ---------------------------------------------------------------------------
# Truncated block SVD of a two-site tensor — MatrixAlgebraKit svd_trunc
# ---------------------------------------------------------------------------


def _row_charge_group(st, nrow):
    from .planning import _leg_codes
    codes = [[c for c, _ in _leg_codes(leg)] for leg in st.legs[:nrow]]
    return lambda key: sum(codes[l][x] for l, x in enumerate(key[:nrow]))


def _charge_of(st, key, nrow):
    return tuple(int(v) for v in np.sum([st.legs[l].sectors[x][0] for l, x in enumerate(key[:nrow])],
                                        axis=0))


def svd_split(theta, nrow, trunc_rank, absorb=None, bond_slot=0, vh_ncod=1):
    """Truncated SVD of an all-codomain device sector tensor across legs
    (0..nrow-1 | nrow..): one dense matrix per total charge q of the row legs.
    The singular values of all sectors are ranked together and the largest
    ``trunc_rank(spectrum)`` kept (ties in sector order).  Returns
    (U, S, Vh, spectrum, discarded norm, bond leg): U :: row legs <- bond
    (device), S a dict bond sector -> device singular values, Vh the column
    legs with the bond inserted at ``bond_slot`` (device, ``vh_ncod`` codomain
    legs), spectrum descending (numpy).  ``absorb`` "u" or "vh" multiplies
    the singular values into that factor."""
    st = theta.blocks
    N = len(st.legs)
    grp = _row_charge_group(st, nrow)
    M = _maps(st, list(range(nrow)), list(range(nrow, N)), ("svd", nrow), grp)
    buf = M.gather(theta)
    states, spectra, charges = [], [], []
    for g in M.groups:
        mat = M.matrix(buf, g)
        s, state = B.svd_spectrum(mat)
        states.append((mat, state))
        spectra.append(s)
        rt = M.rows[M.gindex[g]][0]
        charges.append(_charge_of(st, rt + (0,) * (N - nrow), nrow))
    spectrum = np.sort(np.concatenate(spectra))[::-1] if spectra else np.zeros(0)
    keep_total = trunc_rank(spectrum)
    thr = spectrum[keep_total - 1] if keep_total > 0 else np.inf
    order = sorted(range(len(M.groups)), key=lambda i: charges[i])
    kept, bond, parts = 0, [], {}
    for i in order:
        S = spectra[i]
        r = int(np.sum(S > thr))
        r += min(int(np.sum(S == thr)), keep_total - kept - r) if thr > 0 else 0
        r = max(0, min(r, keep_total - kept))
        kept += r
        if r == 0:
            continue
        parts[i] = (len(bond), r)
        bond.append((charges[i], r))
    kept_sq = 0.0
    Ub, Sb, Vb = {}, {}, {}
    for i, (sidx, r) in parts.items():
        mat, state = states[i]
        U, S, Vh = B.svd_kept(mat, state, r)
        if absorb == "u":
            U = U * S.to(U.dtype)[None, :]
        elif absorb == "vh":
            Vh = Vh * S.to(Vh.dtype)[:, None]
        Ub[i], Sb[sidx], Vb[i] = U, S, Vh
        kept_sq += float(np.sum(S.detach().cpu().numpy() ** 2))
    disc = float(np.sqrt(max(0.0, float(np.sum(spectrum ** 2)) - kept_sq)))
    bond_leg = SectorLeg(bond)
    return (_assemble_u(st, nrow, M, parts, Ub, bond_leg), Sb,
            _assemble_vh(st, nrow, M, parts, Vb, bond_leg, bond_slot, vh_ncod), spectrum, disc,
            bond_leg)


def _assemble_u(st, nrow, M, parts, Ub, bond_leg):
    """U :: (row legs) <- bond as a device block tensor (legs: row legs, bond dual)."""
    legs = st.legs[:nrow] + [bond_leg.dual()]
    blocks = []
    for i, (sidx, _) in parts.items():
        for rt in M.rows[i]:
            blocks.append(rt + (sidx,))
    ust = BlockStructure(legs, sorted(blocks), _int_tuples=True)
    MU = _maps(ust, list(range(nrow)), [nrow], "u", lambda key: key[nrow])
    buf = torch.cat([Ub[i].T.reshape(-1) for i, _ in sorted(parts.items(), key=lambda x: x[1][0])])
    return MU.scatter(buf, ust, nrow)


def _assemble_vh(st, nrow, M, parts, Vb, bond_leg, k, ncod):
    """Vh :: bond <- (column legs) as a device block tensor whose legs are the
    column legs with the bond inserted at slot k."""
    N = len(st.legs)
    cl = st.legs[nrow:]
    legs = cl[:k] + [bond_leg] + cl[k:]
    blocks = []
    for i, (sidx, _) in parts.items():
        for ct in M.cols[i]:
            blocks.append(ct[:k] + (sidx,) + ct[k:])
    vst = BlockStructure(legs, sorted(blocks), _int_tuples=True)
    cols = [j for j in range(N - nrow + 1) if j != k]
    MV = _maps(vst, [k], cols, ("vh", k), lambda key: key[k])
    buf = torch.cat([Vb[i].T.reshape(-1) for i, _ in sorted(parts.items(), key=lambda x: x[1][0])])
    return MV.scatter(buf, vst, ncod)
