"""Abelian sector tensors (U(1), U(1) x U(1)) on the host side of the device path.

Device tensors with sectors are ``DTensor``s carrying a ``BlockStructure``
(flat legs as ``SectorLeg`` lists, blocks = tuples of sector indices whose
flat charges sum to zero, packed in block order).  The effective maps and
environments run as grouped block GEMMs in libgraft_heff.so.  This module
holds the gauge factorizations of such tensors (block QR, truncated block
SVD) and their construction; the reference performs these in
MatrixAlgebraKit through TensorKit (Backend.jl:466-609).  The sweeps call the
device forms (``sector_linalg``: per-sector matrices gathered on the device,
factorized there); the host forms here build initial states and serve as the
reference the device forms are tested against.  Absorbing a link factor into
a neighbour is one planned pair contraction on the device.

Charge conventions (TensorKit, flat-leg picture): a codomain leg carries the
charges of its space V, a domain leg those of dual(V) (negated).  A TTNS
tensor A :: (children) (x) P <- V_parent therefore has blocks with
sum(children) + phys - parent = 0, so bond charges count the subtree below.
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np

from .planning import SectorLeg
from .tensors import BlockStructure, DTensor


def make_leg(charge_degs):
    """SectorLeg from an ordered list of (charge tuple, degeneracy), dropping empty sectors."""
    return SectorLeg([(tuple(q), int(d)) for q, d in charge_degs if d > 0])


def invariant_blocks(legs):
    return BlockStructure.invariant(legs).blocks


class HostBlocks:
    """Host copy of a sector tensor: legs, ncod and {sector tuple: ndarray}."""

    def __init__(self, legs, blocks, ncod):
        self.legs = list(legs)
        self.blocks = dict(blocks)
        self.ncod = ncod

    @staticmethod
    def from_device(t):
        bs = t.blocks
        host = t.data.detach().cpu().numpy()
        blocks = {}
        for i, b in enumerate(bs.blocks):
            o = int(bs.offsets[i])
            blocks[b] = host[o:o + bs.sizes[i]].reshape(bs.block_dims[i], order="F").copy()
        return HostBlocks(bs.legs, blocks, t.ncod)

    def to_device(self, complete=True):
        """Upload.  ``complete``: store every charge-conserving block of the legs
        (absent ones as zeros), so that maps and Krylov vectors built on this
        tensor share one block structure."""
        blocks = dict(self.blocks)
        if complete:
            for key in invariant_blocks(self.legs):
                if key not in blocks:
                    blocks[key] = np.zeros(self.dims_of(key), dtype=np.complex128)
        keys = sorted(blocks)
        struct = BlockStructure(self.legs, keys)
        return DTensor.from_blocks(struct, [blocks[k] for k in keys], self.ncod)

    def dims_of(self, key):
        return tuple(self.legs[l].sectors[s][1] for l, s in enumerate(key))

    def dense(self):
        out = np.zeros(tuple(l.dim for l in self.legs), dtype=np.complex128)
        starts = [np.concatenate([[0], np.cumsum(l.degs())[:-1]]) for l in self.legs]
        for key, blk in self.blocks.items():
            sl = tuple(slice(int(starts[l][s]), int(starts[l][s]) + blk.shape[l]) for l, s in enumerate(key))
            out[sl] = blk
        return out

    def permute(self, perm, ncod=None):
        return HostBlocks([self.legs[p] for p in perm],
                          {tuple(k[p] for p in perm): np.transpose(v, perm) for k, v in self.blocks.items()},
                          self.ncod if ncod is None else ncod)


def random_blocks(rng, legs, ncod):
    """Invariant sector tensor with i.i.d. complex normal entries."""
    blocks = {}
    for key in invariant_blocks(legs):
        shape = tuple(legs[l].sectors[s][1] for l, s in enumerate(key))
        blocks[key] = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)
    return HostBlocks(legs, blocks, ncod)


def _row_index(T, nrow):
    """Group blocks by the sector tuple of legs nrow.. (the column part)."""
    groups = {}
    for key in T.blocks:
        groups.setdefault(key[nrow:], []).append(key[:nrow])
    return groups


def qr_left(T):
    """T (legs 0..N-2 | last leg) = Q . R per sector of the last leg.

    Returns (Q, R) with Q's last leg the new bond (same charges as T's last
    leg, reduced degeneracies) and R :: new <- old as a 2-leg tensor."""
    N = len(T.legs)
    last = T.legs[-1]
    groups = _row_index(T, N - 1)
    qblocks, rblocks, new_degs = {}, {}, {}
    for col, rows in groups.items():
        s = col[0]
        rows = sorted(rows)
        mats = [T.blocks[r + col].reshape(-1, last.sectors[s][1], order="F") for r in rows]
        M = np.vstack(mats)
        Q, R = np.linalg.qr(M, mode="reduced")
        new_degs[s] = Q.shape[1]
        off = 0
        for r, m in zip(rows, mats):
            qblocks[(r, s)] = Q[off:off + m.shape[0]]
            off += m.shape[0]
        rblocks[s] = R
    order = [s for s in range(len(last.sectors)) if new_degs.get(s, 0) > 0]
    remap = {s: i for i, s in enumerate(order)}
    # the new leg has T's last-leg sectors (flat, i.e. dual charges) with reduced degeneracy
    new_leg_flat = SectorLeg([(last.sectors[s][0], new_degs[s]) for s in order])
    Qb = {}
    for (r, s), q in qblocks.items():
        if s in remap:
            shape = T.dims_of(r + (s,))[:-1] + (q.shape[1],)
            Qb[r + (remap[s],)] = q.reshape(shape, order="F")
    Q = HostBlocks(T.legs[:-1] + [new_leg_flat], Qb, T.ncod)
    # R :: V_new <- V_old; flat legs (dual of new_leg_flat, last)
    Rb = {(remap[s], s): rblocks[s] for s in order}
    R = HostBlocks([new_leg_flat.dual(), last], Rb, 1)
    return Q, R


def contract_leg(T, k, C, ck, cnew):
    """Contract leg k of T with leg ck of the 2-leg tensor C; C's other leg cnew replaces leg k."""
    out = {}
    for key, blk in T.blocks.items():
        for ckey, cblk in C.blocks.items():
            if ckey[ck] != key[k]:
                continue
            m = np.tensordot(blk, cblk, axes=([k], [ck]))      # new leg last
            m = np.moveaxis(m, -1, k)
            nk = key[:k] + (ckey[cnew],) + key[k + 1:]
            if nk in out:
                out[nk] = out[nk] + m
            else:
                out[nk] = m
    legs = list(T.legs)
    legs[k] = C.legs[cnew]
    return HostBlocks(legs, out, T.ncod)


def svd_split(theta, nrow, trunc_rank):
    """Truncated SVD of an all-codomain sector tensor across legs (0..nrow-1 | nrow..).

    Singular values of all coupled sectors are ranked together and the
    largest ``trunc_rank(spectrum)`` kept.  Returns (U, S dict, Vh, spectrum,
    discarded norm, bond SectorLeg) with U's last leg and Vh's first leg the
    new bond (charges of the row part's total for U)."""
    from .planning import _leg_codes
    codes = [[c for c, _ in _leg_codes(leg)] for leg in theta.legs[:nrow]]
    rowsets, colsets, charge_of = {}, {}, {}
    for key in theta.blocks:
        r, c = key[:nrow], key[nrow:]
        code = sum(codes[l][x] for l, x in enumerate(r))       # integer charge code of the row part
        q = charge_of.get(code)
        if q is None:
            q = charge_of[code] = tuple(np.sum([theta.legs[l].sectors[x][0] for l, x in enumerate(r)],
                                               axis=0))
        rowsets.setdefault(q, set()).add(r)
        colsets.setdefault(q, set()).add(c)
    pieces = []
    for q in sorted(rowsets):
        rows = sorted(rowsets[q])
        cols = sorted(colsets[q])
        rdims = [int(np.prod(theta.dims_of(r + cols[0])[:nrow])) for r in rows]
        cdims = [int(np.prod(theta.dims_of(rows[0] + c)[nrow:])) for c in cols]
        M = np.zeros((sum(rdims), sum(cdims)), dtype=np.complex128)
        ro = np.concatenate([[0], np.cumsum(rdims)])
        co = np.concatenate([[0], np.cumsum(cdims)])
        for i, r in enumerate(rows):
            for j, c in enumerate(cols):
                b = theta.blocks.get(r + c)
                if b is not None:
                    M[ro[i]:ro[i + 1], co[j]:co[j + 1]] = b.reshape(rdims[i], cdims[j], order="F")
        U, S, Vh = np.linalg.svd(M, full_matrices=False)
        pieces.append((q, rows, cols, rdims, cdims, ro, co, U, S, Vh))
    spectrum = np.sort(np.concatenate([p[8] for p in pieces]))[::-1]
    keep_total = trunc_rank(spectrum)
    thr = spectrum[keep_total - 1] if keep_total > 0 else np.inf
    # keep the keep_total largest values (ties broken by sector order)
    kept = 0
    bond, Ub, Sb, Vb = [], {}, {}, {}
    for q, rows, cols, rdims, cdims, ro, co, U, S, Vh in pieces:
        r = int(np.sum(S > thr))
        r += min(int(np.sum(S == thr)), keep_total - kept - r) if thr > 0 else 0
        r = max(0, min(r, keep_total - kept))
        kept += r
        if r == 0:
            continue
        sidx = len(bond)
        bond.append((q, r))
        for i, rw in enumerate(rows):
            Ub[rw + (sidx,)] = U[ro[i]:ro[i + 1], :r]
        for j, c in enumerate(cols):
            Vb[(sidx,) + c] = Vh[:r, co[j]:co[j + 1]]
        Sb[sidx] = S[:r]
    disc = float(np.sqrt(max(0.0, np.sum(spectrum ** 2) - sum(np.sum(x ** 2) for x in Sb.values()))))
    bond_leg = SectorLeg(bond)                       # charges of the row part, +q
    Ublocks = {}
    for k, v in Ub.items():
        shape = tuple(theta.legs[l].sectors[x][1] for l, x in enumerate(k[:-1])) + (v.shape[1],)
        Ublocks[k] = v.reshape(shape, order="F")
    Vblocks = {}
    for k, v in Vb.items():
        shape = (v.shape[0],) + tuple(theta.legs[nrow + l].sectors[x][1] for l, x in enumerate(k[1:]))
        Vblocks[k] = v.reshape(shape, order="F")
    U = HostBlocks(theta.legs[:nrow] + [bond_leg.dual()], Ublocks, nrow)
    Vh = HostBlocks([bond_leg] + theta.legs[nrow:], Vblocks, 1)
    return U, Sb, Vh, spectrum, disc, bond_leg


def scale_bond(T, S, leg):
    """Multiply T's leg ``leg`` (the bond) by the singular values S[sector]."""
    out = {}
    for k, v in T.blocks.items():
        sv = S[k[leg]]
        shape = [1] * v.ndim
        shape[leg] = len(sv)
        out[k] = v * sv.reshape(shape)
    return HostBlocks(T.legs, out, T.ncod)


# ---------------------------------------------------------------------------
# Gauge moves and two-site split for sector TTNS (ttns.jl:181-203,
# effective.jl:827-876), executed per charge sector on the host.
# ---------------------------------------------------------------------------


def move_center_edge_(psi, m, cache):
    from .trees import childslot
    t = psi.topo
    n = psi.center
    if t.parent[n] == m:
        Q, R = link_split(psi.tensors[n])                   # R :: new <- old
        psi.tensors[n] = Q
        psi.tensors[m] = absorb_link(psi.tensors[m], childslot(t, m, n), R, 1, 0)
    else:
        Q, R = link_split(psi.tensors[n], childslot(t, n, m))
        psi.tensors[n] = Q
        P = psi.tensors[m]
        psi.tensors[m] = absorb_link(P, P.rank - 1, R, 1, 0)
    psi.center = m
    if cache is not None:
        cache.invalidate_edge_(n, m)
    return psi


def link_split(A, k=None):
    """Block QR of a device sector tensor that isolates leg k (the last leg when
    k is None): A = Q . R per sector of that leg, Q isometric from the other
    legs onto a new bond in slot k, R :: new <- old (flat legs: the new bond's
    dual, leg k), both on the device — the QR link splits of tdvp.jl:291-301
    and Backend.jl:597-609.  ``qr_isolating`` below is the host form."""
    from .sector_linalg import qr_isolating as device_qr
    return device_qr(A, k)


def qr_isolating(T, k=None):
    """Host block QR isolating leg k of T (the last leg when k is None): Q keeps
    T's leg order with leg k replaced by the new bond (degeneracy per sector =
    the rank bound min(rows, columns) of that sector's matrix), R :: new <- old."""
    N = len(T.legs)
    if k is None or k == N - 1:
        return qr_left(T)
    perm = [j for j in range(N) if j != k] + [k]
    Q, R = qr_left(T.permute(perm))
    back = [0] * N
    for pos, j in enumerate(perm):
        back[j] = pos
    return Q.permute(back, T.ncod), R


def to_structure(T, struct, ncod):
    """Upload host blocks into a given block structure whose legs are T's legs;
    blocks of the structure that T lacks are stored as zeros."""
    if list(T.legs) != list(struct.legs):
        raise ValueError("to_structure: the tensor's legs differ from the structure's")
    keys = set(struct.blocks)
    for key, blk in T.blocks.items():
        if key not in keys and np.any(blk):
            raise ValueError(f"to_structure: block {key} lies outside the structure")
    data = [T.blocks[b] if b in T.blocks else np.zeros(T.dims_of(b), dtype=np.complex128)
            for b in struct.blocks]
    return DTensor.from_blocks(struct, data, ncod)


_ABSORB_PLANS = OrderedDict()      # pair plans of link absorbs and two-site tensors
ABSORB_PLAN_CACHE = 256


def _remember_plan(key, entry):
    _ABSORB_PLANS[key] = entry
    if len(_ABSORB_PLANS) > ABSORB_PLAN_CACHE:
        _ABSORB_PLANS.popitem(last=False)


def absorb_link(P, k, C, ck, cnew):
    """Leg k of the device sector tensor P contracted with leg ck of the 2-leg
    device tensor C, whose leg cnew takes its place: one planned pair
    contraction on the device (grouped block GEMMs), written into every
    charge-conserving block of the new legs.  Plans are kept per operand
    structure (least recently used first out)."""
    from .contractions import NativePlan
    from .planning import ContractionPlan, PairStep
    if C.rank != 2 or {ck, cnew} != {0, 1}:
        raise ValueError("absorb_link: C must be a 2-leg tensor and (ck, cnew) its two legs")
    if C.blocks.legs[ck] != P.blocks.legs[k].dual():
        raise ValueError("absorb_link: the contracted legs are not dual spaces")
    key = (P.blocks.key(), C.blocks.key(), k, ck)
    entry = _ABSORB_PLANS.get(key)
    if entry is None:
        N = P.rank
        legs = list(P.blocks.legs)
        legs[k] = C.blocks.legs[cnew]
        out = BlockStructure.invariant(legs)
        perm = tuple(i if i < k else (N - 1 if i == k else i - 1) for i in range(N))
        step = PairStep(0, 1, 2, tuple(j for j in range(N) if j != k), (k,), (cnew,), (ck,),
                        False, False, (perm, ()))
        entry = (NativePlan(ContractionPlan(3, 2, [step]), [P, C], out), out)
        _remember_plan(key, entry)
    else:
        _ABSORB_PLANS.move_to_end(key)
    plan, out = entry
    res = DTensor.empty(out.dims, P.ncod, out)
    plan.execute([P, C], res)
    return res


def two_site_host(A, Bm, k):
    """Theta = A (parent leg) x Bm (child slot k) on the full two-site space.

    The result carries every charge-conserving block of its legs
    (two_site_space, effective.jl:808-843): blocks that the product leaves
    empty are stored as zeros, because H_eff and the Krylov vectors reach them."""
    na = len(A.legs)
    out = {}
    for ka, va in A.blocks.items():
        for kb, vb in Bm.blocks.items():
            if kb[k] != ka[-1]:
                continue
            key = ka[:-1] + kb[:k] + kb[k + 1:]
            v = np.tensordot(va, vb, axes=([na - 1], [k]))
            out[key] = out[key] + v if key in out else v
    legs = A.legs[:-1] + Bm.legs[:k] + Bm.legs[k + 1:]
    full = {}
    for key in invariant_blocks(legs):
        blk = out.pop(key, None)
        full[key] = blk if blk is not None else np.zeros(
            tuple(legs[l].sectors[s][1] for l, s in enumerate(key)), dtype=np.complex128)
    if out:
        raise ValueError("two-site product produced a block outside the charge-conserving space")
    return HostBlocks(legs, full, len(legs))


def two_site_tensor(psi, n, m):
    """Theta on the device: one planned pair contraction of A_n and A_m over
    their bond, written into every charge-conserving block of Theta's legs
    (blocks the product leaves empty come out as zeros)."""
    from .contractions import NativePlan
    from .planning import ContractionPlan, PairStep
    from .tensors import BlockStructure
    from .trees import childslot
    k = childslot(psi.topo, m, n)
    A, Bm = psi.tensors[n], psi.tensors[m]
    key = ("theta", A.blocks.key(), Bm.blocks.key(), k)
    entry = _ABSORB_PLANS.get(key)
    if entry is None:
        pn = A.rank - 1
        bopen = tuple(j for j in range(Bm.rank) if j != k)
        legs = A.blocks.legs[:-1] + [Bm.blocks.legs[j] for j in bopen]
        out = BlockStructure.invariant(legs)
        step = PairStep(0, 1, 2, tuple(range(pn)), (pn,), bopen, (k,), False, False,
                        (tuple(range(len(legs))), ()))
        entry = (NativePlan(ContractionPlan(3, 2, [step]), [A, Bm], out), out)
        _remember_plan(key, entry)
    else:
        _ABSORB_PLANS.move_to_end(key)
    plan, out = entry
    theta = DTensor.empty(out.dims, len(out.legs), out)
    plan.execute([A, Bm], theta)
    return theta


def split_two_site_(psi, theta, n, m, trunc_rank, center_on="n"):
    """split_two_site! (effective.jl:852-876) for sector states: the truncated
    block SVD runs on the device per row charge; the singular values go to the
    new centre and the bond takes child slot k of A_m."""
    from .sector_linalg import svd_split as device_svd
    from .trees import childslot
    k = childslot(psi.topo, m, n)
    pn = psi.tensors[n].ncod
    Km = psi.tensors[m].ncod
    U, S, Am, spectrum, disc, bond = device_svd(theta, pn, trunc_rank,
                                                 absorb="u" if center_on == "n" else "vh",
                                                 bond_slot=k, vh_ncod=Km)
    psi.tensors[n] = U
    psi.tensors[m] = Am
    psi.center = n if center_on == "n" else m
    return psi, dict(discarded=disc, spectrum=spectrum, kept=int(sum(d for _, d in bond.sectors)))


# ---------------------------------------------------------------------------
# Random symmetric states with a Gaussian spread of bond sectors.
# ---------------------------------------------------------------------------


def _subtree_sizes(topo):
    from .trees import postorder
    size = [1] * len(topo.ids)
    for n in postorder(topo):
        for c in topo.children[n]:
            size[n] += size[c]
    return size


def _hubbard_multiplicity(s, N, sz2):
    from math import comb
    if (N + sz2) % 2:
        return 0
    nu, nd = (N + sz2) // 2, (N - sz2) // 2
    if nu < 0 or nd < 0 or nu > s or nd > s:
        return 0
    return comb(s, nu) * comb(s, nd)


def hubbard_bond_leg(s, n, D, sigma, n_total, sz2_total=0):
    """Bond sectors (N, 2 S_z) below a subtree of s of the n sites: a discretized
    Gaussian around the mean filling, capped by both sides' state counts."""
    mean = s * n_total / n
    cand = []
    for N in range(0, 2 * s + 1):
        for sz2 in range(-2 * s, 2 * s + 1):
            m_in = _hubbard_multiplicity(s, N, sz2)
            m_out = _hubbard_multiplicity(n - s, n_total - N, sz2_total - sz2)
            if m_in and m_out:
                w = np.exp(-((N - mean) ** 2 + sz2 ** 2) / (2 * sigma ** 2))
                cand.append(((N, sz2), w, min(m_in, m_out)))
    tot = sum(w for _, w, _ in cand)
    sectors = []
    for q, w, cap in sorted(cand):
        d = min(cap, int(round(D * w / tot)))
        if d > 0:
            sectors.append((q, d))
    return SectorLeg(sectors)


def random_hubbard_ttns(rng, topo, D, sigma=None, n_total=None, sz2_total=None):
    """Random canonical U(1) x U(1) TTNS on the device (centre at the root), normalized."""
    from .networks import TTNS
    host = random_hubbard_ttns_host(rng, topo, D, sigma, n_total, sz2_total)
    return TTNS(topo, [h.to_device() for h in host], topo.root, [True] * len(topo.ids)), host


def random_hubbard_ttns_host(rng, topo, D, sigma=None, n_total=None, sz2_total=None):
    """Host blocks of a random canonical U(1) x U(1) TTNS (centre at the root).

    Defaults: half filling (N = number of sites) and the lowest |2 S_z|."""
    from .operators import hubbard_phys_leg
    from .trees import childslot, postorder
    n = len(topo.ids)
    n_total = n if n_total is None else n_total
    sz2_total = n_total % 2 if sz2_total is None else sz2_total
    sigma = max(1.0, np.sqrt(n) / 1.5) if sigma is None else sigma
    size = _subtree_sizes(topo)
    ph = hubbard_phys_leg()
    bond = {c: hubbard_bond_leg(size[c], n, D, sigma, n_total, sz2_total)
            for c in range(n) if topo.parent[c] >= 0}
    host = []
    for v in range(n):
        up = SectorLeg([((n_total, sz2_total), 1)]) if topo.parent[v] < 0 else bond[v]
        legs = [bond[c] for c in topo.children[v]] + [ph, up.dual()]
        host.append(random_blocks(rng, legs, len(legs) - 1))
    for v in postorder(topo):
        if topo.parent[v] < 0:
            continue
        Q, R = qr_left(host[v])
        host[v] = Q
        p = topo.parent[v]
        host[p] = contract_leg(host[p], childslot(topo, p, v), R, 1, 0)
    r = topo.root
    nrm = np.sqrt(sum(np.sum(np.abs(b) ** 2) for b in host[r].blocks.values()))
    host[r] = HostBlocks(host[r].legs, {k: b / nrm for k, b in host[r].blocks.items()}, host[r].ncod)
    return host
