"""Time evolution — lib/GraftEvolution/src/Evolution/tdvp.jl (TDVP1).

One-site TDVP with the tree projector-splitting sweep (tdvp.jl:197-251):
every node is forward-evolved once per sweep with the one-site map, and an
edge is backward-evolved with the zero-site (link) map exactly when the
centre crosses it child -> parent (parent -> child on the reverse sweep).
Krylov exponentials go through ``evolution_exponentiate_backend``
(Evolution.jl:134-148), here the device Lanczos of ``krylov.exponentiate``.
States may be dense or U(1) x U(1) block-sparse; for the latter the link
splits are per-sector block QRs and the link map carries the bond's block
structure.  Controlled bond expansion (TDVP1_CBE) selects its directions
from the factor frame of the edge, for dense and for block-sparse states.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from . import backend as B
from . import krylov
from . import sector_linalg as SL
from . import sectors as S
from . import tensors as T
from . import contractions as C
from .contractions import EnvCache, eff_h0, eff_h1
from .networks import move_center_, update_tensor_
from .trees import childslot, path_between, postorder


def evolution_exponentiate_backend(effective, dz, x, **kw):
    """Hook point of Evolution.jl:134-135: exp(dz * H_eff) x."""
    return krylov.exponentiate(effective, dz, x, **kw)


@dataclass
class TDVP1:
    """TDVP1(; order=2, krylovdim=30, tol=1e-12, eager=true) — tdvp.jl:31-45."""
    order: int = 2
    krylovdim: int = 30
    tol: float = 1e-12
    eager: bool = True
    precision: str = "fp64"
    cache: object = None
    tune_tiles: bool = False     # autotuned tiles for heavy sector maps, per step shape class


@dataclass
class PredictorCBE:
    """PredictorCBE(; max_add=32, spawn_threshold=1e-10, neigs=2, krylovdim=4) —
    cbe.jl:115-149: projected two-site energy-variation selector.  Directions are
    the leading left singular vectors of the double-complement core
    M = (1 - Q Q^H) . H_eff(Theta) . (1 - P_partner), held in factor-frame form;
    at most min(room, max_add, neigs) are spawned, each only if
    |dz| * sigma / ||Theta|| >= spawn_threshold.  Dense cores: block Krylov with
    Rayleigh-Ritz to a residual of 1e-10 sigma_1 (the reference: a Golub-Kahan
    svdsolve of dimension ``krylovdim``); small or unconverged cores and sector
    cores: exact SVD, as the reference's fallback."""
    max_add: int = 32
    spawn_threshold: float = 1e-10
    neigs: int = 2
    krylovdim: int = 4

    def __post_init__(self):
        if self.max_add < 0:
            raise ValueError("PredictorCBE max_add must be nonnegative")
        if self.spawn_threshold < 0:
            raise ValueError("PredictorCBE spawn_threshold must be nonnegative")
        if self.neigs < 1:
            raise ValueError("PredictorCBE neigs must be positive")
        if self.krylovdim < self.neigs:
            raise ValueError("PredictorCBE krylovdim must be at least neigs")


@dataclass
class TDVP1_CBE:
    """TDVP1_CBE (cbe.jl:240-261): TDVP1 whose link splits spawn new bond
    directions (controlled bond expansion) before the backward link evolution.
    ``enabled=False`` is TDVP1 exactly."""
    cbe: PredictorCBE = field(default_factory=PredictorCBE)
    order: int = 2
    trunc: B.TruncationScheme = field(default_factory=lambda: B.TruncationScheme(maxdim=100))
    enabled: bool = True
    krylovdim: int = 30
    tol: float = 1e-12
    eager: bool = True
    precision: str = "fp64"
    cache: object = None
    last_cbe_info: dict = None
    tune_tiles: bool = False     # autotuned tiles for heavy sector maps, per step shape class


CBE_BLOCK_EXTRA = 4      # Krylov block size = budget + this
CBE_MAX_BLOCKS = 16      # block Krylov steps before the exact fallback
CBE_RESIDUAL = 1e-10     # converged when every kept Ritz residual <= this * sigma_1


class _ProjectedCore:
    """The double-complement core Z = (1 - Q Q^H) . L R^T . (1 - conj(P) P^T)
    of edge u -> v in factor-frame form (cbe.jl:294-327): L and R are the two
    site factor actions (external legs x (bond, operator channel)), so
    L R^T is the two-site map applied to Theta = A_u A_v without forming it;
    Q spans the active tensor's range, P is the partner (canonical, isometric).
    Only products with a few vectors are taken."""

    def __init__(self, L, R, Q, P):
        self.L, self.R, self.Q, self.P = L, R, Q, P
        self.shape = (L.shape[0], R.shape[0])

    def _pu(self, Y):
        return Y - B.matmul(self.Q, B.matmul(self.Q, Y, "H"))

    def _pv(self, X):
        return X - B.matmul(self.P, B.matmul(self.P, X, "T"), "C")

    def mul(self, X):
        """Z X for X (cols of v's side x p)."""
        return self._pu(B.matmul(self.L, B.matmul(self.R, self._pv(X), "T")))

    def rmul(self, Y):
        """Z^H Y for Y (rows of u's side x p)."""
        return self._pv(B.matmul(self.R, B.matmul(self.L, self._pu(Y), "H"), "C"))

    def dense(self):
        """Z itself, formed from its smaller side (Z = (Z^H I)^H when it has fewer rows)."""
        rows, cols = self.shape
        if rows < cols:
            eye = torch.eye(rows, dtype=self.L.dtype, device=self.L.device)
            return self.rmul(eye).conj().T.resolve_conj()
        eye = torch.eye(cols, dtype=self.L.dtype, device=self.L.device)
        return self.mul(eye)


def _orth_against(Kb, G, rel=1e-10):
    """Orthonormal columns spanning G minus its components along the blocks Kb
    (orthonormal), or None when nothing of G is left.  Columns that lose all
    but ``rel`` of their norm to the projection are dropped (the Krylov space
    has become invariant in those directions); the survivors are projected
    again after normalization, so the new block is orthogonal to Kb to
    working precision."""
    def project(X):
        for _ in range(2):
            for K in Kb:
                X = X - B.matmul(K, B.matmul(K, X, "H"))
        return X
    g0 = torch.linalg.vector_norm(G, dim=0)
    G = project(G)
    keep = (torch.linalg.vector_norm(G, dim=0) > rel * g0).nonzero().flatten()
    if keep.numel() == 0:
        return None
    Qg, _ = torch.linalg.qr(G[:, keep], mode="reduced")
    Qg, _ = torch.linalg.qr(project(Qg), mode="reduced")
    return Qg


SMALL_SIDE = 4096        # a core side up to this size gets an explicit complement basis


def _complement(M):
    """Orthonormal basis of the orthogonal complement of the isometry M's columns."""
    Qf, _ = torch.linalg.qr(M, mode="complete")
    return Qf[:, M.shape[1]:]


def _top_from_gram(X, budget, Qproj=None):
    """Leading singular values of (1 - Qproj Qproj^H) X and the matching
    eigenvectors of its Gram matrix X^H (1 - Qproj Qproj^H) X (its right
    singular vectors), from a Hermitian eigendecomposition on the small side."""
    G = B.matmul(X, X, "H")
    if Qproj is not None:
        QX = B.matmul(Qproj, X, "H")
        G = G - B.matmul(QX, QX, "H")
    G = (G + G.conj().T) / 2
    scale = float(torch.diagonal(G).real.max())
    scale = scale if scale > 0 else 1.0
    lam, V = torch.linalg.eigh(G / scale)
    lam = torch.flip(lam, (0,)).clamp(min=0.0) * scale
    V = torch.flip(V, (1,))[:, :budget]
    sig = torch.sqrt(lam[:budget])
    return sig, V


def _leading_left_vectors(Z, budget, solver=None):
    """Leading ``budget`` singular values and left vectors of the projected core
    Z = (1 - Q Q^H) L R^T (1 - conj(P) P^T).  Returns (sigmas (numpy,
    descending), U (rows x budget) or None, solver name).

    "gram_v" / "gram_u" (default when one side has at most SMALL_SIDE
    entries): an explicit basis N of that side's complement reduces Z to a
    matrix with rank(N) columns (v side: (1 - QQ^H) L R^T conj(N_p)) or rows
    (u side: N_u^H L R'^T), whose small-side Gram matrix is diagonalized; the
    big side is touched by one GEMM with rank(N) columns.  "block_krylov"
    (both sides large): block Krylov on Z Z^H with a Rayleigh-Ritz step per
    block from a fixed-seed start (the reference runs KrylovKit's svdsolve
    from a fixed random probe, cbe.jl:491-528) until every wanted Ritz pair
    has residual ||Z v - sigma u|| <= CBE_RESIDUAL * sigma_1; "exact_svd" the
    SVD of Z formed from its smaller side (the fallback)."""
    rows, cols = Z.shape
    if solver is None:
        solver = ("gram_v" if cols <= min(rows, SMALL_SIDE) else
                  "gram_u" if rows <= SMALL_SIDE else "block_krylov")
    if solver == "gram_v":
        Np = _complement(Z.P)                               # ext_v x (cols - r_P)
        if Np.shape[1] == 0:
            return np.zeros(0), None, solver
        X = B.matmul(Z.L, B.matmul(Z.R, Np, "T", "C"))       # L R^T conj(Np)
        sig, V = _top_from_gram(X, budget, Qproj=Z.Q)
        U = Z._pu(B.matmul(X, V)) / sig.clamp(min=1e-300).to(X.dtype)[None, :]
        return sig.detach().cpu().numpy(), U, solver
    if solver == "gram_u":
        Nu = _complement(Z.Q)                               # ext_u x (rows - r_Q)
        if Nu.shape[1] == 0:
            return np.zeros(0), None, solver
        Y = B.matmul(Z.R, B.matmul(Z.L, Nu, "T", "C"))       # R L^T conj(Nu): cols x m_u
        Y = Y - B.matmul(Z.P, B.matmul(Z.P, Y, "H"))         # R' = (1 - P P^H) R
        sig, V = _top_from_gram(Y, budget)                   # W W^H = conj(Y^H Y), W = Y^T
        return sig.detach().cpu().numpy(), B.matmul(Nu, V.conj()), solver
    if solver == "block_krylov":
        p = budget + CBE_BLOCK_EXTRA
        g = torch.Generator(device=Z.L.device)
        g.manual_seed(0xCBE25031)
        Om = torch.randn((cols, p), dtype=Z.L.dtype, device=Z.L.device, generator=g)
        K0 = _orth_against([], Z.mul(Om))
        if K0 is None:                                     # Z vanishes
            return np.zeros(0), None, "block_krylov(0)"
        Kb, Wb = [K0], [Z.rmul(K0)]
        for it in range(CBE_MAX_BLOCKS):
            K = torch.cat(Kb, dim=1)
            Wt = torch.cat(Wb, dim=1)                      # Z^H K
            U1, S, V1h = torch.linalg.svd(Wt, full_matrices=False)
            sig = S.detach().cpu().numpy()
            Ub = B.matmul(K, V1h[:budget, :], "N", "H")    # left Ritz vectors
            Vb = U1[:, :budget]
            res = Z.mul(Vb) - Ub * S[:budget].to(Ub.dtype)[None, :]
            rnorm = torch.linalg.vector_norm(res, dim=0).detach().cpu().numpy()
            if sig[0] == 0.0 or float(rnorm.max()) <= CBE_RESIDUAL * sig[0]:
                return sig[:budget], Ub, f"block_krylov({it + 1})"
            if K.shape[1] + p > min(rows, cols):
                break
            Kn = _orth_against(Kb, Z.mul(Wb[-1]))
            if Kn is None:
                break
            Kb.append(Kn)
            Wb.append(Z.rmul(Kn))
    Zd = Z.dense()
    s, state = B.svd_spectrum(Zd)
    U, _, _ = B.svd_kept(Zd, state, min(budget, len(s)))
    return s[:budget], U, "exact_svd"


def _cbe_directions(ev, psi, H, n, m, active_is_n, Q, dz, room, solver=None):
    """cbe.jl:332-352, 459-528, 590-619: new orthonormal directions for the
    active site, orthogonal to its current range Q, or None.  The projected
    core is held in factor-frame form (two site factor actions, never the
    two-site map applied to Theta); at most min(room, max_add, neigs)
    directions are spawned, each only if sigma >= spawn_threshold * ||Theta||
    / |dz| (trunctol's >=, cbe.jl:590-619)."""
    strat = ev.cbe
    budget = min(room, strat.max_add, strat.neigs)
    info = {"selected": 0, "budget": budget}
    ev.last_cbe_info = info
    if budget <= 0 or dz == 0:
        return None
    t = psi.topo
    u, v = (n, m) if active_is_n else (m, n)
    L = C.site_factor_action(ev.cache, psi, H, u, v)
    R = C.site_factor_action(ev.cache, psi, H, v, u)
    Pt = psi.tensors[v]
    kv = Pt.rank - 1 if t.parent[v] == u else childslot(t, v, u)
    if kv != Pt.rank - 1:
        Pt = B.permute(Pt, tuple(j for j in range(Pt.rank) if j != kv) + (kv,))
    Z = _ProjectedCore(B.matrix_view(L, L.rank - 2), B.matrix_view(R, R.rank - 2), Q,
                       B.matrix_view(Pt, Pt.rank - 1))
    budget = min(budget, *Z.shape)
    if budget <= 0:
        return None
    norm_theta = T.norm(psi.tensors[u])          # the active site is the orthogonality centre
    cutoff = strat.spawn_threshold * max(norm_theta, np.finfo(float).eps) / abs(complex(dz))
    sig, U, solver = _leading_left_vectors(Z, budget, solver)
    info["solver"] = solver
    info["sigma_max"] = float(sig[0]) if len(sig) else 0.0
    sel = [i for i in range(min(len(sig), budget)) if sig[i] >= cutoff]
    if not sel or U is None:
        return None
    E = U[:, sel]
    # exact orthogonality to the kept range, then orthonormal columns
    E = E - B.matmul(Q, B.matmul(Q, E, "H"))
    E, _ = torch.linalg.qr(E, mode="reduced")
    info["selected"] = len(sel)
    return E


def _cbe_split(ev, psi, H, u, v, dz):
    """_cbe_projected_split + _cbe_state_preserving_splice (cbe.jl:575-584, 743-760):
    A_u = U . R with U = [Q, E] and R = U^H A_u = [R_qr; 0]; the state is unchanged.
    Returns (U as the site tensor in A_u's leg order, R :: new bond <- old bond)."""
    t = psi.topo
    A = psi.tensors[u]
    if A.blocks is not None:
        return _cbe_split_sector(ev, psi, H, u, v, dz)
    up = t.parent[u] == v
    if up:
        k = A.rank - 1
        At = A
    else:
        k = childslot(t, u, v)
        At = B.permute(A, tuple(j for j in range(A.rank) if j != k) + (k,))
    Ma = B.matrix_view(At, At.rank - 1)
    Q, _ = B.thin_qr(Ma)
    room = max(0, ev.trunc.maxdim - Q.shape[1])
    n, m = (u, v) if up else (v, u)
    E = _cbe_directions(ev, psi, H, n, m, up, Q, dz, room)
    U = Q if E is None else torch.cat([Q, E], dim=1)
    R = B.matmul(U, Ma, "H")
    Ut = B.from_matrix(U, At.dims[:-1] + (U.shape[1],), At.rank - 1)
    if not up:
        Ut = B.permute(Ut, B._restore_perm(A.rank, k), A.ncod)
    Rt = B.from_matrix(R, (R.shape[0], R.shape[1]), 1)
    return Ut, Rt


def _row_code_group(legs, rows):
    """Group key of a block: the integer charge code of its tuple on ``rows``."""
    from .planning import _leg_codes
    codes = [[c for c, _ in _leg_codes(legs[l])] for l in rows]
    return lambda key: sum(codes[i][key[l]] for i, l in enumerate(rows))


def _cbe_split_sector(ev, psi, H, u, v, dz):
    """_cbe_split for U(1) x U(1) states: the reference's exact sector-aware
    selector for graded maps (cbe.jl:459-463, 497-498).  The double-complement
    core Z = (1 - Q Q^H) L R^T (1 - conj(P) P^T) is block diagonal in the
    charge q of the active site's external legs; each block is formed from the
    two site factor actions by one planned pair contraction (never the
    two-site map applied to Theta), projected per charge, and factorized
    exactly.  The singular values of all charges are ranked together; the
    leading min(room, max_add, neigs) with sigma >= cutoff are spawned, in
    existing bond sectors or in new ones of charge q.  U = [Q, E] per charge,
    R = U^H A_u; the state is unchanged."""
    from .contractions import NativePlan
    from .planning import ContractionPlan, PairStep
    from .planning import SectorLeg
    from .tensors import BlockStructure, DTensor
    t = psi.topo
    A = psi.tensors[u]
    st = A.blocks
    N = A.rank
    k = N - 1 if t.parent[u] == v else childslot(t, u, v)
    ext = [l for l in range(N) if l != k]
    Q, _ = SL.qr_isolating(A, k)
    strat = ev.cbe
    room = max(0, ev.trunc.maxdim - Q.blocks.legs[k].dim)
    budget = min(room, strat.max_add, strat.neigs)
    info = {"selected": 0, "budget": budget, "solver": "exact_svd_sector"}
    ev.last_cbe_info = info
    grp_u = _row_code_group(st.legs, ext)
    sel = {}
    newrows = {}                  # rows of charges the old bond does not carry
    if budget > 0 and dz != 0:
        L = C.site_factor_action(ev.cache, psi, H, u, v)          # ext_u..., a, o
        R = C.site_factor_action(ev.cache, psi, H, v, u)          # ext_v..., a', o'
        nu, nv = L.rank - 2, R.rank - 2
        zlegs = L.blocks.legs[:nu] + R.blocks.legs[:nv]
        zst = BlockStructure.invariant(zlegs)
        step = PairStep(0, 1, 2, tuple(range(nu)), (nu, nu + 1), tuple(range(nv)), (nv, nv + 1),
                        False, False, (tuple(range(nu + nv)), ()))
        Z = DTensor.empty(zst.dims, nu + nv, zst)
        NativePlan(ContractionPlan(3, 2, [step]), [L, R], zst).execute([L, R], Z)
        MZ = SL._maps(zst, list(range(nu)), list(range(nu, nu + nv)), ("cbe_z", nu),
                      _row_code_group(zlegs, list(range(nu))))
        zbuf = MZ.gather(Z)
        MQ = SL._maps(Q.blocks, ext, [k], ("cbe_q", k), _row_code_group(Q.blocks.legs, ext))
        qbuf = MQ.gather(Q)
        Pt = psi.tensors[v]
        kv = Pt.rank - 1 if t.parent[v] == u else childslot(t, v, u)
        pext = [l for l in range(Pt.rank) if l != kv]
        MP = SL._maps(Pt.blocks, pext, [kv], ("cbe_p", kv), _row_code_group(Pt.blocks.legs, pext))
        pbuf = MP.gather(Pt)
        cand = []
        for q in MZ.groups:
            Zq = MZ.matrix(zbuf, q)
            if -q in MP.gindex:
                Pq = MP.matrix(pbuf, -q)
                Zq = Zq - B.matmul(B.matmul(Zq, Pq, "N", "C"), Pq, "N", "T")
            if q in MQ.gindex:
                Qq = MQ.matrix(qbuf, q)
                Zq = Zq - B.matmul(Qq, B.matmul(Qq, Zq, "H"))
            if min(Zq.shape) == 0:
                continue
            sq, state = B.svd_spectrum(Zq)
            r = min(budget, len(sq))
            Uq, _, _ = B.svd_kept(Zq, state, r)
            for i in range(r):
                cand.append((-float(sq[i]), q, i, Uq))
            newrows[q] = MZ.rows[MZ.gindex[q]]
        norm_a = float(torch.linalg.vector_norm(A.data))
        cutoff = strat.spawn_threshold * max(norm_a, np.finfo(float).eps) / abs(complex(dz))
        cand.sort(key=lambda c: (c[0], c[1], c[2]))
        info["sigma_max"] = -cand[0][0] if cand else 0.0
        for negs, q, i, Uq in cand[:budget]:
            if -negs >= cutoff:
                sel.setdefault(q, (Uq, []))[1].append(i)
    # U = [Q, E] per charge of the external legs; R = U^H A per charge
    MA = SL._maps(st, ext, [k], ("cbe_a", k), grp_u)
    abuf = MA.gather(A)
    MQ = SL._maps(Q.blocks, ext, [k], ("cbe_q", k), _row_code_group(Q.blocks.legs, ext))
    qbuf = MQ.gather(Q)
    leg_old = st.legs[k]
    ucols, charges = {}, {}
    for q in sorted(set(MA.groups) | set(sel)):
        Uq = MQ.matrix(qbuf, q) if q in MQ.gindex else None
        if q in sel:
            Uv, idx = sel[q]
            E = Uv[:, idx]
            if Uq is not None:
                E = E - B.matmul(Uq, B.matmul(Uq, E, "H"))
            E, _ = torch.linalg.qr(E, mode="reduced")
            Uq = E if Uq is None else torch.cat([Uq, E], dim=1)
            info["selected"] += len(idx)
        if Uq is not None and Uq.shape[1] > 0:
            ucols[q] = Uq
    # the new bond leg: flat charge of each group = minus the external legs' charge
    rows_of = {}
    for key in st.blocks:
        rows_of.setdefault(grp_u(key), set()).add(tuple(key[l] for l in ext))
    for q in sel:
        rows_of.setdefault(q, set()).update(newrows[q])
    for q in ucols:
        rt = next(iter(rows_of[q]))
        charges[q] = tuple(-int(x) for x in np.sum([st.legs[l].sectors[rt[i]][0]
                                                    for i, l in enumerate(ext)], axis=0))
    order = sorted(ucols, key=lambda q: charges[q])
    new_leg = SectorLeg([(charges[q], ucols[q].shape[1]) for q in order])
    sidx = {q: i for i, q in enumerate(order)}
    ulegs = list(st.legs)
    ulegs[k] = new_leg
    ublocks = sorted(rt[:k] + (sidx[q],) + rt[k:] for q in order for rt in rows_of[q])
    ust = BlockStructure(ulegs, ublocks, _int_tuples=True)
    MU = SL._maps(ust, ext, [k], ("cbe_u", k), _row_code_group(ulegs, ext))
    ubuf = torch.cat([ucols[q].T.reshape(-1) for q in sorted(ucols)])
    Ut = MU.scatter(ubuf, ust, A.ncod)
    old_idx = {}
    for s_old, (qc, _) in enumerate(leg_old.sectors):
        old_idx[qc] = s_old
    rblocks, pieces = [], {}
    for q in order:
        so = old_idx.get(charges[q])
        if so is None or q not in MA.gindex:
            continue
        Rq = B.matmul(ucols[q], MA.matrix(abuf, q), "H")           # (new deg) x (old deg)
        rblocks.append((sidx[q], so))
        pieces[(sidx[q], so)] = Rq.T.reshape(-1)
    rst = BlockStructure([new_leg.dual(), leg_old], sorted(rblocks), _int_tuples=True)
    Rt = DTensor(rst.dims, torch.cat([pieces[b] for b in rst.blocks]), 1, rst)
    return Ut, Rt


def step_(ev, psi, H, dz):
    """step! — tdvp.jl:171-195."""
    if ev.cache is None or ev.cache.topo != psi.topo:
        ev.cache = EnvCache(psi.topo, tune_orders=True, tune_tiles=getattr(ev, "tune_tiles", False))
    if ev.order == 1:
        _tdvp1_sweep_(ev, psi, H, dz, rev=False)
    elif ev.order == 2:
        _tdvp1_sweep_(ev, psi, H, dz / 2, rev=False, terminal_dz=dz)
        _tdvp1_sweep_(ev, psi, H, dz / 2, rev=True, skip_initial_site=True)
    else:
        raise ValueError("order must be 1 or 2")
    return psi


def evolve_(ev, psi, H, dz, nsteps, callback=None):
    """evolve! — Evolution.jl:166-173."""
    for s in range(nsteps):
        step_(ev, psi, H, dz)
        if callback is not None:
            callback(s, psi)
    return psi


def _tdvp1_sweep_(ev, psi, H, dz, rev, terminal_dz=None, skip_initial_site=False):
    """tdvp.jl:197-251."""
    t = psi.topo
    cache = ev.cache
    order = postorder(t)
    if rev:
        order = list(reversed(order))
    herm = H.ishermitian
    move_center_(psi, order[0], cache)
    for i, n in enumerate(order):
        assert psi.center == n
        if not (skip_initial_site and i == 0):
            site_dz = terminal_dz if (terminal_dz is not None and i == len(order) - 1) else dz
            h1 = eff_h1(cache, psi, H, n, precision=ev.precision)
            A, _ = evolution_exponentiate_backend(h1, site_dz, psi.tensors[n], ishermitian=herm,
                                                  krylovdim=ev.krylovdim, tol=ev.tol, eager=ev.eager)
            update_tensor_(psi, n, A, caches=(cache,))
        if i == len(order) - 1:
            break
        seg = path_between(t, n, order[i + 1])
        for u, v in zip(seg[:-1], seg[1:]):
            if (t.parent[u] == v) != rev:
                _evolve_link_and_move_(ev, psi, H, u, v, dz, herm)
            else:
                move_center_(psi, v, cache)
    return psi


def _evolve_link_and_move_(ev, psi, H, u, v, dz, herm):
    """tdvp.jl:257-288 with the QR splits of tdvp.jl:291-301."""
    if psi.tensors[u].blocks is not None:
        if hasattr(psi.tensors[u].blocks, "dense_from_flat"):
            from . import su2
            return su2.evolve_link_and_move_(ev, psi, H, u, v, dz, herm,
                                             evolution_exponentiate_backend,
                                             cbe=isinstance(ev, TDVP1_CBE) and ev.enabled)
        return _evolve_sector_link_(ev, psi, H, u, v, dz, herm)
    t = psi.topo
    cache = ev.cache
    assert psi.center == u
    cbe = isinstance(ev, TDVP1_CBE) and ev.enabled
    if t.parent[u] == v:
        if cbe:
            Q, C = _cbe_split(ev, psi, H, u, v, dz)    # expanded: C :: V_new <- V_e
        else:
            Q, C = B.left_orth(psi.tensors[u])        # C :: V_new <- V_e
        psi.tensors[u] = Q
        cache.invalidate_node_(u)
        k0 = eff_h0(cache, psi, H, u, v, precision=ev.precision)
        C, _ = evolution_exponentiate_backend(k0, -dz, C, ishermitian=herm, krylovdim=ev.krylovdim,
                                              tol=ev.tol, eager=ev.eager)
        psi.tensors[v] = B.absorb_on_leg(psi.tensors[v], C, childslot(t, v, u))
    else:
        k = childslot(t, u, v)
        if cbe:
            Q, Cd = _cbe_split(ev, psi, H, u, v, dz)
        else:
            Q, Cd = B.orth_factor_leg(psi.tensors[u], k)   # Cd :: Y <- dual(V_e)
        psi.tensors[u] = Q
        C = B.transpose2(Cd)                           # :: V_e <- Y
        cache.invalidate_node_(u)
        psi.center = u
        k0 = eff_h0(cache, psi, H, v, u, precision=ev.precision)
        C, _ = evolution_exponentiate_backend(k0, -dz, C, ishermitian=herm, krylovdim=ev.krylovdim,
                                              tol=ev.tol, eager=ev.eager)
        psi.tensors[v] = B.mul_domain(psi.tensors[v], C)
    psi.center = v
    cache.invalidate_node_(v)
    return psi


def _evolve_sector_link_(ev, psi, H, u, v, dz, herm):
    """_evolve_link_and_move_ for U(1) x U(1) block-sparse states (tdvp.jl:257-301):
    the link splits are per-sector block QRs, the link tensor carries the block
    structure of its bond spaces and the backward step is the sector zero-site map."""
    t = psi.topo
    cache = ev.cache
    assert psi.center == u
    kw = dict(ishermitian=herm, krylovdim=ev.krylovdim, tol=ev.tol, eager=ev.eager)
    cbe = isinstance(ev, TDVP1_CBE) and ev.enabled
    if t.parent[u] == v:
        if cbe:
            Q, R = _cbe_split_sector(ev, psi, H, u, v, dz)
        else:
            Q, R = S.link_split(psi.tensors[u])         # R :: V_new <- V_e
        psi.tensors[u] = Q
        cache.invalidate_node_(u)
        k0 = eff_h0(cache, psi, H, u, v, precision=ev.precision)
        C = SL.to_structure(R, k0.blocks, 1)
        C, _ = evolution_exponentiate_backend(k0, -dz, C, **kw)
        psi.tensors[v] = S.absorb_link(psi.tensors[v], childslot(t, v, u), C, 1, 0)
    else:
        if cbe:
            Q, R = _cbe_split_sector(ev, psi, H, u, v, dz)
        else:
            Q, R = S.link_split(psi.tensors[u], childslot(t, u, v))   # R :: Y <- dual(V_e)
        psi.tensors[u] = Q
        cache.invalidate_node_(u)
        psi.center = u
        k0 = eff_h0(cache, psi, H, v, u, precision=ev.precision)
        C = SL.to_structure(SL.permute2(R), k0.blocks, 1)       # :: V_e <- Y
        C, _ = evolution_exponentiate_backend(k0, -dz, C, **kw)
        P = psi.tensors[v]
        psi.tensors[v] = S.absorb_link(P, P.rank - 1, C, 0, 1)
    psi.center = v
    cache.invalidate_node_(v)
    return psi
