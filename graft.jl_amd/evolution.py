"""Time evolution — lib/GraftEvolution/src/Evolution/tdvp.jl (TDVP1).

One-site TDVP with the tree projector-splitting sweep (tdvp.jl:197-251):
every node is forward-evolved once per sweep with the one-site map, and an
edge is backward-evolved with the zero-site (link) map exactly when the
centre crosses it child -> parent (parent -> child on the reverse sweep).
Krylov exponentials go through ``evolution_exponentiate_backend``
(Evolution.jl:134-148), here the device Lanczos of ``krylov.exponentiate``.
States may be dense or U(1) x U(1) block-sparse; for the latter the link
splits are per-sector block QRs and the link map carries the bond's block
structure.  Controlled bond expansion (TDVP1_CBE) is built for dense states.
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
from .contractions import EnvCache, eff_h0, eff_h1, eff_h2, two_site_tensor
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


@dataclass
class PredictorCBE:
    """PredictorCBE(; max_add=32, spawn_threshold=1e-10, neigs=2, krylovdim=4) —
    cbe.jl:115-149: projected two-site energy-variation selector.  Directions are
    the leading left singular vectors of the double-complement core
    M = (1 - Q Q^H) . H_eff(Theta) . (1 - P_partner); at most min(room, max_add,
    neigs) are spawned, each only if |dz| * sigma / ||Theta|| > spawn_threshold.
    The selection here is an exact eigendecomposition of the core's Gram matrix
    on its smaller side (the reference keeps the same exact path for sector maps
    and a 4-dimensional Golub-Kahan solve for dense ones; ``krylovdim`` is kept
    for the interface)."""
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


def _cbe_directions(ev, psi, H, n, m, active_is_n, Q, dz, room):
    """cbe.jl:332-352, 590-619: new orthonormal directions for the active site,
    orthogonal to its current range Q, or None."""
    strat = ev.cbe
    budget = min(room, strat.max_add, strat.neigs)
    info = {"selected": 0, "budget": budget}
    ev.last_cbe_info = info
    if budget <= 0 or dz == 0:
        return None
    theta = two_site_tensor(psi, n, m)
    Y = eff_h2(ev.cache, psi, H, n, m)(theta)
    pn = psi.tensors[n].ncod
    k = childslot(psi.topo, m, n)
    Am = psi.tensors[m]
    Am_k = B.permute(Am, tuple(j for j in range(Am.rank) if j != k) + (k,))
    Ym = B.matrix_view(Y, pn)                         # (n legs) x (m legs but k)
    if active_is_n:
        Z = Ym.clone()
        Bp = B.matrix_view(Am_k, Am_k.rank - 1)       # partner (m legs but k) x bond
    else:
        Z = Ym.T.contiguous()
        An = psi.tensors[n]
        Bp = B.matrix_view(An, An.rank - 1)           # partner (n legs) x bond
    # double complement: Theta = M_a . Bp^T, so the partner's range acts on the right
    Z = Z - B.matmul(B.matmul(Z, Bp, "N", "C"), Bp, "N", "T")
    Z = Z - B.matmul(Q, B.matmul(Q, Z, "H"))
    norm_theta = T.norm(theta)
    cutoff = strat.spawn_threshold * max(norm_theta, np.finfo(float).eps) / abs(complex(dz))
    rows, cols = Z.shape
    if cols <= rows:
        lam, V = torch.linalg.eigh(B.matmul(Z, Z, "H"))
        lam, V = torch.flip(lam, (0,)), torch.flip(V, (1,))
        sig = torch.sqrt(lam.clamp(min=0.0))
        sel = [i for i in range(min(budget, len(lam))) if float(sig[i]) > cutoff]
        if not sel:
            return None
        E = B.matmul(Z, V[:, sel]) / sig[sel]
    else:
        lam, U = torch.linalg.eigh(B.matmul(Z, Z, "N", "H"))
        lam, U = torch.flip(lam, (0,)), torch.flip(U, (1,))
        sig = torch.sqrt(lam.clamp(min=0.0))
        sel = [i for i in range(min(budget, len(lam))) if float(sig[i]) > cutoff]
        if not sel:
            return None
        E = U[:, sel]
    # exact orthogonality to the kept range, then orthonormal columns
    E = E - B.matmul(Q, B.matmul(Q, E, "H"))
    E, _ = torch.linalg.qr(E, mode="reduced")
    info["selected"] = len(sel)
    info["sigma_max"] = float(sig[0])
    return E


def _cbe_split(ev, psi, H, u, v, dz):
    """_cbe_projected_split + _cbe_state_preserving_splice (cbe.jl:575-584, 743-760):
    A_u = U . R with U = [Q, E] and R = U^H A_u = [R_qr; 0]; the state is unchanged.
    Returns (U as the site tensor in A_u's leg order, R :: new bond <- old bond)."""
    t = psi.topo
    A = psi.tensors[u]
    if A.blocks is not None:
        raise NotImplementedError("controlled bond expansion is built for dense states only")
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


def step_(ev, psi, H, dz):
    """step! — tdvp.jl:171-195."""
    if (any(t.blocks is not None for t in psi.tensors) and isinstance(ev, TDVP1_CBE)
            and ev.enabled):
        raise NotImplementedError("controlled bond expansion is built for dense states only")
    if ev.cache is None or ev.cache.topo != psi.topo:
        ev.cache = EnvCache(psi.topo, tune_orders=True)
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
    if t.parent[u] == v:
        Q, R = S.link_split(psi.tensors[u])             # R :: V_new <- V_e
        psi.tensors[u] = Q
        cache.invalidate_node_(u)
        k0 = eff_h0(cache, psi, H, u, v, precision=ev.precision)
        C = SL.to_structure(R, k0.blocks, 1)
        C, _ = evolution_exponentiate_backend(k0, -dz, C, **kw)
        psi.tensors[v] = S.absorb_link(psi.tensors[v], childslot(t, v, u), C, 1, 0)
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
