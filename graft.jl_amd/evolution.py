"""Time evolution — lib/GraftEvolution/src/Evolution/tdvp.jl (TDVP1).

One-site TDVP with the tree projector-splitting sweep (tdvp.jl:197-251):
every node is forward-evolved once per sweep with the one-site map, and an
edge is backward-evolved with the zero-site (link) map exactly when the
centre crosses it child -> parent (parent -> child on the reverse sweep).
Krylov exponentials go through ``evolution_exponentiate_backend``
(Evolution.jl:134-148), here the device Lanczos of ``krylov.exponentiate``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

from . import backend as B
from . import krylov
from .contractions import EnvCache, eff_h0, eff_h1
from .networks import move_center_, update_tensor_
from .trees import childslot, nnodes, path_between, postorder


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


def step_(ev, psi, H, dz):
    """step! — tdvp.jl:171-195."""
    if ev.cache is None or ev.cache.topo != psi.topo:
        ev.cache = EnvCache(psi.topo)
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
    t = psi.topo
    cache = ev.cache
    assert psi.center == u
    if t.parent[u] == v:
        Q, C = B.left_orth(psi.tensors[u])            # C :: V_new <- V_e
        psi.tensors[u] = Q
        cache.invalidate_node_(u)
        k0 = eff_h0(cache, psi, H, u, v, precision=ev.precision)
        C, _ = evolution_exponentiate_backend(k0, -dz, C, ishermitian=herm, krylovdim=ev.krylovdim,
                                              tol=ev.tol, eager=ev.eager)
        psi.tensors[v] = B.absorb_on_leg(psi.tensors[v], C, childslot(t, v, u))
    else:
        k = childslot(t, u, v)
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
