"""Ground-state sweeps — lib/GraftGroundState/src/GroundState/GroundState.jl.

``dmrg2_`` is dmrg2! (GroundState.jl:102-143): post-order bond sweep and its
reverse, a Lanczos ground state of the two-site effective Hamiltonian on every
bond, and a truncated split.  Before each sweep a binary root without a
physical leg has both child bonds completed by zero-weight directions
(GroundState.jl:119-120), else each root-edge split stays rank-limited by the
other root bond.  The local solve goes through
``groundstate_eigsolve_backend`` (GroundState.jl:25-26), here the device
Lanczos of ``krylov.eigsolve``.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import backend as B
from . import krylov
from .contractions import EnvCache, eff_h2, split_two_site_, two_site_tensor
from .networks import move_center_
from .trees import childslot, postorder


def groundstate_eigsolve_backend(effective, x, howmany, which, precision="fp64", **kw):
    """Hook point of GroundState.jl:25-26.

    ``precision`` other than "fp64" runs the two-phase solver: reduced-precision
    products first, then FP64 refinement to the same tolerance."""
    if precision == "fp64":
        return krylov.eigsolve(effective, x, howmany, which, **kw)
    return krylov.eigsolve_mixed(effective, x, precision, howmany, which, **kw)


def _energy_converged(energies, tol):
    return len(energies) >= 2 and abs(energies[-1] - energies[-2]) < tol


class _Phases:
    """Wall time per sweep phase (device-synchronized at phase ends) and apply FLOPs."""

    def __init__(self, out):
        import torch
        self.out, self.sync = out, torch.cuda.synchronize
        self.t = time.perf_counter()

    def mark(self, name):
        self.sync()
        now = time.perf_counter()
        self.out[name] = self.out.get(name, 0.0) + now - self.t
        self.t = now


def _physless_root_two_site_targets(psi, trunc):
    """_physless_root_two_site_targets (expansion.jl:542-549): at a binary root
    without a physical leg, both root-child bonds are opened to the ordinary
    two-site cap, bounded by each child's codomain dimension."""
    t = psi.topo
    r = t.root
    kids = t.children[r]
    if psi.hasphys[r] or len(kids) != 2:
        return []
    target = min(trunc.maxdim, min(int(np.prod(psi.tensors[n].dims[:-1])) for n in kids))
    return [(n, target) for n in kids]


def _null_enrich_split(A, maxdim):
    """_null_enrich_split (expansion.jl:560-575): A == U . R with U isometric and
    up to ``maxdim`` columns: the column space of A completed by orthonormal
    directions of its complement (zero weight, so the state is unchanged).
    The complement directions come from a fixed-seed random probe projected
    off range(A) (any orthonormal complement basis serves: the following
    variational updates choose which directions acquire weight).  Returns
    (U as a site tensor, R :: new <- old) or (A, None) when there is no room."""
    if A.blocks is not None:
        raise NotImplementedError("the physless-root bootstrap is built for dense and SU(2) states")
    M = B.matrix_view(A, A.rank - 1)
    rows, old = M.shape
    room = min(maxdim, rows) - old
    if room <= 0:
        return A, None
    Q, _ = B.thin_qr(M)
    g = torch.Generator(device=M.device)
    g.manual_seed(0x5EED)
    G = torch.randn((rows, room), dtype=M.dtype, device=M.device, generator=g)
    for _ in range(2):                        # twice is enough (Kahan, Parlett)
        G = G - B.matmul(Q, B.matmul(Q, G, "H"))
    E, _ = torch.linalg.qr(G, mode="reduced")
    U = torch.cat([Q, E], dim=1)
    R = B.matmul(U, M, "H")
    Ut = B.from_matrix(U, A.dims[:-1] + (U.shape[1],), A.ncod)
    Rt = B.from_matrix(R, (R.shape[0], R.shape[1]), 1)
    return Ut, Rt


def _bootstrap_physless_root_(psi, cache, targets):
    """_bootstrap_physless_root! (expansion.jl:584-602): complete both root-child
    subspaces before the sweep, so that neither root-edge split is rank-limited
    by the other root bond."""
    t = psi.topo
    r = t.root
    for n, target in targets:
        if t.parent[n] != r:
            raise ValueError("root bootstrap target is not a root child")
        if psi.tensors[n].dims[-1] >= target:
            continue
        move_center_(psi, n, cache)
        A = psi.tensors[n]
        if A.blocks is not None and hasattr(A.blocks, "dense_from_flat"):
            from . import su2                     # SU(2) blocks: per-spin completion
            U, R, new = su2.null_enrich_split(A, target)
            if R is None:
                continue
            psi.tensors[n] = U
            psi.tensors[r] = su2._absorb_on_leg(psi.tensors[r], childslot(t, r, n), R, new.irreps)
        else:
            U, R = _null_enrich_split(A, target)
            if R is None:
                continue
            psi.tensors[n] = U
            psi.tensors[r] = B.absorb_on_leg(psi.tensors[r], R, childslot(t, r, n))
        psi.center = r
        cache.invalidate_edge_(n, r)
    return psi


def dmrg2_(psi, H, trunc=B.NO_TRUNCATION, nsweeps=10, tol=1e-10, krylovdim=20, eig_tol=1e-12,
           precision="fp64", verbose=False, callback=None, timings=None, tune_tiles=False):
    """Two-site DMRG; returns (psi, energies) with one energy per sweep.

    ``precision`` selects the product arithmetic of the local Lanczos
    ("fp64", or "fp32" / "bf16x3" with FP64 refinement of every local solve).
    ``timings``: a dict that receives seconds per phase (gauge move, two-site
    tensor, map build with environment updates, eigensolve, split) and the
    FP64/reduced apply FLOPs; phases are then synchronized, which costs a little.
    ``tune_tiles``: autotune the tiles of heavy sector maps once per step shape
    class (EnvCache tune_tiles); the first map of each class pays the tuning."""
    if not H.ishermitian:
        raise ValueError("dmrg2_: DMRG requires ishermitian(H) == true")
    t = psi.topo
    cache = EnvCache(t, tune_orders=True, tune_tiles=tune_tiles)
    energies = []
    bonds = [n for n in postorder(t) if t.parent[n] >= 0]
    schedule = [(n, "m") for n in bonds] + [(n, "n") for n in reversed(bonds)]
    for sweep in range(1, nsweeps + 1):
        t0 = time.perf_counter()
        _bootstrap_physless_root_(psi, cache, _physless_root_two_site_targets(psi, trunc))
        E = float("nan")
        ph = _Phases(timings) if timings is not None else None
        for n, center_on in schedule:
            m = t.parent[n]
            move_center_(psi, n, cache)
            ph and ph.mark("move_center")
            theta = two_site_tensor(psi, n, m)
            ph and ph.mark("two_site")
            h2 = eff_h2(cache, psi, H, n, m, out_structure=theta.blocks)
            ph and ph.mark("map_build")
            vals, vecs, info = groundstate_eigsolve_backend(
                h2, theta, 1, "SR", precision=precision, ishermitian=True, krylovdim=krylovdim,
                tol=eig_tol)
            E = float(vals[0].real)
            if ph:
                ph.mark("eigsolve")
                timings["flops_fp64"] = timings.get("flops_fp64", 0.0) + info.numops * h2.flops
                timings["flops_lowp"] = (timings.get("flops_lowp", 0.0)
                                         + getattr(info, "lowp_ops", 0) * h2.flops)
            cache.invalidate_edge_(n, m)
            split_two_site_(psi, vecs[0], n, m, trunc=trunc, center_on=center_on)
            ph and ph.mark("split")
            if callback is not None:
                callback(sweep, n, m, E, info)
        energies.append(E)
        if verbose:
            print(f"dmrg2_ sweep {sweep}: E = {E:.14f} ({time.perf_counter() - t0:.2f} s)")
        if _energy_converged(energies, tol):
            break
    return psi, energies
