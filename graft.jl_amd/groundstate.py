"""Ground-state sweeps — lib/GraftGroundState/src/GroundState/GroundState.jl.

``dmrg2_`` is dmrg2! (GroundState.jl:102-143): post-order bond sweep and its
reverse, a Lanczos ground state of the two-site effective Hamiltonian on every
bond, and a truncated split.  The local solve goes through
``groundstate_eigsolve_backend`` (GroundState.jl:25-26), here the device
Lanczos of ``krylov.eigsolve``.
"""
from __future__ import annotations

import time

from . import backend as B
from . import krylov
from .contractions import EnvCache, eff_h2, split_two_site_, two_site_tensor
from .networks import move_center_
from .trees import postorder


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


def dmrg2_(psi, H, trunc=B.NO_TRUNCATION, nsweeps=10, tol=1e-10, krylovdim=20, eig_tol=1e-12,
           precision="fp64", verbose=False, callback=None, timings=None):
    """Two-site DMRG; returns (psi, energies) with one energy per sweep.

    ``precision`` selects the product arithmetic of the local Lanczos
    ("fp64", or "fp32" / "bf16x3" with FP64 refinement of every local solve).
    ``timings``: a dict that receives seconds per phase (gauge move, two-site
    tensor, map build with environment updates, eigensolve, split) and the
    FP64/reduced apply FLOPs; phases are then synchronized, which costs a little."""
    if not H.ishermitian:
        raise ValueError("dmrg2_: DMRG requires ishermitian(H) == true")
    t = psi.topo
    cache = EnvCache(t, tune_orders=True)
    energies = []
    bonds = [n for n in postorder(t) if t.parent[n] >= 0]
    schedule = [(n, "m") for n in bonds] + [(n, "n") for n in reversed(bonds)]
    for sweep in range(1, nsweeps + 1):
        t0 = time.perf_counter()
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
