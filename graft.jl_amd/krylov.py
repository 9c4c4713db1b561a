"""Krylov solvers on device vectors — the map hooks of the reference's solvers.

``eigsolve`` replaces KrylovKit.eigsolve(map, x, 1, :SR; ishermitian=true,
krylovdim) behind groundstate_eigsolve_backend
(lib/GraftGroundState/src/GroundState/GroundState.jl:25-39); ``exponentiate``
replaces KrylovKit.exponentiate(map, dz, x; ishermitian, krylovdim, tol,
eager) behind evolution_exponentiate_backend
(lib/GraftEvolution/src/Evolution/Evolution.jl:134-148).

Vectors stay in HBM; every vector operation is a libgraft_heff.so kernel
(one-pass multi-dot / multi-axpy for the Gram-Schmidt steps).  Only the small
projected problem (k x k, k <= krylovdim) is solved on the host in FP64.
Orthogonalization is classical Gram-Schmidt with one full re-orthogonalization
pass (CGS2) against every basis vector.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import tensors as T


@dataclass
class ConvergenceInfo:
    converged: int
    residual: float
    numiter: int
    numops: int


def _orthogonalize(V, w):
    """w <- w - V V^H w twice (CGS2); returns the accumulated coefficients."""
    h = T.multi_inner(V, w)
    T.multi_axpy_(V, -h, w)
    h2 = T.multi_inner(V, w)
    T.multi_axpy_(V, -h2, w)
    return h + h2


def eigsolve(A, x0, howmany=1, which="SR", ishermitian=True, krylovdim=20, tol=1e-12,
             maxiter=100, verbosity=0):
    """Lowest eigenpairs of a hermitian map by thick-restart Lanczos (Krylov–Schur).

    Returns (vals, vecs, info) like KrylovKit.eigsolve.  Convergence: residual
    norm of each requested Ritz pair <= tol * max(1, |theta|)."""
    if which != "SR" or not ishermitian:
        raise NotImplementedError("eigsolve supports the hermitian :SR problem used by dmrg2!")
    nrm = T.norm(x0)
    if nrm == 0:
        raise ValueError("eigsolve: zero starting vector")
    v = x0.clone()
    T.scale_(v, 1.0 / nrm)
    V = [v]
    m = max(krylovdim, howmany + 2)
    H = np.zeros((m + 1, m + 1), dtype=np.complex128)
    k = 0            # number of vectors whose A-image is in H
    numops = 0
    theta = U = None
    beta = 0.0
    resid = np.inf
    for it in range(1, maxiter + 1):
        while k < m:
            w = A(V[k])
            numops += 1
            h = _orthogonalize(V, w)
            H[: len(V), k] = h
            beta = T.norm(w)
            H[k + 1, k] = beta
            k += 1
            if beta < 1e-14 * max(1.0, abs(h[-1])):
                del w
                break
            T.scale_(w, 1.0 / beta)
            V.append(w)
        Hk = H[:k, :k]
        Hk = (Hk + Hk.conj().T) / 2
        theta, U = np.linalg.eigh(Hk)
        last = U[k - 1, :] if k > 0 else np.zeros(0)
        res = np.abs(beta * last[:howmany])
        resid = float(res.max()) if len(res) else 0.0
        converged = bool(np.all(res <= tol * np.maximum(1.0, np.abs(theta[:howmany]))))
        invariant = len(V) == k          # breakdown: no residual vector was appended
        if converged or invariant or it == maxiter:
            vecs = [T.lincomb(V[:k], U[:, j], V[0]) for j in range(howmany)]
            nconv = int(np.sum(res <= tol * np.maximum(1.0, np.abs(theta[:howmany])))) if not invariant else howmany
            return (list(theta[:howmany]), vecs,
                    ConvergenceInfo(nconv, resid, it, numops))
        # thick restart: keep p lowest Ritz vectors plus the residual direction
        p = max(howmany, min(k - 1, m // 2))
        Y = [T.lincomb(V[:k], U[:, j], V[0]) for j in range(p)]
        r = V[k]
        V = Y + [r]
        H[:] = 0
        H[np.arange(p), np.arange(p)] = theta[:p]
        H[p, :p] = beta * last[:p]
        H[:p, p] = np.conj(beta * last[:p])
        k = p
    raise RuntimeError("unreachable")


def _expm_hermitian_small(Hk, dz):
    w, U = np.linalg.eigh((Hk + Hk.conj().T) / 2)
    return (U * np.exp(dz * w)[None, :]) @ U.conj().T


def exponentiate(A, dz, x, ishermitian=True, krylovdim=30, tol=1e-12, eager=True, maxiter=100):
    """exp(dz * A) x by Lanczos with adaptive sub-stepping (KrylovKit.exponentiate semantics).

    Returns (y, info).  The local error estimate of a step of size tau is
    ||x|| * |beta_k * [exp(tau T_k) e_1]_k|; a step is accepted when the
    estimate is <= tol * |tau| / |dz|, else tau is halved."""
    if not ishermitian:
        raise NotImplementedError("exponentiate supports hermitian maps (Lanczos)")
    total = complex(dz)
    if total == 0:
        return x.clone(), ConvergenceInfo(1, 0.0, 0, 0)
    y = x.clone()
    done = 0.0          # fraction of dz integrated
    frac = 1.0
    numops = 0
    numiter = 0
    worst = 0.0
    while done < 1.0 - 1e-15:
        frac = min(frac, 1.0 - done)
        nrm = T.norm(y)
        if nrm == 0:
            break
        v = y.clone()
        T.scale_(v, 1.0 / nrm)
        V = [v]
        H = np.zeros((krylovdim + 1, krylovdim + 1), dtype=np.complex128)
        k = 0
        beta = 0.0
        accepted = False
        while True:
            w = A(V[k])
            numops += 1
            h = _orthogonalize(V, w)
            H[: len(V), k] = h
            beta = T.norm(w)
            k += 1
            breakdown = beta < 1e-14
            Hk = H[:k, :k]
            tau = frac * total
            E = _expm_hermitian_small(Hk, tau)
            err = nrm * abs(beta * E[k - 1, 0])
            if breakdown or (eager and err <= tol * frac) or k == krylovdim:
                if breakdown or err <= tol * frac:
                    coef = nrm * E[:, 0]
                    y = T.lincomb(V[:k], coef, V[0])
                    done += frac
                    worst = max(worst, err)
                    accepted = True
                break
            T.scale_(w, 1.0 / beta)
            H[k, k - 1] = beta
            V.append(w)
        numiter += 1
        if not accepted:
            frac /= 2.0
            if numiter > maxiter:
                raise RuntimeError("exponentiate: step size underflow")
    return y, ConvergenceInfo(1, worst, numiter, numops)


LOWP_TOL = {"fp32": 1e-6, "bf16x3": 1e-4}


def eigsolve_mixed(A, x0, lowp="bf16x3", howmany=1, which="SR", ishermitian=True, krylovdim=20,
                   tol=1e-12, lowp_tol=None, maxiter=100, lowp_maxiter=4):
    """Two-phase eigensolver: reduced-precision products, then FP64 refinement.

    Phase 1 runs the thick-restart Lanczos on the map with FP32 or split-BF16x3
    products until the Ritz residual reaches ``lowp_tol`` (the products' own
    accuracy bounds it) or ``lowp_maxiter`` restarts pass.  Phase 2 restarts the same Lanczos in FP64 from that
    Ritz vector and converges to ``tol``; the returned eigenvalue and vector
    are those of the FP64 operator.  Returns (vals, vecs, info) with
    info.numops counting FP64 applies and ``info.lowp_ops`` the reduced ones.
    """
    Alp = A.with_precision(lowp)
    lowp_tol = LOWP_TOL[lowp] if lowp_tol is None else lowp_tol
    _, v1, i1 = eigsolve(Alp, x0, howmany, which, ishermitian, krylovdim, lowp_tol, lowp_maxiter)
    vals, vecs, i2 = eigsolve(A, v1[0], howmany, which, ishermitian, krylovdim, tol, maxiter)
    i2.lowp_ops = i1.numops
    return vals, vecs, i2
