"""L0 Backend on the device — lib/GraftFoundation/src/Backend/Backend.jl.

``contract_pair`` runs one binary contraction in libgraft_heff.so (the same
grouped strided MFMA GEMM as the effective maps).  The gauge factorizations
(``left_orth``, ``orth_factor_leg``, ``split_svd``) call rocSOLVER through
torch.linalg on the device; they are outside the H_eff hot path.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat
from .planning import PairStep
from .tensors import DTensor

_pair_plans = {}


def _pair_plan(step, A, B):
    key = (step, A.structure_key(), B.structure_key())
    p = _pair_plans.get(key)
    if p is None:
        ctx = nat.default_context()
        descs = (nat.TensorDescC * 2)(A.desc().c, B.desc().c)
        s = nat.make_step(PairStep(0, 1, 2, step.oindA, step.cindA, step.oindB, step.cindB,
                                   step.conjA, step.conjB, step.out))
        h = ctypes.c_void_p()
        nat.check(nat.lib().graft_plan_create(ctx.handle, ctypes.byref(s), 1, 3, 2, descs, 2, None,
                                              ctypes.byref(h)))
        outn = ctypes.c_uint64()
        nat.check(nat.lib().graft_plan_info(h, None, None, ctypes.byref(outn)))
        p = (h, outn.value)
        _pair_plans[key] = p
    return p


def contract_pair(A, oindA, cindA, conjA, B, cindB, oindB, conjB, out):
    """Backend.contract_pair (Backend.jl:166-170) with 0-based expert-mode tuples.

    ``out`` = (codomain positions, domain positions) over (A open..., B open...)."""
    if A.blocks is not None or B.blocks is not None:
        raise NotImplementedError("sector tensors use planned maps (graft_plan_*), not contract_pair")
    step = PairStep(0, 1, 2, tuple(oindA), tuple(cindA), tuple(oindB), tuple(cindB),
                    bool(conjA), bool(conjB), (tuple(out[0]), tuple(out[1])))
    h, n = _pair_plan(step, A, B)
    inter = [A.dims[i] for i in oindA] + [B.dims[i] for i in oindB]
    flat = list(out[0]) + list(out[1])
    C = DTensor.empty([inter[i] for i in flat], len(out[0]))
    ctx = nat.default_context()
    ins = nat.ptr_array([A.ptr, B.ptr])
    nat.check(nat.lib().graft_plan_execute(ctx.handle, h, ins, ctypes.c_void_p(C.ptr), nat.PREC_FP64))
    return C


def _others(N, k):
    return tuple(j for j in range(N) if j != k)


def _restore_perm(N, k):
    """Output order after (others..., new) so that the new leg sits at k (Backend.jl:515-521)."""
    return tuple(j if j < k else (N - 1 if j == k else j - 1) for j in range(N))


def absorb_on_leg(A, C, k):
    """A' = C applied to codomain leg k of A (C :: new <- old) — Backend.jl:533-543."""
    N = A.rank
    if C.dims[1] != A.dims[k]:
        raise ValueError(f"absorb_on_leg: domain of C ({C.dims[1]}) != leg {k} of A ({A.dims[k]})")
    order = _restore_perm(N, k)
    return contract_pair(A, _others(N, k), (k,), False, C, (1,), (0,), False,
                         (order[:A.ncod], order[A.ncod:]))


def mul_domain(A, C):
    """A * C on A's last (domain) leg: A'[..., i] = sum_j A[..., j] C[j, i]."""
    N = A.rank
    return contract_pair(A, tuple(range(N - 1)), (N - 1,), False, C, (0,), (1,), False,
                         (tuple(range(A.ncod)), tuple(range(A.ncod, N))))


# -- torch views (column-major DTensor <-> torch row-major) -------------------


def matrix_view(A, nrows_legs):
    """(prod first legs) x (prod remaining legs) matrix view of a dense DTensor."""
    r = int(np.prod(A.dims[:nrows_legs])) if nrows_legs else 1
    c = A.numel // r
    return A.data.view(c, r).T


def from_matrix(M, dims, ncod):
    """Store an (r x c) torch matrix column-major as a DTensor with the given leg dims."""
    return DTensor(dims, M.T.contiguous().view(-1), ncod)


def permute(A, perm, ncod=None):
    """Data-movement permute of a dense DTensor (new leg q = old leg perm[q])."""
    N = A.rank
    t = A.data.view(tuple(reversed(A.dims)))
    tp = [N - 1 - perm[N - 1 - j] for j in range(N)]
    data = t.permute(tp).contiguous().view(-1)
    return DTensor([A.dims[p] for p in perm], data, A.ncod if ncod is None else ncod)


def left_orth(A):
    """QR across (codomain | last leg): A = Q . C, Q isometric — MatrixAlgebraKit left_orth."""
    M = matrix_view(A, A.rank - 1)
    Q, R = torch.linalg.qr(M, mode="reduced")
    Qt = from_matrix(Q, A.dims[:-1] + (Q.shape[1],), A.rank - 1)
    Ct = from_matrix(R, (R.shape[0], R.shape[1]), 1)
    return Qt, Ct


def orth_factor_leg(A, k):
    """Backend.jl:602-609: Q isometric away from leg k, C :: Y <- dual(V_old)."""
    N = A.rank
    perm = _others(N, k) + (k,)
    t = permute(A, perm)
    Q, C = left_orth(t)
    back = _restore_perm(N, k)
    return permute(Q, back, A.ncod), C


def transpose2(C):
    return permute(C, (1, 0), 1)


@dataclass
class TruncationScheme:
    """Backend.jl:466-471."""
    maxdim: int = 2 ** 62
    atol: float = 0.0
    rtol: float = 0.0
    discarded_weight: float = 0.0


NO_TRUNCATION = TruncationScheme()


def truncation_rank(s, ts):
    """Number of singular values kept (sorted descending) under truncrank & trunctol & truncerror."""
    n = len(s)
    keep = min(n, ts.maxdim)
    if ts.atol > 0 or ts.rtol > 0:
        thr = max(ts.atol, ts.rtol * (s[0] if n else 0.0))
        keep = min(keep, int(np.sum(s > thr)))
    if ts.discarded_weight > 0 and n:
        tot = math.sqrt(float(np.sum(s ** 2)))
        tail = np.sqrt(np.cumsum((s ** 2)[::-1]))[::-1]    # tail[i] = norm of s[i:]
        k = n
        for i in range(n):
            if tail[i] <= ts.discarded_weight * tot:
                k = i
                break
        keep = min(keep, k)
    return max(1, keep)


GRAM_SVD_MIN = 256   # smaller matrix side from which split_svd takes the Gram route


def split_svd(M, ts=NO_TRUNCATION, method=None):
    """Truncated SVD of a torch matrix -> (U, S, Vh, discarded norm, full spectrum).

    ``method`` "direct": the library SVD.  "gram" (the default from
    GRAM_SVD_MIN on): the full spectrum from the Hermitian eigendecomposition
    of the smaller Gram matrix G = M M^H (or M^H M), then the kept part from an
    exact SVD of the projection Q_r^H M onto the top-r eigenvectors.  The kept
    singular values and vectors are those of M to working precision when the
    spectrum has a gap at the cut; the discarded tail (sqrt of Gram eigenvalues)
    carries absolute error ~ eps * sigma_1, so a tail value sigma_i has relative
    error ~ eps (sigma_1 / sigma_i)^2.  On the GPU the library SVD of a
    2048 x 2048 complex matrix costs 18x the Gram route (MI355X, ROCm 7)."""
    rows, cols = M.shape
    if method is None:
        method = "gram" if min(rows, cols) >= GRAM_SVD_MIN else "direct"
    if method == "direct":
        U, S, Vh = torch.linalg.svd(M, full_matrices=False)
        s = S.detach().cpu().numpy()
        r = truncation_rank(s, ts)
        disc = float(np.sqrt(np.sum(s[r:] ** 2)))
        return U[:, :r], S[:r], Vh[:r, :], disc, s
    wide = rows <= cols
    G = M @ M.conj().T if wide else M.conj().T @ M
    lam, Q = torch.linalg.eigh(G)
    lam = torch.flip(lam, (0,)).clamp(min=0.0)
    Q = torch.flip(Q, (1,))
    s = np.sqrt(lam.detach().cpu().numpy())
    r = truncation_rank(s, ts)
    if 2 * r > min(rows, cols):          # little truncation: the projection saves nothing
        return split_svd(M, ts, "direct")
    Qr = Q[:, :r]
    if wide:
        Ub, S, Vh = torch.linalg.svd(Qr.conj().T @ M, full_matrices=False)
        U = Qr @ Ub
    else:
        U, S, Vb = torch.linalg.svd(M @ Qr, full_matrices=False)
        Vh = Vb @ Qr.conj().T
    s = s.copy()
    s[:r] = S.detach().cpu().numpy()
    disc = float(np.sqrt(np.sum(s[r:] ** 2)))
    return U, S, Vh, disc, s


def ones_tensor(dims):
    """Backend.jl:144-150 for trivial sectors: every entry one."""
    return DTensor(dims, torch.ones(int(np.prod(dims)), dtype=torch.complex128,
                                    device=torch.device("cuda", torch.cuda.current_device())),
                   len(dims))
