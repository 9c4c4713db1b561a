"""L0 Backend on the device — lib/GraftFoundation/src/Backend/Backend.jl.

``contract_pair`` runs one binary contraction in libgraft_heff.so (the same
grouped strided MFMA GEMM as the effective maps); ``matmul`` puts the tall
matrix products of the gauge and expansion steps on it.  The factorizations
themselves (QR, Hermitian eigensolves, small SVDs) call rocSOLVER through
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


def _matrix_operand(M):
    """(DTensor over the memory of the 2-D torch matrix M, leg of its row index,
    leg of its column index); no copy when M or M^T is contiguous."""
    if M.is_conj():
        M = M.resolve_conj()
    if M.is_contiguous():
        return DTensor((M.shape[1], M.shape[0]), M.reshape(-1), 2), 1, 0
    if M.T.is_contiguous():
        return DTensor((M.shape[0], M.shape[1]), M.T.reshape(-1), 2), 0, 1
    M = M.contiguous()
    return DTensor((M.shape[1], M.shape[0]), M.reshape(-1), 2), 1, 0


def matmul(X, Y, op_x="N", op_y="N"):
    """op(X) @ op(Y) for complex128 device matrices on the library's GEMM
    (split-K for long contractions: the Gram and projection products of the
    gauge and expansion steps have K of 10^5 and a few output tiles).  op is
    "N", "T" (transpose), "C" (conjugate) or "H" (conjugate transpose).
    Returns a row-major torch matrix."""
    A, ra, ca = _matrix_operand(X)
    B, rb, cb = _matrix_operand(Y)
    i_leg, k_a = (ca, ra) if op_x in ("T", "H") else (ra, ca)
    k_b, j_leg = (cb, rb) if op_y in ("T", "H") else (rb, cb)
    if A.dims[k_a] != B.dims[k_b]:
        raise ValueError(f"matmul: inner dimensions {A.dims[k_a]} and {B.dims[k_b]} differ")
    C = contract_pair(A, (i_leg,), (k_a,), op_x in ("C", "H"), B, (k_b,), (j_leg,),
                      op_y in ("C", "H"), ((1, 0), ()))
    return C.data.view(A.dims[i_leg], B.dims[j_leg])


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


TALL_QR_MIN_COLS = 64     # Cholesky route from this many columns ...
TALL_QR_ASPECT = 8        # ... and at least this many rows per column


def _householder_qr(M):
    return torch.linalg.qr(M, mode="reduced")


def tall_qr(M):
    """Thin QR of a tall matrix by shifted Cholesky QR in three passes
    (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa, "Shifted Cholesky QR
    for computing the QR factorization of ill-conditioned matrices", SIAM J.
    Sci. Comput. 42 (2020) A477): G = M^H M + s I with s = 11 (m n + n (n+1))
    eps ||M||_F^2, R1 = chol(G), Q1 = M R1^-1, then two plain Cholesky QR passes
    on Q1.  The Gram and triangular-factor products run on the library GEMM
    (split-K over the 10^5 rows), where rocSOLVER's Householder panels are
    sequences of matrix-vector products.  Falls back to Householder when a
    Cholesky factorization breaks down (numerically rank-deficient M) or the
    last Gram matrix is not within 1e-6 of the identity.  R has a positive
    real diagonal (Householder's may carry phases: the gauge differs, the
    factorization does not)."""
    m, n = M.shape
    eye = torch.eye(n, dtype=M.dtype, device=M.device)
    G = matmul(M, M, "H")
    G = (G + G.conj().T) / 2
    shift = 11.0 * (m * n + n * (n + 1)) * np.finfo(np.float64).eps * float(torch.diagonal(G).real.sum())
    R, info = torch.linalg.cholesky_ex(G + shift * eye, upper=True)
    if int(info) != 0:
        return _householder_qr(M)
    Q = matmul(M, torch.linalg.solve_triangular(R, eye, upper=True))
    Rtot = R
    dev = math.inf
    for _ in range(2):
        G = matmul(Q, Q, "H")
        G = (G + G.conj().T) / 2
        dev = float((G - eye).abs().max())
        R, info = torch.linalg.cholesky_ex(G, upper=True)
        if int(info) != 0:
            return _householder_qr(M)
        Q = matmul(Q, torch.linalg.solve_triangular(R, eye, upper=True))
        Rtot = R @ Rtot
    if not dev <= 1e-6:                  # (also catches NaN from a near-singular pass)
        return _householder_qr(M)
    return Q, Rtot


def thin_qr(M):
    """Thin QR: shifted Cholesky QR for tall matrices with many columns, else Householder."""
    rows, cols = M.shape
    if cols >= TALL_QR_MIN_COLS and rows >= TALL_QR_ASPECT * cols:
        return tall_qr(M)
    return _householder_qr(M)


def left_orth(A):
    """QR across (codomain | last leg): A = Q . C, Q isometric — MatrixAlgebraKit left_orth."""
    M = matrix_view(A, A.rank - 1)
    Q, R = thin_qr(M)
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
    G = matmul(M, M, "N", "H") if wide else matmul(M, M, "H", "N")
    lam, Q = torch.linalg.eigh(G)
    lam = torch.flip(lam, (0,)).clamp(min=0.0)
    Q = torch.flip(Q, (1,))
    s = np.sqrt(lam.detach().cpu().numpy())
    r = truncation_rank(s, ts)
    if 2 * r > min(rows, cols):          # little truncation: the projection saves nothing
        return split_svd(M, ts, "direct")
    Qr = Q[:, :r]
    if wide:
        Ub, S, Vh = torch.linalg.svd(matmul(Qr, M, "H"), full_matrices=False)
        U = matmul(Qr, Ub)
    else:
        U, S, Vb = torch.linalg.svd(matmul(M, Qr), full_matrices=False)
        Vh = matmul(Vb, Qr, "N", "H")
    s = s.copy()
    s[:r] = S.detach().cpu().numpy()
    disc = float(np.sqrt(np.sum(s[r:] ** 2)))
    return U, S, Vh, disc, s


def ones_tensor(dims):
    """Backend.jl:144-150 for trivial sectors: every entry one."""
    return DTensor(dims, torch.ones(int(np.prod(dims)), dtype=torch.complex128,
                                    device=torch.device("cuda", torch.cuda.current_device())),
                   len(dims))
