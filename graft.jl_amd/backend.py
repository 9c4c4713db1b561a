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
    """Number of singular values kept under truncrank & trunctol & truncerror
    (Backend.jl:476-483) for a spectrum sorted in descending order.

    The rules are MatrixAlgebraKit's documented ones (the package itself,
    0.6.9, is not in the reference tree): truncrank(maxdim) keeps the maxdim
    largest; trunctol(atol, rtol) keeps every value >= max(atol, rtol * ||s||_2);
    truncerror(rtol=discarded_weight) keeps the fewest largest values whose
    discarded 2-norm is <= discarded_weight * ||s||_2; ``&`` keeps what every
    rule keeps.  At least one value is kept, so a bond never becomes empty."""
    s = np.asarray(s, dtype=np.float64)
    n = len(s)
    keep = min(n, ts.maxdim)
    tot = math.sqrt(float(np.sum(s ** 2))) if n else 0.0
    if ts.atol > 0 or ts.rtol > 0:
        thr = max(ts.atol, ts.rtol * tot)
        keep = min(keep, int(np.sum(s >= thr)))
    if ts.discarded_weight > 0 and n:
        tail = np.sqrt(np.cumsum((s ** 2)[::-1]))[::-1]    # tail[i] = 2-norm of s[i:]
        k = n
        for i in range(n):
            if tail[i] <= ts.discarded_weight * tot:
                k = i
                break
        keep = min(keep, k)
    return max(1, keep)


GRAM_SVD_MIN = 256   # smaller matrix side from which split_svd takes the Gram route
GRAM_ACCEPT = 1e-6   # a Gram level keeps the eigenvalues >= this fraction of its largest
TAIL_FLOOR = 1e-7    # deflate until what is left lies below this fraction of sigma_1
GRAM_MAX_LEVELS = 8


def gram_spectrum(M, wide):
    """Full singular spectrum of M (descending) and an orthonormal basis of
    its smaller side ordered like it, by deflated Gram eigendecompositions.

    Level 0 diagonalizes G = M M^H (``wide``) or M^H M.  An eigenvalue of a
    Gram matrix carries an absolute error ~ eps * (its largest eigenvalue), so
    a level keeps only the eigenvalues within GRAM_ACCEPT of its largest
    (singular values within 1e-3 of its top: relative error <= ~1e-10); the
    rest of the basis R is deflated: the next level diagonalizes the Gram
    matrix of R^H M (or M R), formed from M itself, whose top is the first
    value not yet kept.  Levels continue while the remainder's top exceeds
    TAIL_FLOOR * sigma_1.  Every value sigma >= 1e-5 sigma_1 is then accurate
    to 1e-10 relative (the product R^H M rounds at eps * sigma_1, as a
    backward-stable SVD does), and smaller ones to ~1e-14 sigma_1 absolute."""
    G = matmul(M, M, "N", "H") if wide else matmul(M, M, "H", "N")
    side = G.shape[0]
    vals, vecs = [], []
    R = None                              # remaining basis (None = identity)
    top0 = None
    for level in range(GRAM_MAX_LEVELS):
        G = (G + G.conj().T) / 2
        # unit scale: the device eigensolver's eigenvalue error does not follow
        # a matrix whose norm is far below one (a deflated level's is 1e-6 and less)
        scale = float(torch.diagonal(G).real.max())
        scale = scale if scale > 0 else 1.0
        lam, W = torch.linalg.eigh(G / scale)
        lam = torch.flip(lam, (0,)).clamp(min=0.0) * scale
        W = torch.flip(W, (1,))
        Rw = W if R is None else matmul(R, W)
        lam_h = lam.detach().cpu().numpy()
        if top0 is None:
            top0 = float(lam_h[0]) if len(lam_h) else 0.0
        acc = max(1, int(np.sum(lam_h >= GRAM_ACCEPT * lam_h[0])))
        last = (acc == len(lam_h) or level == GRAM_MAX_LEVELS - 1
                or lam_h[acc] < TAIL_FLOOR ** 2 * top0)
        if last:
            vals.append(lam_h)
            vecs.append(Rw)
            break
        vals.append(lam_h[:acc])
        vecs.append(Rw[:, :acc])
        # The deflated basis must be orthonormal and orthogonal to every kept
        # vector to working precision: a tail vector with a component theta
        # along a kept one of value s_j lifts its own squared value by about
        # (theta s_j)^2.  Two projections off the kept vectors, then a
        # Householder QR.
        R = Rw[:, acc:].contiguous()
        K = torch.cat(vecs, dim=1) if len(vecs) > 1 else vecs[0]
        for _ in range(2):
            R = R - matmul(K, matmul(K, R, "H"))
        R, _ = torch.linalg.qr(R, mode="reduced")
        B = matmul(R, M, "H", "N") if wide else matmul(M, R)
        G = matmul(B, B, "N", "H") if wide else matmul(B, B, "H", "N")
    lam_all = np.concatenate(vals)
    Q = torch.cat(vecs, dim=1) if len(vecs) > 1 else vecs[0]
    order = np.argsort(-lam_all, kind="stable")
    if np.any(order != np.arange(len(order))):
        lam_all = lam_all[order]
        Q = Q[:, torch.as_tensor(order, device=Q.device)]
    assert Q.shape == (side, side)
    return np.sqrt(lam_all), Q


def svd_spectrum(M, method=None):
    """Full singular spectrum of a torch matrix (numpy, descending) and the
    state ``svd_kept`` needs to produce the kept factors: the library SVD
    ("direct"), or from GRAM_SVD_MIN on the deflated Gram spectrum and basis."""
    rows, cols = M.shape
    if method is None:
        method = "gram" if min(rows, cols) >= GRAM_SVD_MIN else "direct"
    if method == "direct":
        U, S, Vh = torch.linalg.svd(M, full_matrices=False)
        return S.detach().cpu().numpy(), ("direct", U, S, Vh)
    wide = rows <= cols
    s, Q = gram_spectrum(M, wide)
    return s, ("gram", wide, Q)


def svd_kept(M, state, r):
    """The r leading singular triplets (U, S, Vh) of M from ``svd_spectrum``'s
    state: slices of the library SVD, or an exact SVD of the projection of M
    onto the top-r Gram basis (the library SVD when r exceeds half the
    smaller side, where the projection saves nothing)."""
    if state[0] == "direct":
        _, U, S, Vh = state
        return U[:, :r], S[:r], Vh[:r, :]
    _, wide, Q = state
    if 2 * r > min(M.shape):
        U, S, Vh = torch.linalg.svd(M, full_matrices=False)
        return U[:, :r], S[:r], Vh[:r, :]
    Qr = Q[:, :r]
    if wide:
        Ub, S, Vh = torch.linalg.svd(matmul(Qr, M, "H"), full_matrices=False)
        return matmul(Qr, Ub), S, Vh
    U, S, Vb = torch.linalg.svd(matmul(M, Qr), full_matrices=False)
    return U, S, matmul(Vb, Qr, "N", "H")


def split_svd(M, ts=NO_TRUNCATION, method=None):
    """Truncated SVD of a torch matrix -> (U, S, Vh, discarded norm, full spectrum).

    ``method`` "direct": the library SVD.  "gram" (the default from
    GRAM_SVD_MIN on): the full spectrum from ``gram_spectrum`` (deflated
    Hermitian eigendecompositions on the smaller side: every value down to
    1e-5 sigma_1 to 1e-10 relative), then the kept part from an exact SVD of
    the projection Q_r^H M onto the top-r basis vectors.  The kept singular
    values and vectors are those of M to working precision when the spectrum
    has a gap at the cut.  On the GPU the library SVD of a 2048 x 2048
    complex matrix costs 18x a Gram eigendecomposition (MI355X, ROCm 7)."""
    s, state = svd_spectrum(M, method)
    r = truncation_rank(s, ts)
    U, S, Vh = svd_kept(M, state, r)
    s = s.copy()
    s[:r] = S.detach().cpu().numpy()
    disc = float(np.sqrt(np.sum(s[r:] ** 2)))
    return U, S, Vh, disc, s


def ones_tensor(dims):
    """Backend.jl:144-150 for trivial sectors: every entry one."""
    return DTensor(dims, torch.ones(int(np.prod(dims)), dtype=torch.complex128,
                                    device=torch.device("cuda", torch.cuda.current_device())),
                   len(dims))
