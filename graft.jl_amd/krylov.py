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

For a native effective map the Lanczos expansion runs device-resident: the
basis lives in fixed buffers, the Gram-Schmidt coefficients and norms stay in
device memory, and a run of iterations (apply, two multi-dot/multi-axpy
passes, norm, scale) is captured once as a HIP graph and replayed at every
restart; the host reads the coefficients once per run of iterations.
"""
from __future__ import annotations

import contextlib
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from . import tensors as T


# A Lanczos step whose residual norm is below BREAKDOWN times the largest
# projected-matrix entry has found an invariant subspace: the next vector
# would be rounding noise divided by rounding noise.
BREAKDOWN = 1e-13

# Smallest Krylov tolerance a map with reduced-precision products can meet:
# its products carry relative errors near these (FP32: 2^-24 per operation,
# split BF16x3: the dropped lo*lo term and FP32 accumulation), so residuals
# and exponential error estimates stall there.
LOWP_TOL = {"fp32": 1e-6, "bf16x3": 1e-4}


def _floor_tol(A, tol):
    return max(tol, LOWP_TOL.get(getattr(A, "precision", "fp64"), 0.0))


@dataclass
class ConvergenceInfo:
    converged: int
    residual: float
    numiter: int
    numops: int


REORTH_ETA = 1.0 / np.sqrt(2.0)   # KrylovKit's ModifiedGramSchmidtIR default


def _orthogonalize(V, w):
    """w <- w - V V^H w (classical Gram-Schmidt), once more when the pass
    removed more than 1 - REORTH_ETA of the norm: the refinement rule of
    KrylovKit's ModifiedGramSchmidtIR (||w_1|| < eta ||w_0||, with ||w_0||^2 =
    ||w_1||^2 + sum |h|^2 for an orthonormal V).  Returns the coefficients."""
    h = T.multi_inner(V, w)
    T.multi_axpy_(V, -h, w)
    n1 = T.norm(w)
    if n1 * n1 < REORTH_ETA ** 2 * (n1 * n1 + float(np.sum(np.abs(h) ** 2))):
        h2 = T.multi_inner(V, w)
        T.multi_axpy_(V, -h2, w)
        h = h + h2
    return h


# ---------------------------------------------------------------------------
# Device-resident Lanczos expansion
# ---------------------------------------------------------------------------

_STREAMS = {}
USE_GRAPHS = True          # capture expansions as HIP graphs (False: same launches, eager)
MAX_DEVICE_BASIS = 32      # basis vectors the device driver handles (block combine limit)


@contextlib.contextmanager
def _krylov_stream():
    """Run on the engine's own (default-priority) stream: graph capture needs a
    stream other than the legacy default one."""
    import torch
    dev = torch.cuda.current_device()
    s = _STREAMS.get(dev)
    if s is None:
        s = _STREAMS[dev] = torch.cuda.Stream(dev)
    prev = torch.cuda.current_stream(dev)
    if prev == s:
        yield prev, s
        return
    s.wait_stream(prev)
    try:
        with torch.cuda.stream(s):
            yield prev, s
    finally:
        prev.wait_stream(s)


def _handoff(t, prev, s):
    """A tensor made on the Krylov stream and returned to the caller's stream."""
    if prev != s:
        t.data.record_stream(prev)
    return t


class _Graph:
    def __init__(self, handle):
        self.handle = handle

    def launch(self):
        nat.check(nat.lib().graft_graph_launch(nat.default_context().handle, self.handle))

    def close(self):
        if self.handle is not None and nat._lib is not None:
            nat.lib().graft_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _capture(issue):
    L = nat.lib()
    ctx = nat.default_context().handle
    nat.check(L.graft_graph_capture_begin(ctx))
    err = None
    try:
        issue()
    except Exception as e:       # close the capture before re-raising
        err = e
    h = ctypes.c_void_p()
    rc = L.graft_graph_capture_end(ctx, ctypes.byref(h))
    if err is not None:
        if h.value:
            L.graft_graph_destroy(h)
        raise err
    nat.check(rc)
    return _Graph(h)


class DeviceLanczos:
    """Fixed Krylov basis V_0..V_m of one map, with device-resident coefficients.

    Iteration j: V_{j+1} = A V_j; a CGS pass against V_0..V_j and a second one
    when the first cancelled most of the norm (``_orthogonalize``'s rule, decided
    on the device), coefficients in device memory; beta_j = ||V_{j+1}||;
    V_{j+1} /= beta_j.
    ``expand(j0, j1)`` queues iterations j0..j1-1 (replaying a captured graph
    when enabled) and returns their coefficients after one host read."""

    def __init__(self, A, like, m, use_graphs=None):
        import torch
        if m + 1 > MAX_DEVICE_BASIS:
            raise ValueError(f"device Lanczos keeps at most {MAX_DEVICE_BASIS} basis vectors")
        self.A, self.m = A, m
        self.V = [T.DTensor.empty(like.dims, like.ncod, like.blocks) for _ in range(m + 1)]
        self.n = like.numel
        self.ptrs = [v.ptr for v in self.V]
        self.stride = 4 * (m + 1) + 2            # [pass-1 coeffs | pass-2 coeffs | beta, 2nd pass]
        self.dcoef = torch.zeros(m * self.stride, dtype=torch.float64, device="cuda")
        self.hcoef = torch.empty(m * self.stride, dtype=torch.float64, pin_memory=True)
        self.use_graphs = USE_GRAPHS if use_graphs is None else use_graphs
        self.graphs = {}
        self.numops = 0

    def _issue(self, j0, j1):
        L = nat.lib()
        ctx = nat.default_context().handle
        n, m = self.n, self.m
        base0 = self.dcoef.data_ptr()
        for j in range(j0, j1):
            c1 = base0 + 8 * j * self.stride
            c2 = c1 + 8 * 2 * (m + 1)
            cb = c1 + 8 * 4 * (m + 1)
            self.A(self.V[j], out=self.V[j + 1])
            Vp = nat.ptr_array(self.ptrs[: j + 1])
            w = ctypes.c_void_p(self.ptrs[j + 1])
            nat.check(L.graft_vec_orthonormalize_dev(ctx, n, j + 1, Vp, w, REORTH_ETA,
                                                     ctypes.c_void_p(c1), ctypes.c_void_p(c2),
                                                     ctypes.c_void_p(cb)))

    def expand(self, j0, j1):
        """Run iterations j0..j1-1; returns (h[j] = CGS coefficients over V_0..V_j, beta[j])."""
        import torch
        if self.use_graphs:
            g = self.graphs.get((j0, j1))
            if g is None:
                # captured launches address the workspace as it is now
                nat.default_context().reserve_workspace(self.A.native.workspace_bytes)
                g = self.graphs[(j0, j1)] = _capture(lambda: self._issue(j0, j1))
            g.launch()
        else:
            self._issue(j0, j1)
        self.numops += j1 - j0
        lo, hi = j0 * self.stride, j1 * self.stride
        self.hcoef[lo:hi].copy_(self.dcoef[lo:hi], non_blocking=True)
        torch.cuda.current_stream().synchronize()
        raw = self.hcoef[lo:hi].numpy().reshape(j1 - j0, self.stride)
        m1 = self.m + 1
        hs, betas = [], []
        for r, j in enumerate(range(j0, j1)):
            a = raw[r, : 2 * m1].view(np.complex128)[: j + 1]
            b = raw[r, 2 * m1: 4 * m1].view(np.complex128)[: j + 1]
            hs.append(a + b)
            betas.append(float(raw[r, 4 * m1]))
        return hs, betas

    def start(self, x):
        """V_0 = x / ||x||; returns ||x||."""
        nrm = T.norm(x)
        if nrm == 0:
            raise ValueError("Krylov: zero starting vector")
        self.V[0].data.copy_(x.data)
        T.scale_(self.V[0], 1.0 / nrm)
        return nrm

    def combine(self, k_in, coef, outs):
        """outs[j] = sum_l coef[l, j] V_l over l < k_in (outs may be basis vectors)."""
        coef = np.asarray(coef, dtype=np.complex128)
        assert coef.shape == (k_in, len(outs))
        arr = np.ascontiguousarray(coef.T).reshape(-1)      # column-major k_in x p
        buf = (ctypes.c_double * (2 * arr.size)).from_buffer_copy(arr.view(np.float64).tobytes())
        nat.check(nat.lib().graft_vec_block_combine(
            nat.default_context().handle, self.n, k_in, nat.ptr_array(self.ptrs[:k_in]), len(outs),
            buf, nat.ptr_array([o.ptr for o in outs])))

    def close(self):
        for g in self.graphs.values():
            g.close()
        self.graphs.clear()


def _device_capable(A, x0, m):
    return (getattr(A, "native", None) is not None and m + 1 <= MAX_DEVICE_BASIS
            and x0.numel > 0)


def _eigsolve_device(A, x0, howmany, krylovdim, tol, maxiter):
    m = min(max(krylovdim, howmany + 2), x0.numel)
    with _krylov_stream() as (prev, s):
        L = DeviceLanczos(A, x0, m)
        try:
            L.start(x0)
            H = np.zeros((m + 1, m + 1), dtype=np.complex128)
            k = 0
            for it in range(1, maxiter + 1):
                j0 = k
                hs, betas = L.expand(j0, m)
                invariant = m == x0.numel      # the basis spans the whole space
                k = m
                for r, j in enumerate(range(j0, m)):
                    H[: j + 1, j] = hs[r]
                    H[j + 1, j] = betas[r]
                    if betas[r] < BREAKDOWN * max(1.0, np.abs(H[: j + 2, : j + 1]).max()):
                        k, invariant = j + 1, True
                        break
                beta = float(H[k, k - 1].real)
                Hk = H[:k, :k]
                Hk = (Hk + Hk.conj().T) / 2
                theta, U = np.linalg.eigh(Hk)
                last = U[k - 1, :]
                res = np.abs(beta * last[:howmany])
                resid = float(res.max())
                ok = res <= tol
                if bool(np.all(ok)) or invariant or it == maxiter:
                    vecs = [T.DTensor.empty(x0.dims, x0.ncod, x0.blocks) for _ in range(howmany)]
                    L.combine(k, U[:, :howmany], vecs)
                    vecs = [_handoff(v, prev, s) for v in vecs]
                    nconv = howmany if invariant else int(np.sum(ok))
                    return list(theta[:howmany]), vecs, ConvergenceInfo(nconv, resid, it, L.numops)
                # thick restart in place: V_i <- V U[:, i] (i < p), V_p <- V_k;
                # p follows KrylovKit's Lanczos eigsolve: (3 m + 2 converged) / 5
                p = max(howmany, min(k - 1, (3 * m + 2 * int(np.sum(ok))) // 5))
                coef = np.zeros((k + 1, p + 1), dtype=np.complex128)
                coef[:k, :p] = U[:, :p]
                coef[k, p] = 1.0
                L.combine(k + 1, coef, L.V[: p + 1])
                H[:] = 0
                H[np.arange(p), np.arange(p)] = theta[:p]
                H[p, :p] = beta * last[:p]
                H[:p, p] = np.conj(beta * last[:p])
                k = p
        finally:
            L.close()
    raise RuntimeError("unreachable")


def eigsolve(A, x0, howmany=1, which="SR", ishermitian=True, krylovdim=20, tol=1e-12,
             maxiter=100, verbosity=0, device=None):
    """Lowest eigenpairs of a hermitian map by thick-restart Lanczos (Krylov–Schur).

    Returns (vals, vecs, info) like KrylovKit.eigsolve.  Convergence: residual
    norm of each requested Ritz pair <= tol (absolute, as KrylovKit's Lanczos
    eigsolve compares it; the reference calls it with the default 1e-12,
    GroundState.jl:28-39).  Native maps use
    the device-resident driver (``device=False`` forces the per-call path)."""
    if which != "SR" or not ishermitian:
        raise NotImplementedError("eigsolve supports the hermitian :SR problem used by dmrg2!")
    tol = _floor_tol(A, tol)
    m = min(max(krylovdim, howmany + 2), x0.numel)
    if m < howmany:
        raise ValueError("eigsolve: more eigenpairs requested than the space has dimensions")
    if device is None:
        device = _device_capable(A, x0, m)
    if device:
        return _eigsolve_device(A, x0, howmany, krylovdim, tol, maxiter)
    nrm = T.norm(x0)
    if nrm == 0:
        raise ValueError("eigsolve: zero starting vector")
    v = x0.clone()
    T.scale_(v, 1.0 / nrm)
    V = [v]
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
            if beta < BREAKDOWN * max(1.0, np.abs(H[: k + 1, :k]).max()):
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
        converged = bool(np.all(res <= tol))
        # breakdown (no residual vector appended) or a basis of the whole space
        invariant = len(V) == k or k == x0.numel
        if converged or invariant or it == maxiter:
            vecs = [T.lincomb(V[:k], U[:, j], V[0]) for j in range(howmany)]
            nconv = int(np.sum(res <= tol)) if not invariant else howmany
            return (list(theta[:howmany]), vecs,
                    ConvergenceInfo(nconv, resid, it, numops))
        # thick restart: keep p lowest Ritz vectors plus the residual direction,
        # p = (3 m + 2 converged) / 5 as in KrylovKit's Lanczos eigsolve
        nok = int(np.sum(res <= tol))
        p = max(howmany, min(k - 1, (3 * m + 2 * nok) // 5))
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


EXP_CHUNK = 4     # iterations per device expansion between error checks (eager mode)


def _exponentiate_device(A, dz, x, krylovdim, tol, eager, maxiter):
    """Device-resident form of the loop below: the same Krylov spaces, step
    sizes and acceptance decisions; the expansion runs in chunks of EXP_CHUNK
    iterations (captured graphs) and the error estimate is evaluated for every
    dimension of a chunk, so the accepted dimension is the one the per-
    iteration loop picks.  The propagated vector is combined into V_0 in place."""
    total = complex(dz)
    m = krylovdim
    whole = m == x.numel           # the Krylov space can reach the whole space: exact there
    with _krylov_stream() as (prev, s):
        L = DeviceLanczos(A, x, m)
        try:
            nrm = L.start(x)
            done, frac, numiter, worst = 0.0, 1.0, 0, 0.0
            while done < 1.0 - 1e-15:
                frac = min(frac, 1.0 - done)
                H = np.zeros((m + 1, m + 1), dtype=np.complex128)
                k, accepted, stop = 0, False, False
                while not stop:
                    j1 = min(m, k + EXP_CHUNK) if eager else m
                    hs, betas = L.expand(k, j1)
                    for r, j in enumerate(range(k, j1)):
                        H[: j + 1, j] = hs[r]
                        beta = betas[r]
                        kk = j + 1
                        scale = max(1.0, np.abs(H[:kk, :kk]).max(), np.abs(beta))
                        breakdown = beta < BREAKDOWN * scale or (whole and kk == m)
                        E = _expm_hermitian_small(H[:kk, :kk], frac * total)
                        err = nrm * abs(beta * E[kk - 1, 0])
                        if breakdown or (eager and err <= tol * frac) or kk == m:
                            if breakdown or err <= tol * frac:
                                L.combine(kk, (nrm * E[:, 0])[:, None], [L.V[0]])
                                done += frac
                                worst = max(worst, err)
                                accepted = True
                            stop = True
                            break
                        H[kk, kk - 1] = beta
                    k = j1
                numiter += 1
                if accepted:
                    if done < 1.0 - 1e-15:
                        nrm = T.norm(L.V[0])
                        if nrm == 0:
                            break
                        T.scale_(L.V[0], 1.0 / nrm)
                else:
                    frac /= 2.0
                    if numiter > maxiter:
                        raise RuntimeError(f"exponentiate: step size underflow (dim {x.numel}, "
                                           f"krylovdim {m}, last error {err:.3e}, beta {beta:.3e}, "
                                           f"fraction {frac:.3e})")
            y = _handoff(L.V[0].clone(), prev, s)
            return y, ConvergenceInfo(1, worst, numiter, L.numops)
        finally:
            L.close()


def exponentiate(A, dz, x, ishermitian=True, krylovdim=30, tol=1e-12, eager=True, maxiter=100,
                 device=None):
    """exp(dz * A) x by Lanczos with adaptive sub-stepping (KrylovKit.exponentiate semantics).

    Returns (y, info).  The local error estimate of a step of size tau is
    ||x|| * |beta_k * [exp(tau T_k) e_1]_k|; a step is accepted when the
    estimate is <= tol * |tau| / |dz|, else tau is halved.  A map with
    reduced-precision products (FP32, split BF16x3) cannot meet a tolerance
    below its products' accuracy: ``tol`` is raised to LOWP_TOL for it."""
    if not ishermitian:
        raise NotImplementedError("exponentiate supports hermitian maps (Lanczos)")
    total = complex(dz)
    if total == 0:
        return x.clone(), ConvergenceInfo(1, 0.0, 0, 0)
    tol = _floor_tol(A, tol)
    krylovdim = min(krylovdim, x.numel)
    whole = krylovdim == x.numel
    if device is None:
        device = _device_capable(A, x, krylovdim)
    if device:
        return _exponentiate_device(A, dz, x, krylovdim, tol, eager, maxiter)
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
            breakdown = (beta < BREAKDOWN * max(1.0, np.abs(H[:k, :k]).max())
                         or (whole and k == krylovdim))
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
