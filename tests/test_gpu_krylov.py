"""GPU checks of the device-resident Krylov driver (capturable vector kernels,
in-place basis recombination, HIP-graph replays of the Lanczos expansion).

The device driver must reproduce the per-call driver (same Krylov spaces and
decisions) to 1e-12, dense linear algebra on the map's matrix to 1e-10, and a
replayed graph must give bit-identical coefficients to the eager launches.
"""
import ctypes

import numpy as np
import pytest

from oracle import heff_oracle as ref

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def g():
    import torch
    import graft_amd
    from graft_amd import _native
    torch.cuda.set_device(0)
    _native.lib()
    return graft_amd


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def _h2_map(D=8, seed=3):
    from graft_amd import contractions as C, networks as N, operators as O, trees as Tr
    topo = Tr.binary_topology(3)
    rng = np.random.default_rng(seed)
    psi = N.random_ttns(rng, topo, 2, D)
    H = O.heisenberg_ttno(topo, 1.0)
    cache = C.EnvCache(topo)
    h2 = C.eff_h2(cache, psi, H, 1, 0)
    return h2, C.two_site_tensor(psi, 1, 0)


def _dense_matrix(A, like):
    from graft_amd.tensors import DTensor
    n = int(np.prod(like.dims))
    M = np.zeros((n, n), dtype=complex)
    e = np.zeros(n, dtype=complex)
    for i in range(n):
        e[:] = 0
        e[i] = 1
        M[:, i] = A(DTensor.from_numpy(e.reshape(like.dims, order="F"), like.ncod)).to_numpy().reshape(-1, order="F")
    return M


@pytest.mark.parametrize("k", [7, 21])
def test_device_vector_ops(g, k):
    """k = 21 takes two passes of the multi-dot kernel (16 vectors per pass)."""
    import torch
    from graft_amd import _native as nat, krylov as K, tensors as T
    from graft_amd.tensors import DTensor
    rng = np.random.default_rng(5)
    n = 70001
    _say("device vector ops", n, k)
    xs = [ref.random_complex(rng, n) for _ in range(k + 1)]
    ds = [DTensor.from_numpy(x) for x in xs]
    L = nat.lib()
    ctx = nat.default_context().handle
    dv = torch.zeros(2 * k + 2, dtype=torch.float64, device="cuda")
    V = nat.ptr_array([d.ptr for d in ds[:k]])
    y = ctypes.c_void_p(ds[k].ptr)
    nat.check(L.graft_vec_multi_dotc_dev(ctx, n, k, V, y, ctypes.c_void_p(dv.data_ptr())))
    got = dv[: 2 * k].cpu().numpy().view(np.complex128)
    want = np.array([np.vdot(x, xs[k]) for x in xs[:k]])
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
    # y -= V c with c read from device memory
    nat.check(L.graft_vec_multi_axpy_dev(ctx, n, k, V, ctypes.c_void_p(dv.data_ptr()), -1.0, y))
    ref_y = xs[k] - sum(c * x for c, x in zip(got, xs[:k]))
    assert _rel(ds[k].to_numpy(), ref_y) < 1e-13
    nat.check(L.graft_vec_nrm2_dev(ctx, n, y, ctypes.c_void_p(dv.data_ptr() + 8 * 2 * k)))
    nrm = float(dv[2 * k].item())
    assert abs(nrm - np.linalg.norm(ref_y)) < 1e-12 * nrm
    nat.check(L.graft_vec_scale_inv_dev(ctx, n, ctypes.c_void_p(dv.data_ptr() + 8 * 2 * k), y))
    assert _rel(ds[k].to_numpy(), ref_y / nrm) < 1e-13
    # in-place block combine: outputs alias the first p + 1 inputs
    basis = [DTensor.from_numpy(x) for x in xs[:k]]
    host = [b.to_numpy() for b in basis]
    coef = rng.standard_normal((k, 4)) + 1j * rng.standard_normal((k, 4))
    lz = K.DeviceLanczos.__new__(K.DeviceLanczos)
    lz.n, lz.ptrs = n, [b.ptr for b in basis]
    lz.combine(k, coef, basis[:4])
    for j in range(4):
        assert _rel(basis[j].to_numpy(), sum(coef[l, j] * host[l] for l in range(k))) < 1e-13
    for j in range(4, k):
        assert np.array_equal(basis[j].to_numpy(), host[j])


def test_graph_replays_match_eager_launches(g):
    """A captured expansion replayed three times, with eager work and host
    synchronizations between the replays, matches the eager launches bit for bit."""
    import torch
    from graft_amd import krylov as K
    h2, x = _h2_map()
    _say("graph replays", h2.xdims)
    h2(x)                                    # the plan's workspace exists before capture
    with K._krylov_stream():
        eager = K.DeviceLanczos(h2, x, 8, use_graphs=False)
        eager.start(x)
        h_e, b_e = eager.expand(0, 6)
        v_e = eager.V[6].to_numpy()
        graph = K.DeviceLanczos(h2, x, 8, use_graphs=True)
        try:
            for r in range(3):
                graph.start(x)               # eager work between replays
                torch.cuda.current_stream().synchronize()
                h_g, b_g = graph.expand(0, 6)
                assert len(graph.graphs) == 1
                assert b_g == b_e, r
                for a, b in zip(h_g, h_e):
                    assert np.array_equal(a, b), r
                assert np.array_equal(graph.V[6].to_numpy(), v_e), r
        finally:
            graph.close()
    # the expansion is a Lanczos process: orthonormal basis, tridiagonal projection
    Vm = np.stack([v.to_numpy().reshape(-1) for v in eager.V[:7]], axis=1)
    assert np.abs(Vm.conj().T @ Vm - np.eye(7)).max() < 1e-12
    for j, h in enumerate(h_e):
        assert np.abs(h[: max(0, j - 1)]).max(initial=0.0) < 1e-10


def test_graph_launch_refused_after_workspace_moves(g):
    from graft_amd import _native as nat, krylov as K
    h2, x = _h2_map()
    h2(x)
    _say("graph workspace guard")
    with K._krylov_stream():
        lz = K.DeviceLanczos(h2, x, 4, use_graphs=True)
        lz.start(x)
        lz.expand(0, 2)
        nat.default_context().release_workspace()
        with pytest.raises(nat.GraftError, match="workspace moved"):
            lz.graphs[(0, 2)].launch()
        lz.close()


def test_device_eigsolve_matches_host_driver_and_dense(g):
    from graft_amd import krylov as K, tensors as T
    h2, x = _h2_map()
    _say("device eigsolve", h2.xdims)
    M = _dense_matrix(h2, x)
    assert np.abs(M - M.conj().T).max() < 1e-11 * np.abs(M).max()
    e0 = np.linalg.eigvalsh(M)[0]
    vd, xd, infod = K.eigsolve(h2, x, 1, "SR", krylovdim=20, tol=1e-12, device=True)
    vh, xh, infoh = K.eigsolve(h2, x, 1, "SR", krylovdim=20, tol=1e-12, device=False)
    assert abs(vd[0] - e0) < 1e-10 * abs(e0)
    assert abs(vd[0] - vh[0]) < 1e-12 * abs(e0)
    assert infod.numiter == infoh.numiter
    ov = abs(T.inner(xd[0], xh[0]))
    assert abs(ov - 1) < 1e-10
    r = M @ xd[0].to_numpy().reshape(-1, order="F") - vd[0] * xd[0].to_numpy().reshape(-1, order="F")
    assert np.linalg.norm(r) < 1e-9


def test_device_exponentiate_matches_host_driver_and_dense(g):
    from graft_amd import krylov as K
    h2, x = _h2_map()
    _say("device exponentiate", h2.xdims)
    M = _dense_matrix(h2, x)
    x0 = x.to_numpy().reshape(-1, order="F")
    w, U = np.linalg.eigh((M + M.conj().T) / 2)
    for dz in (-0.05j, -0.7j, -0.3):
        want = U @ (np.exp(dz * w) * (U.conj().T @ x0))
        yd, infod = K.exponentiate(h2, dz, x, krylovdim=30, tol=1e-12, device=True)
        yh, infoh = K.exponentiate(h2, dz, x, krylovdim=30, tol=1e-12, device=False)
        assert _rel(yd.to_numpy(), yh.to_numpy()) < 1e-12, dz
        assert _rel(yd.to_numpy().reshape(-1, order="F"), want) < 1e-10, dz
        assert infod.numiter == infoh.numiter


@pytest.mark.parametrize("shape,decay", [((600, 900), 25.0), ((1100, 700), 25.0),
                                         ((700, 2000), 40.0 / np.log(10.0))])
def test_gram_route_svd_matches_dense_svd(g, shape, decay):
    """The Gram-route truncated SVD against LAPACK on matrices whose spectra
    decay over 11 to 15 decades: every singular value >= 1e-5 sigma_1 to 1e-10
    relative (the truncation spectrum), all of them to 1e-13 sigma_1
    absolute, the kept values to 1e-12, the discarded norm, and the truncated
    reconstruction; cuts by maxdim, by atol and by discarded weight that fall
    deep inside the tail keep LAPACK's rank."""
    import torch
    from graft_amd import backend as B
    rows, cols = shape
    _say("gram svd", shape, decay)
    rng = np.random.default_rng(rows)
    n = min(shape)
    U0, _ = np.linalg.qr(ref.random_complex(rng, (rows, n)))
    V0, _ = np.linalg.qr(ref.random_complex(rng, (cols, n)))
    sig = np.exp(-np.arange(n) / decay)
    M = (U0 * sig) @ V0.conj().T
    Md = torch.from_numpy(M).cuda()
    s_ref = np.linalg.svd(M, compute_uv=False)
    ts = B.TruncationScheme(maxdim=64)
    U, S, Vh, disc, spec = B.split_svd(Md, ts, method="gram")
    assert U.shape == (rows, 64) and Vh.shape == (64, cols)
    big = s_ref >= 1e-5 * s_ref[0]
    assert np.abs(spec[big] / s_ref[big] - 1).max() < 1e-10
    assert np.abs(spec - s_ref).max() < 1e-13 * s_ref[0]
    assert np.abs(S.cpu().numpy() - s_ref[:64]).max() < 1e-12 * s_ref[0]
    assert abs(disc - np.sqrt(np.sum(s_ref[64:] ** 2))) < 1e-10 * max(disc, 1e-300)
    low = (U * S) @ Vh
    want = (U0[:, :64] * sig[:64]) @ V0[:, :64].conj().T
    assert np.abs(low.cpu().numpy() - want).max() < 1e-11
    Ud, Sd, Vhd, discd, _ = B.split_svd(Md, ts, method="direct")
    assert np.abs(Sd.cpu().numpy() - S.cpu().numpy()).max() < 1e-12
    # cuts inside the tail: thresholds midway (geometrically) between two values
    mid = lambda i: float(np.sqrt(sig[i] * sig[i + 1]))
    for cut in (B.TruncationScheme(atol=mid(150)),
                B.TruncationScheme(maxdim=n, rtol=mid(120) / np.linalg.norm(sig)),
                B.TruncationScheme(discarded_weight=float(np.sqrt(np.sum(sig[100:] ** 2))
                                                          / np.linalg.norm(sig)) * 1.0001)):
        r_ref = B.truncation_rank(s_ref, cut)
        r = B.truncation_rank(spec, cut)
        assert r == r_ref, (cut, r, r_ref)
        Ug, Sg, _, dg, _ = B.split_svd(Md, cut, method="gram")
        assert Ug.shape[1] == r_ref
        assert abs(dg - np.sqrt(np.sum(s_ref[r_ref:] ** 2))) < 1e-10 * max(dg, 1e-300)


@pytest.mark.parametrize("shape", [(7, 5, 9), (300, 40000, 48)])
def test_library_matmul_matches_torch(g, shape):
    """backend.matmul (the library GEMM, split-K for the long contraction of
    the second shape) against torch for every op pair, on contiguous and
    transposed-view operands."""
    import torch
    from graft_amd import backend as B
    m, k, n = shape
    _say("library matmul", shape)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(m)
    rnd = lambda r, c: torch.randn(r, c, dtype=torch.complex128, device="cuda", generator=gen)
    ops = {"N": lambda M: M, "T": lambda M: M.T, "C": lambda M: M.conj(), "H": lambda M: M.conj().T}
    for ox in "NTCH":
        for oy in "NTCH":
            X = rnd(m, k) if ox in "NC" else rnd(k, m)
            Y = rnd(k, n) if oy in "NC" else rnd(n, k)
            for view in (False, True):
                Xv = X.T.contiguous().T if view else X        # same values, transposed storage
                got = B.matmul(Xv, Y, ox, oy)
                want = ops[ox](X) @ ops[oy](Y)
                err = float((got - want).abs().max() / want.abs().max())
                assert got.shape == (m, n) and err < 1e-13, (ox, oy, view, err)


@pytest.mark.parametrize("case", ["random", "cond1e12", "rank_deficient", "empty_blocks"])
def test_tall_qr(g, case):
    """Shifted Cholesky QR (three passes) on tall matrices: orthonormal Q and
    Q R = M to 1e-12 relative for a random matrix and one with singular values
    down to 1e-12; a numerically rank-deficient matrix, and one whose rank
    deficiency comes from exactly empty row blocks (the dense form of a tensor
    with charge sectors of fewer rows than columns), take the Householder
    fallback and satisfy the same bounds."""
    import torch
    from graft_amd import backend as B
    m, n = 40000, 96
    _say("tall qr", case)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(17)
    rnd = lambda r, c: torch.randn(r, c, dtype=torch.complex128, device="cuda", generator=gen)
    M = rnd(m, n)
    if case != "random":
        U, _ = torch.linalg.qr(rnd(m, n))
        V, _ = torch.linalg.qr(rnd(n, n))
        sig = torch.logspace(0, -12, n, dtype=torch.float64, device="cuda")
        if case == "rank_deficient":
            sig[n // 2:] = 0.0
        M = (U * sig) @ V.conj().T
    if case == "empty_blocks":
        M = torch.zeros(m, n, dtype=torch.complex128, device="cuda")
        rows, cols = m // 12, n // 12
        for b in range(12):
            r = cols // 2 if b % 3 == 0 else rows           # every third block: fewer rows than columns
            M[b * rows:b * rows + r, b * cols:(b + 1) * cols] = rnd(r, cols)
    Q, R = B.tall_qr(M)
    eye = torch.eye(n, dtype=torch.complex128, device="cuda")
    assert float((Q.conj().T @ Q - eye).abs().max()) < 1e-12 * n
    assert float((Q @ R - M).abs().max()) < 1e-12 * float(M.abs().max())
    assert float(torch.tril(R, -1).abs().max()) == 0.0


@pytest.mark.parametrize("near_span", [False, True])
def test_device_orthonormalize(g, near_span):
    """graft_vec_orthonormalize_dev: a generic vector takes one Gram-Schmidt
    pass (second-pass flag 0, zero second coefficients); a vector within 1e-8
    of span(V) takes the second pass (flag 1).  Either way the result is a unit
    vector orthogonal to V to 1e-12 and w_0 = V (c1 + c2) + beta w."""
    import torch
    from graft_amd import _native as nat, krylov as K
    from graft_amd.tensors import DTensor
    rng = np.random.default_rng(23)
    n, k = 50001, 9
    _say("device orthonormalize", near_span)
    Q, _ = np.linalg.qr(ref.random_complex(rng, (n, k)))
    c = ref.random_complex(rng, k)
    w0 = Q @ c + (1e-8 if near_span else 1.0) * ref.random_complex(rng, n)
    V = [DTensor.from_numpy(np.ascontiguousarray(Q[:, j])) for j in range(k)]
    w = DTensor.from_numpy(w0)
    out = torch.zeros(4 * k + 2, dtype=torch.float64, device="cuda")
    p = out.data_ptr()
    nat.check(nat.lib().graft_vec_orthonormalize_dev(
        nat.default_context().handle, n, k, nat.ptr_array([v.ptr for v in V]), ctypes.c_void_p(w.ptr),
        K.REORTH_ETA, ctypes.c_void_p(p), ctypes.c_void_p(p + 16 * k), ctypes.c_void_p(p + 32 * k)))
    o = out.cpu().numpy()
    c1, c2 = o[: 2 * k].view(np.complex128), o[2 * k: 4 * k].view(np.complex128)
    beta, flag = o[4 * k], o[4 * k + 1]
    assert flag == (1.0 if near_span else 0.0)
    if not near_span:
        assert np.all(c2 == 0)
    wn = w.to_numpy()
    assert abs(np.linalg.norm(wn) - 1) < 1e-13
    assert np.abs(Q.conj().T @ wn).max() < 1e-12
    assert np.linalg.norm(Q @ (c1 + c2) + beta * wn - w0) < 1e-13 * np.linalg.norm(w0)
