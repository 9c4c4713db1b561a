"""GPU parity of the U(1) x U(1) sector path (grouped block GEMM launches).

The sector engine runs the spinful Hubbard model; every check compares it
with the oracle on the dense embedding of the same tensors, with exact
diagonalization in the charge sector, or with the plain dense GPU form.
"""
import numpy as np
import pytest

from oracle import heff_oracle as ref

pytestmark = pytest.mark.gpu

TOL = 1e-12


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


def _tree(topo):
    return ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])


def _dense_state(psi):
    return [t.to_numpy() for t in psi.tensors]


def test_sector_envs_and_maps_match_oracle(g):
    from graft_amd import contractions as C, operators as O, sectors as S, trees as Tr
    topo = Tr.star_topology(3, 1)
    t = _tree(topo)
    _say("sector maps star(3,1) hubbard")
    rng = np.random.default_rng(3)
    psi, host = S.random_hubbard_ttns(rng, topo, 12)
    H = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    Wd = O.hubbard_host(topo, 1.0, 4.0)
    dense = _dense_state(psi)
    envs = ref.all_envs(t, dense, Wd)
    cache = C.EnvCache(topo)
    for (u, v), E in envs.items():
        got = C.env_(cache, psi, H, u, v)
        assert got.blocks is not None
        assert _rel(got.to_numpy(), E) < TOL, (u, v)
    h1 = C.eff_h1(cache, psi, H, 0)
    assert _rel(h1(psi.tensors[0]).to_numpy(), ref.h1_apply(t, Wd, envs, 0, dense[0])) < TOL
    assert abs(C.expect(psi, H, cache) - ref.env_closure(t, dense, Wd, envs, 0)) < 1e-12
    theta = C.two_site_tensor(psi, 1, 0)
    theta_ref = ref.two_site_tensor(t, dense, 1, 0)
    assert _rel(theta.to_numpy(), theta_ref) < TOL
    h2 = C.eff_h2(cache, psi, H, 1, 0, out_structure=theta.blocks)
    assert _rel(h2(theta).to_numpy(), ref.h2_apply(t, Wd, envs, 1, 0, theta_ref)) < TOL
    info = h2.native.step_info()
    assert max(s["blocks"] for s in info) > 1          # genuinely grouped launches


def test_sector_dmrg2_matches_sector_exact_diagonalization(g):
    from graft_amd import backend as B, groundstate as GS, operators as O, sectors as S, trees as Tr
    topo = Tr.star_topology(3, 1)
    t = _tree(topo)
    _say("sector dmrg2 star(3,1) hubbard U=4")
    rng = np.random.default_rng(4)
    psi, _ = S.random_hubbard_ttns(rng, topo, 4)
    H = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    Hd, c = ref.hubbard_dense(t, 1.0, 4.0)
    N, Sz2 = ref.number_operators(c, t.n)
    E0 = ref.sector_ground_energy(Hd, N, Sz2, t.n, 0)
    _, Es = GS.dmrg2_(psi, H, trunc=B.TruncationScheme(maxdim=256), nsweeps=6)
    assert abs(Es[-1] - E0) < 1e-10, (Es, E0)


def _full_vector(t, psi):
    return ref.ttns_to_dense(t, _dense_state(psi), list(psi.hasphys))


def test_sector_tdvp1_is_exact_at_full_bonds(g):
    """TDVP1 on block-sparse states (tdvp.jl:171-301, link splits by per-sector
    QR).  On star(3,1) every leaf bond holds the leaf's whole local space, so a
    leaf step and the backward link step after it cancel and the root step is
    the full propagator: TDVP1 of order 1 and 2 equals exp(dz H) on the
    half-filling sector to the Krylov tolerance.  H is the dense matrix of the
    same TTNO (ttno_to_dense), the states their dense embeddings."""
    from graft_amd import evolution as EV, operators as O, sectors as S, trees as Tr
    topo = Tr.star_topology(3, 1)
    t = _tree(topo)
    _say("sector TDVP1 star(3,1) hubbard, full bonds, vs exact propagation")
    rng = np.random.default_rng(11)
    psi, _ = S.random_hubbard_ttns(rng, topo, 4)
    H = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    Hd = ref.ttno_to_dense(t, O.hubbard_host(topo, 1.0, 4.0))
    v0 = _full_vector(t, psi)
    vex = ref.exact_evolve(Hd, v0, -0.15j)
    assert np.linalg.norm(vex - v0) > 1e-2             # the evolution moves the state
    for order in (1, 2):
        p = psi.copy()
        ev = EV.TDVP1(order=order)
        for _ in range(3):
            EV.step_(ev, p, H, -0.05j)
        assert all(a.blocks is not None for a in p.tensors)
        err = np.linalg.norm(_full_vector(t, p) - vex)
        assert err < 1e-9, (order, err)


def test_sector_tdvp1_matches_dense_tdvp1_and_conserves(g):
    """Truncated bonds (7-node binary tree, D=16 of 64 subtree states): every
    per-sector link QR is full rank, so the dense TDVP1 on the embedded tensors
    and the sector TDVP1 give one physical state; both conserve the norm and
    the energy, each sub-step being the exponential of a hermitian map."""
    from graft_amd import contractions as C, evolution as EV, networks as N, operators as O
    from graft_amd import sectors as S, trees as Tr
    from graft_amd.tensors import DTensor
    topo = Tr.binary_topology(2)
    t = _tree(topo)
    _say("sector TDVP1 vs dense TDVP1, binary(2) hubbard, D=16")
    rng = np.random.default_rng(12)
    psi, _ = S.random_hubbard_ttns(rng, topo, 16, sigma=1.6)
    assert max(psi.bond_dims()) == 16
    Hs = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    Wd = O.hubbard_host(topo, 1.0, 4.0)
    Hd = N.TTNO(topo, [DTensor.from_numpy(w, w.ndim - 2) for w in Wd], ishermitian=True)
    psi_d = N.TTNS(topo, [DTensor.from_numpy(a, a.ndim - 1) for a in _dense_state(psi)], psi.center)
    E0 = C.expect(psi, Hs, C.EnvCache(topo)).real
    evs, evd = EV.TDVP1(order=2), EV.TDVP1(order=2)
    for _ in range(3):
        EV.step_(evs, psi, Hs, -0.05j)
        EV.step_(evd, psi_d, Hd, -0.05j)
    vs, vd = _full_vector(t, psi), _full_vector(t, psi_d)
    assert np.linalg.norm(vs - vd) < 1e-9, np.linalg.norm(vs - vd)
    assert abs(np.linalg.norm(vs) - 1) < 1e-10
    E1 = C.expect(psi, Hs, C.EnvCache(topo)).real
    assert abs(E1 - E0) < 1e-9 * abs(E0), (E0, E1)


def test_sector_and_dense_forms_agree_at_scale(g):
    """16-site tree, Hubbard U=4 at half filling: the grouped sector apply and the
    plain dense FP64 apply on the embedded tensors agree, and so do their
    Lanczos ground energies (1e-10 relative)."""
    from graft_amd import contractions as C, krylov as K, networks as N, operators as O, sectors as S
    from graft_amd import trees as Tr
    from graft_amd.tensors import DTensor
    edges = [(str((i - 1) // 2), str(i)) for i in range(1, 16)]
    topo = Tr.TreeTopology("0", edges)
    _say("sector vs dense, 16-site tree, D=64")
    rng = np.random.default_rng(5)
    psi, host = S.random_hubbard_ttns(rng, topo, 64, sigma=1.6)
    Hs = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    Wd = O.hubbard_host(topo, 1.0, 4.0)
    Hd = N.TTNO(topo, [DTensor.from_numpy(w, w.ndim - 2) for w in Wd], ishermitian=True)
    psi_d = N.TTNS(topo, [DTensor.from_numpy(a, a.ndim - 1) for a in _dense_state(psi)], psi.center)
    n, m = 1, 0
    cs, cd = C.EnvCache(topo), C.EnvCache(topo)
    theta = C.two_site_tensor(psi, n, m)
    theta_d = C.two_site_tensor(psi_d, n, m)
    hs = C.eff_h2(cs, psi, Hs, n, m, out_structure=theta.blocks)
    hd = C.eff_h2(cd, psi_d, Hd, n, m)
    ys = hs(theta).to_numpy()
    yd = hd(theta_d).to_numpy()
    assert _rel(ys, yd) < 1e-12
    vs, _, _ = K.eigsolve(hs, theta, 1, "SR", krylovdim=20, tol=1e-12)
    vd, _, _ = K.eigsolve(hd, theta_d, 1, "SR", krylovdim=20, tol=1e-12)
    assert abs(vs[0] - vd[0]) < 1e-10 * abs(vd[0])


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_sector_fusion_modes_agree(g, mode):
    """Per-block GEMM items (mode 0), fused per-charge matrices (mode 2) and
    per-block streaming for every multi-block step (mode 3: both operand
    sides, conjugated bras, more than 16 small-side elements per block) give
    the oracle's envs and maps; at D=48 on the 16-site tree every mode agrees
    with the automatic choice."""
    from graft_amd import _native as nat, contractions as C, operators as O, sectors as S, trees as Tr
    ctx = nat.default_context()
    try:
        ctx.set_sector_fusion(mode)
        topo = Tr.star_topology(3, 1)
        t = _tree(topo)
        _say("sector fusion mode", mode, "star(3,1)")
        rng = np.random.default_rng(3)
        psi, host = S.random_hubbard_ttns(rng, topo, 12)
        H = O.hubbard_sector_ttno(topo, 1.0, 4.0)
        Wd = O.hubbard_host(topo, 1.0, 4.0)
        dense = _dense_state(psi)
        envs = ref.all_envs(t, dense, Wd)
        cache = C.EnvCache(topo)
        for (u, v), E in envs.items():
            assert _rel(C.env_(cache, psi, H, u, v).to_numpy(), E) < TOL, (u, v)
        theta = C.two_site_tensor(psi, 1, 0)
        h2 = C.eff_h2(cache, psi, H, 1, 0, out_structure=theta.blocks)
        want = ref.h2_apply(t, Wd, envs, 1, 0, ref.two_site_tensor(t, dense, 1, 0))
        assert _rel(h2(theta).to_numpy(), want) < TOL
        # 16-site tree: this mode against the automatic one
        edges = [(str((i - 1) // 2), str(i)) for i in range(1, 16)]
        topo = Tr.TreeTopology("0", edges)
        _say("sector fusion mode", mode, "16-site D=48")
        rng = np.random.default_rng(8)
        psi, _ = S.random_hubbard_ttns(rng, topo, 48, sigma=1.6)
        Hs = O.hubbard_sector_ttno(topo, 1.0, 4.0)
        theta = C.two_site_tensor(psi, 1, 0)
        y_mode = C.eff_h2(C.EnvCache(topo), psi, Hs, 1, 0, out_structure=theta.blocks)(theta).to_numpy()
        ctx.set_sector_fusion(1)
        y_auto = C.eff_h2(C.EnvCache(topo), psi, Hs, 1, 0, out_structure=theta.blocks)(theta).to_numpy()
        assert _rel(y_mode, y_auto) < 1e-13
    finally:
        ctx.set_sector_fusion(1)


def test_sector_cbe_split_directions_match_dense_core(g):
    """One sector CBE link split (the reference's exact sector-aware selector,
    cbe.jl:459-463): U R reproduces the active tensor, U is isometric, and the
    spawned directions span the leading left singular space of the dense
    double-complement core the oracle forms from the two-site map on the
    embedded tensors."""
    import torch
    from graft_amd import backend as B, contractions as C, evolution as EV, networks as N
    from graft_amd import operators as O, sectors as S, trees as Tr
    topo = Tr.binary_topology(2)
    t = _tree(topo)
    _say("sector CBE split vs dense core, binary(2) hubbard, D=8")
    psi, _ = S.random_hubbard_ttns(np.random.default_rng(21), topo, 8, sigma=1.6)
    Hs = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    Wd = O.hubbard_host(topo, 1.0, 4.0)
    n, m = 1, 0
    cache = C.EnvCache(topo)
    N.move_center_(psi, n, cache)
    ev = EV.TDVP1_CBE(trunc=B.TruncationScheme(maxdim=64),
                      cbe=EV.PredictorCBE(max_add=8, neigs=3, spawn_threshold=1e-12))
    ev.cache = cache
    dense = _dense_state(psi)
    A = dense[n]
    U, R = EV._cbe_split(ev, psi, Hs, n, m, -0.05j)
    assert ev.last_cbe_info["selected"] == 3
    Ud, Rd = U.to_numpy(), R.to_numpy()
    Um = Ud.reshape(-1, Ud.shape[-1], order="F")
    assert np.abs(Um.conj().T @ Um - np.eye(Um.shape[1])).max() < 1e-12
    rec = np.tensordot(Ud, Rd, axes=([Ud.ndim - 1], [0]))
    assert _rel(rec, A) < 1e-12
    # oracle: two-site map on Theta, double complement, leading left vectors
    envs = ref.all_envs(t, dense, Wd)
    theta = ref.two_site_tensor(t, dense, n, m)
    Y = ref.h2_apply(t, Wd, envs, n, m, theta)
    pn = A.ndim - 1
    Ym = Y.reshape(int(np.prod(Y.shape[:pn])), -1, order="F")
    Am = dense[m]
    k = list(topo.children[m]).index(n)
    perm = [j for j in range(Am.ndim) if j != k] + [k]
    Pp = np.transpose(Am, perm).reshape(-1, Am.shape[k], order="F")
    Ma = A.reshape(-1, A.shape[-1], order="F")
    Qd, sv, _ = np.linalg.svd(Ma, full_matrices=False)
    Qd = Qd[:, sv > 1e-12 * sv[0]]
    Z = Ym - (Ym @ Pp.conj()) @ Pp.T
    Z = Z - Qd @ (Qd.conj().T @ Z)
    Uz, s, _ = np.linalg.svd(Z, full_matrices=False)
    assert s[2] - s[3] > 1e-6 * s[0]
    assert abs(ev.last_cbe_info["sigma_max"] - s[0]) < 1e-10 * s[0]
    E = Um - Qd @ (Qd.conj().T @ Um)                     # the spawned part of U
    Es, es, _ = np.linalg.svd(E, full_matrices=False)
    Es = Es[:, es > 0.5]
    assert Es.shape[1] == 3
    ov = np.linalg.svd(Uz[:, :3].conj().T @ Es, compute_uv=False)
    assert np.all(ov > 1 - 1e-9), ov


def test_sector_tdvp1_cbe_matches_dense_tdvp1_cbe(g):
    """TDVP1_CBE on a U(1) x U(1) Hubbard state (binary(2), D=8, maxdim 16)
    against TDVP1_CBE on the dense embedding of the same tensors: the dense
    core is block diagonal in charge, so both spawn the same directions and
    reach one physical state; bonds grow, norm and energy are conserved."""
    from graft_amd import backend as B, contractions as C, evolution as EV, networks as N
    from graft_amd import operators as O, sectors as S, trees as Tr
    from graft_amd.tensors import DTensor
    topo = Tr.binary_topology(2)
    t = _tree(topo)
    _say("sector TDVP1_CBE vs dense TDVP1_CBE, binary(2) hubbard, D=8")
    psi, _ = S.random_hubbard_ttns(np.random.default_rng(22), topo, 8, sigma=1.6)
    Hs = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    Wd = O.hubbard_host(topo, 1.0, 4.0)
    Hd = N.TTNO(topo, [DTensor.from_numpy(w, w.ndim - 2) for w in Wd], ishermitian=True)
    psi_d = N.TTNS(topo, [DTensor.from_numpy(a, a.ndim - 1) for a in _dense_state(psi)], psi.center)
    E0 = C.expect(psi, Hs, C.EnvCache(topo)).real
    mk = lambda: EV.TDVP1_CBE(order=2, trunc=B.TruncationScheme(maxdim=16),
                              cbe=EV.PredictorCBE(max_add=4, neigs=2))
    evs, evd = mk(), mk()
    b0 = sum(psi.bond_dims())
    for _ in range(2):
        EV.step_(evs, psi, Hs, -0.05j)
        EV.step_(evd, psi_d, Hd, -0.05j)
    assert sum(psi.bond_dims()) > b0
    assert psi.bond_dims() == psi_d.bond_dims()
    vs, vd = _full_vector(t, psi), _full_vector(t, psi_d)
    assert 1 - abs(np.vdot(vs, vd)) < 1e-9
    assert abs(np.linalg.norm(vs) - 1) < 1e-10
    E1 = C.expect(psi, Hs, C.EnvCache(topo)).real
    assert abs(E1 - E0) < 1e-9 * abs(E0), (E0, E1)


@pytest.mark.parametrize("mode", [1, 2])
def test_fused_map_loads_equal_gathers(g, mode):
    """Fused FP64 sector steps that read and write the blocks through the
    element maps inside the GEMM (one launch per step) give bit for bit the
    result of gather, GEMM and scatter, for the automatic and the always-fused
    mode, on the root-edge map of the 16-site tree at D=48; both equal the
    oracle on star(3,1)."""
    from graft_amd import _native as nat, contractions as C, operators as O, sectors as S, trees as Tr
    ctx = nat.default_context()
    edges = [(str((i - 1) // 2), str(i)) for i in range(1, 16)]
    topo = Tr.TreeTopology("0", edges)
    _say("fused map loads vs gathers, mode", mode, "16-site D=48")
    psi, _ = S.random_hubbard_ttns(np.random.default_rng(8), topo, 48, sigma=1.6)
    Hs = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    theta = C.two_site_tensor(psi, 1, 0)
    try:
        ctx.set_sector_fusion(mode)
        h2 = C.eff_h2(C.EnvCache(topo), psi, Hs, 1, 0, out_structure=theta.blocks)
        assert any(s["blocks"] > 1 for s in h2.native.step_info())
        ctx.set_fused_map_loads(0)
        y0 = h2(theta).data.clone()
        ctx.set_fused_map_loads(1)
        y1 = h2(theta).data.clone()
        assert bool((y0 == y1).all()), float((y0 - y1).abs().max())
        topo3 = Tr.star_topology(3, 1)
        t = _tree(topo3)
        psi3, _ = S.random_hubbard_ttns(np.random.default_rng(3), topo3, 12)
        H3 = O.hubbard_sector_ttno(topo3, 1.0, 4.0)
        Wd = O.hubbard_host(topo3, 1.0, 4.0)
        dense = _dense_state(psi3)
        envs = ref.all_envs(t, dense, Wd)
        th3 = C.two_site_tensor(psi3, 1, 0)
        h3 = C.eff_h2(C.EnvCache(topo3), psi3, H3, 1, 0, out_structure=th3.blocks)
        want = ref.h2_apply(t, Wd, envs, 1, 0, ref.two_site_tensor(t, dense, 1, 0))
        assert _rel(h3(th3).to_numpy(), want) < TOL
    finally:
        ctx.set_fused_map_loads(1)
        ctx.set_sector_fusion(1)


def test_lean_stream_equals_general_kernel(g):
    """The lean per-block streaming kernel (small operand staged in LDS, row
    digits by multiply-high) gives the general per-block kernel's result on
    every streaming step of the root-edge map of the 16-site tree at D=48,
    in the automatic and the always-streaming fusion mode; the always-
    streaming map also equals the oracle on star(3,1)."""
    from graft_amd import _native as nat, contractions as C, operators as O, sectors as S, trees as Tr
    ctx = nat.default_context()
    edges = [(str((i - 1) // 2), str(i)) for i in range(1, 16)]
    topo = Tr.TreeTopology("0", edges)
    psi, _ = S.random_hubbard_ttns(np.random.default_rng(8), topo, 48, sigma=1.6)
    Hs = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    theta = C.two_site_tensor(psi, 1, 0)
    try:
        for mode in (1, 3):
            _say("lean streaming vs general kernel, fusion mode", mode, "16-site D=48")
            ctx.set_sector_fusion(mode)
            h2 = C.eff_h2(C.EnvCache(topo), psi, Hs, 1, 0, out_structure=theta.blocks)
            assert any(s["config"] == nat.block_stream_config() for s in h2.native.step_info())
            ctx.set_lean_stream(0)
            y0 = h2(theta).data.clone()
            ctx.set_lean_stream(1)
            y1 = h2(theta).data.clone()
            scale = float(y0.abs().max())
            assert float((y0 - y1).abs().max()) <= 1e-14 * scale
        ctx.set_sector_fusion(3)
        topo3 = Tr.star_topology(3, 1)
        t = _tree(topo3)
        psi3, _ = S.random_hubbard_ttns(np.random.default_rng(3), topo3, 12)
        H3 = O.hubbard_sector_ttno(topo3, 1.0, 4.0)
        Wd = O.hubbard_host(topo3, 1.0, 4.0)
        dense = _dense_state(psi3)
        envs = ref.all_envs(t, dense, Wd)
        th3 = C.two_site_tensor(psi3, 1, 0)
        h3 = C.eff_h2(C.EnvCache(topo3), psi3, H3, 1, 0, out_structure=th3.blocks)
        want = ref.h2_apply(t, Wd, envs, 1, 0, ref.two_site_tensor(t, dense, 1, 0))
        assert _rel(h3(th3).to_numpy(), want) < TOL
    finally:
        ctx.set_lean_stream(1)
        ctx.set_sector_fusion(1)
