"""GPU parity of the SU(2) block path (coupled block-pair launches).

Every device result is compared on the dense embedding of the same tensors:
with the oracle's environments and effective maps, with the host form of
each coupled step, and, for the Heisenberg model, with exact
diagonalization restricted to total spin zero.  Every case prints its name
first so a failing run names its case.
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


def _leg(irreps, dual=False):
    from graft_amd import su2
    return su2.SU2Leg(irreps, dual)


def test_coupled_contract_matches_host_and_dense(g):
    from graft_amd import su2
    cases = [
        ([_leg([(1, 3), (3, 2)]), _leg([(1, 1)]), _leg([(0, 2), (2, 1)]), _leg([(0, 2), (2, 3)], True)],
         [_leg([(1, 3), (3, 2)], True), _leg([(0, 2), (2, 1)], True), _leg([(1, 1)], True),
          _leg([(0, 1), (2, 2)])],
         (0, 1, 2), (0, 2, 1), (1, 0), False, False),
        ([_leg([(1, 40), (3, 24)]), _leg([(1, 1)]), _leg([(0, 2), (2, 1)]), _leg([(0, 30), (2, 41)], True)],
         [_leg([(0, 2), (2, 1)]), _leg([(1, 1)]), _leg([(0, 30), (2, 41)]), _leg([(1, 1)], True)],
         (3,), (2,), (0, 4, 1, 2, 5, 3), True, True),
        ([_leg([(0, 17), (2, 33), (4, 9)]), _leg([(1, 1)]), _leg([(1, 70), (3, 20)], True)],
         [_leg([(1, 70), (3, 20)]), _leg([(1, 1)]), _leg([(0, 11), (2, 21)], True)],
         (2,), (0,), (3, 0, 2, 1), False, False),
    ]
    rng = np.random.default_rng(5)
    splits = []
    for i, (legsA, legsB, cA, cB, perm, conjA, conjB) in enumerate(cases):
        _say("coupled contract", i, "conj", conjA, conjB)
        SA, SB = su2.SU2Structure.invariant(legsA), su2.SU2Structure.invariant(legsB)
        fa, fb = su2.random_flat(rng, SA), su2.random_flat(rng, SB)
        A, B = su2.to_device(SA, fa), su2.to_device(SB, fb)
        C = su2.contract(A, cA, B, cB, conjA, conjB, perm)
        host = C.data.cpu().numpy()
        oA = tuple(x for x in range(A.rank) if x not in cA)
        oB = tuple(x for x in range(B.rank) if x not in cB)
        p = next(iter(reversed(su2._PAIR_PLANS.values())))
        want = su2.host_execute_pair(C.blocks, p.step_pairs[0], p.step_coefs[0], SA, fa, conjA, SB, fb,
                                     conjB, oA, cA, oB, cB, perm)
        assert _rel(host, want) < TOL
        Ad, Bd = SA.to_dense(fa), SB.to_dense(fb)
        Ad = Ad.conj() if conjA else Ad
        Bd = Bd.conj() if conjB else Bd
        nc = len(cA)
        dense = np.tensordot(Ad.transpose(oA + cA), Bd.transpose(cB + oB),
                             axes=(list(range(len(oA), len(oA) + nc)), list(range(nc)))).transpose(perm)
        assert _rel(C.to_numpy(), dense) < TOL
        assert p.npairs > len(C.blocks.blocks)        # blocks sum several pairs
        splits.append(_step_split(p))
    # a tile step with few output tiles cuts its blocks' pair lists into
    # chunks summed afterwards: exercised here, and checked like the rest
    assert max(splits) > 1, splits


def _step_split(p):
    """Split factor of a one-step plan (chunks per output block's pair list
    for a coupled tile step)."""
    import ctypes
    from graft_amd import _native as nat
    cfg, orient, ks = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    nat.check(nat.lib().graft_plan_step_launch(p.handle, 0, ctypes.byref(cfg), ctypes.byref(orient),
                                               ctypes.byref(ks)))
    return ks.value


@pytest.mark.parametrize("name,tphys,maxdeg", [("star31", 1, None), ("binary2", 2, 2)])
def test_su2_envs_and_maps_match_oracle(g, name, tphys, maxdeg):
    from graft_amd import contractions as C, su2, trees as Tr
    topo = Tr.star_topology(3, 1) if name == "star31" else Tr.binary_topology(2)
    t = _tree(topo)
    _say("su2 maps", name, "spin", tphys, "/2, maxdeg", maxdeg)
    rng = np.random.default_rng(13)
    psi = su2.random_ttns(rng, topo, tphys, maxdeg)
    H = su2.random_ttno(rng, topo, tphys)
    dense = [x.to_numpy() for x in psi.tensors]
    Wd = [w.to_numpy() for w in H.tensors]
    envs = ref.all_envs(t, dense, Wd)
    cache = C.EnvCache(topo)
    for (u, v), E in envs.items():
        got = C.env_(cache, psi, H, u, v)
        assert got.blocks is not None
        assert _rel(got.to_numpy(), E) < TOL, (u, v)
    root = psi.center
    h1 = C.eff_h1(cache, psi, H, root)
    assert _rel(h1(psi.tensors[root]).to_numpy(), ref.h1_apply(t, Wd, envs, root, dense[root])) < TOL
    assert abs(C.expect(psi, H, cache) - ref.env_closure(t, dense, Wd, envs, root)) < 1e-11
    child = topo.children[root][0]
    theta = C.two_site_tensor(psi, child, root)
    theta_ref = ref.two_site_tensor(t, dense, child, root)
    assert _rel(theta.to_numpy(), theta_ref) < TOL
    h2 = C.eff_h2(cache, psi, H, child, root, out_structure=theta.blocks)
    assert _rel(h2(theta).to_numpy(), ref.h2_apply(t, Wd, envs, child, root, theta_ref)) < TOL
    assert all(s["blocks"] >= 1 for s in h2.native.step_info())
    h0 = C.eff_h0(cache, psi, H, child, root)
    link = su2.to_device(h0.blocks, su2.random_flat(rng, h0.blocks), 1)
    assert _rel(h0(link).to_numpy(), ref.h0_apply(envs, child, root, link.to_numpy())) < TOL


def _total_spin_squared(nsites, tphys):
    from graft_amd import su2
    S = su2.spin_matrices(tphys)
    d = tphys + 1
    tot = [np.zeros((d ** nsites,) * 2, dtype=complex) for _ in range(3)]
    for i in range(nsites):
        for a in range(3):
            op = np.ones((1, 1))
            for j in reversed(range(nsites)):      # site 0 fastest in the dense vector
                op = np.kron(op, S[a] if j == i else np.eye(d))
            tot[a] += op
    return sum(x @ x for x in tot)


@pytest.mark.parametrize("name,tphys", [("star31", 1), ("binary2", 2)])
def test_su2_heisenberg_lanczos_gives_the_singlet_ground_energy(g, name, tphys):
    """Full bonds: the two-site map at the root edge is H restricted to total
    spin zero, so its lowest eigenvalue is the lowest singlet energy of exact
    diagonalization (star(3,1), spin 1/2: E = -3/4 J exactly, since the
    root spin couples to leaf total spin 1/2)."""
    from graft_amd import contractions as C, krylov as K, su2, trees as Tr
    topo = Tr.star_topology(3, 1) if name == "star31" else Tr.binary_topology(2)
    t = _tree(topo)
    _say("su2 heisenberg lanczos", name, "spin", tphys, "/2")
    rng = np.random.default_rng(17)
    psi = su2.random_ttns(rng, topo, tphys)
    H = su2.heisenberg_ttno(topo, 1.0, tphys)
    n = len(topo.ids)
    Hd = ref.ttno_to_dense(t, [w.to_numpy() for w in H.tensors])
    S2 = _total_spin_squared(n, tphys)
    e_singlet = np.linalg.eigvalsh(Hd + 50.0 * S2)[0]
    if name == "star31":
        assert abs(e_singlet + 0.75) < 1e-12
    cache = C.EnvCache(topo)
    root = psi.center
    child = topo.children[root][0]
    theta = C.two_site_tensor(psi, child, root)
    h2 = C.eff_h2(cache, psi, H, child, root, out_structure=theta.blocks)
    vals, vecs, info = K.eigsolve(h2, theta, 1, "SR", krylovdim=30, tol=1e-12)
    assert abs(vals[0] - e_singlet) < 1e-10, (vals[0], e_singlet)


def test_su2_split_and_center_moves_keep_the_state(g):
    """The SU(2) two-site split without truncation and the centre moves keep
    the dense state (1e-12); the split's factors are dense isometries on the
    non-centre side, and its spectrum is the dense Theta's singular values."""
    from graft_amd import contractions as C, networks as N, su2, trees as Tr
    topo = Tr.binary_topology(2)
    t = _tree(topo)
    _say("su2 split and centre moves, binary2 spin 1, maxdeg 3")
    rng = np.random.default_rng(23)
    psi = su2.random_ttns(rng, topo, 2, 3)
    v0 = ref.ttns_to_dense(t, [a.to_numpy() for a in psi.tensors], list(psi.hasphys))
    child = topo.children[psi.center][0]
    theta = C.two_site_tensor(psi, child, psi.center)
    th = theta.to_numpy()
    pn = psi.tensors[child].ncod
    sv = np.linalg.svd(th.reshape(int(np.prod(th.shape[:pn])), -1, order="F"), compute_uv=False)
    for center_on in ("n", "m"):
        p = N.TTNS(topo, list(psi.tensors), psi.center, psi.hasphys)
        p, info = C.split_two_site_(p, theta, child, topo.parent[child], center_on=center_on)
        v = ref.ttns_to_dense(t, [a.to_numpy() for a in p.tensors], list(p.hasphys))
        assert _rel(v, v0) < 1e-12, center_on
        nz = np.sort(sv[sv > 1e-12 * sv[0]])[::-1]
        got = np.sort(info["spectrum"][info["spectrum"] > 1e-12 * sv[0]])[::-1]
        assert len(got) == len(nz) and np.abs(got - nz).max() < 1e-12 * sv[0]
        iso = child if center_on == "m" else topo.parent[child]
        A = p.tensors[iso].to_numpy()
        if iso == child:
            Mx = A.reshape(-1, A.shape[-1], order="F")
        else:
            k = topo.children[topo.parent[child]].index(child)
            Mx = np.moveaxis(A, k, -1).reshape(-1, A.shape[k], order="F")
        assert np.abs(Mx.conj().T @ Mx - np.eye(Mx.shape[1])).max() < 1e-12
    p = N.TTNS(topo, list(psi.tensors), psi.center, psi.hasphys)
    for target in (5, 2, 6, 0, 3):
        N.move_center_(p, target)
        v = ref.ttns_to_dense(t, [a.to_numpy() for a in p.tensors], list(p.hasphys))
        assert _rel(v, v0) < 1e-12, target
        assert p.center == target


@pytest.mark.parametrize("name,tphys", [("star31", 1), ("chain6", 1), ("binary2", 2)])
def test_su2_dmrg2_reaches_the_singlet_ground_energy(g, name, tphys):
    """dmrg2_ on SU(2) states from small bonds reaches the lowest singlet
    energy of exact diagonalization (1e-10)."""
    from graft_amd import backend as B, groundstate as GS, su2, trees as Tr
    topo = {"star31": lambda: Tr.star_topology(3, 1), "chain6": lambda: Tr.mps_topology(6),
            "binary2": lambda: Tr.binary_topology(2)}[name]()
    t = _tree(topo)
    _say("su2 dmrg2", name, "spin", tphys, "/2")
    rng = np.random.default_rng(29)
    psi = su2.random_ttns(rng, topo, tphys, 1)
    H = su2.heisenberg_ttno(topo, 1.0, tphys)
    n = len(topo.ids)
    Hd = ref.ttno_to_dense(t, [w.to_numpy() for w in H.tensors])
    e_singlet = np.linalg.eigvalsh(Hd + 50.0 * _total_spin_squared(n, tphys))[0]
    _, Es = GS.dmrg2_(psi, H, trunc=B.TruncationScheme(maxdim=4096), nsweeps=6)
    assert abs(Es[-1] - e_singlet) < 1e-10, (Es, e_singlet)


def test_su2_dmrg2_with_a_target_spin(g):
    """A root leg of total spin S: on star(3,1) with spin 1/2 the ground state
    is the triplet, E = -5/4 J exactly (the root spin against leaf total spin
    3/2), and dmrg2_ at S = 1 reaches it; S = 0 gives -3/4 J."""
    from graft_amd import backend as B, contractions as C, groundstate as GS, su2, trees as Tr
    topo = Tr.star_topology(3, 1)
    for target, e_exact in ((2, -1.25), (0, -0.75)):
        _say("su2 dmrg2 star31 target spin", target, "/2")
        rng = np.random.default_rng(31)
        psi = su2.random_ttns(rng, topo, 1, 1, target=target)
        H = su2.heisenberg_ttno(topo)
        _, Es = GS.dmrg2_(psi, H, trunc=B.TruncationScheme(maxdim=64), nsweeps=4)
        assert abs(Es[-1] - e_exact) < 1e-10, (target, Es)
        assert abs(C.expect(psi, H) - e_exact) < 1e-10


def test_su2_link_split_pieces(g):
    """Moving a leg (a contraction with the identity) permutes the dense tensor;
    the per-irrep QR of the last leg gives a dense isometry and rebuilds it."""
    from graft_amd import su2
    _say("su2 leg move and per-irrep QR")
    rng = np.random.default_rng(37)
    legs = [_leg([(1, 2), (3, 1)]), _leg([(1, 1)]), _leg([(0, 2), (2, 3)], True)]
    st = su2.SU2Structure.invariant(legs)
    A = su2.to_device(st, su2.random_flat(rng, st), 2)
    D = A.to_numpy()
    for src, dst in ((0, 2), (2, 0), (1, 2)):
        M = su2._move_leg(A, src, dst)
        assert _rel(M.to_numpy(), np.moveaxis(D, src, dst)) < 1e-13, (src, dst)
    Q, R, new = su2._qr_last(A)
    Qd = Q.to_numpy()
    Qm = Qd.reshape(-1, Qd.shape[-1], order="F")
    assert np.abs(Qm.conj().T @ Qm - np.eye(Qm.shape[1])).max() < 1e-12
    # A = Q . (R (x) 1) on the last leg
    Rd = np.zeros((new.dim, legs[2].dim), dtype=complex)
    ro, co = new.offsets(), legs[2].dual().offsets()
    for s, (tj, dn) in enumerate(new.irreps):
        so = [i for i, (t, _) in enumerate(legs[2].irreps) if t == tj][0]
        Rd[ro[s]:ro[s] + dn * (tj + 1), co[so]:co[so] + legs[2].irreps[so][1] * (tj + 1)] = \
            np.kron(R[tj], np.eye(tj + 1))
    assert _rel(np.tensordot(Qd, Rd, axes=([2], [0])), D) < 1e-12


@pytest.mark.parametrize("order", [1, 2])
def test_su2_tdvp1_is_exact_at_full_bonds(g, order):
    """TDVP1 on SU(2) states (link splits by per-irrep QR): on star(3,1) at full
    bonds every leaf step and the link step after it cancel and the root step
    is the whole propagator, so the state equals exp(dz H) of the dense
    Heisenberg matrix on the embedded vector (1e-9), norm and energy kept."""
    from graft_amd import contractions as C, evolution as EV, su2, trees as Tr
    topo = Tr.star_topology(3, 1)
    t = _tree(topo)
    _say("su2 TDVP1 order", order, "star31 full bonds")
    rng = np.random.default_rng(41)
    psi = su2.random_ttns(rng, topo, 1)
    H = su2.heisenberg_ttno(topo)
    H.ishermitian = True
    Hd = ref.ttno_to_dense(t, [w.to_numpy() for w in H.tensors])
    v0 = ref.ttns_to_dense(t, [a.to_numpy() for a in psi.tensors], list(psi.hasphys))
    vex = ref.exact_evolve(Hd, v0, -0.15j)
    assert np.linalg.norm(vex - v0) > 1e-2
    e0 = C.expect(psi, H)
    ev = EV.TDVP1(order=order)
    for _ in range(3):
        EV.step_(ev, psi, H, -0.05j)
    v = ref.ttns_to_dense(t, [a.to_numpy() for a in psi.tensors], list(psi.hasphys))
    assert np.linalg.norm(v - vex) < 1e-9, np.linalg.norm(v - vex)
    assert abs(np.linalg.norm(v) - np.linalg.norm(v0)) < 1e-10
    assert abs(C.expect(psi, H) - e0) < 1e-9


def test_su2_tdvp1_matches_dense_tdvp1_at_truncated_bonds(g):
    """Truncated bonds (depth-2 binary tree, spin 1, degeneracies <= 2): the
    SU(2) TDVP1 and the plain dense TDVP1 on the embedded tensors follow the
    same trajectory (H is SU(2) invariant, so the dense tangent-space flow
    stays in the symmetric states), to 1e-9; norm and energy are conserved."""
    from graft_amd import contractions as C, evolution as EV, networks as N, su2, tensors as T
    from graft_amd import trees as Tr
    topo = Tr.binary_topology(2)
    t = _tree(topo)
    _say("su2 vs dense TDVP1 binary2 spin 1 maxdeg 2")
    rng = np.random.default_rng(43)
    psi = su2.random_ttns(rng, topo, 2, 2)
    H = su2.heisenberg_ttno(topo, 1.0, 2)
    H.ishermitian = True
    psid = N.TTNS(topo, [T.DTensor.from_numpy(a.to_numpy(), a.ncod) for a in psi.tensors], psi.center,
                  psi.hasphys)
    Hd = N.TTNO(topo, [T.DTensor.from_numpy(w.to_numpy(), w.ncod) for w in H.tensors], ishermitian=True)
    e0 = C.expect(psi, H)
    for p, h in ((psi, H), (psid, Hd)):
        ev = EV.TDVP1(order=2)
        for _ in range(2):
            EV.step_(ev, p, h, -0.05j)
    v = ref.ttns_to_dense(t, [a.to_numpy() for a in psi.tensors], list(psi.hasphys))
    vd = ref.ttns_to_dense(t, [a.to_numpy() for a in psid.tensors], list(psid.hasphys))
    assert np.linalg.norm(v - vd) < 1e-9, np.linalg.norm(v - vd)
    assert abs(np.linalg.norm(v) - 1.0) < 1e-10
    assert abs(C.expect(psi, H) - e0) < 1e-9


def _dense_copies(psi, H):
    from graft_amd import networks as N, tensors as T
    psid = N.TTNS(psi.topo, [T.DTensor.from_numpy(a.to_numpy(), a.ncod) for a in psi.tensors],
                  psi.center, psi.hasphys)
    Hd = N.TTNO(H.topo, [T.DTensor.from_numpy(w.to_numpy(), w.ncod) for w in H.tensors],
                ishermitian=True)
    return psid, Hd


def _dense_link(R, new, old):
    """The dense (new x old) matrix of the per-irrep blocks R[j] (x) identity."""
    Rd = np.zeros((new.dim, old.dim), dtype=complex)
    ro, co = new.offsets(), old.offsets()
    for s, (tj, dn) in enumerate(new.irreps):
        so = [i for i, (t, _) in enumerate(old.irreps) if t == tj]
        if tj in R and so:
            Rd[ro[s]:ro[s] + dn * (tj + 1), co[so[0]]:co[so[0]] + old.irreps[so[0]][1] * (tj + 1)] = \
                np.kron(R[tj], np.eye(tj + 1))
    return Rd


@pytest.mark.parametrize("u,v,maxspin", [(1, 0, None), (0, 2, None), (1, 0, 2), (0, 1, 2)])
def test_su2_cbe_split_matches_the_dense_selector(g, u, v, maxspin):
    """The SU(2) bond-expansion split of edge u -> v (depth-2 binary tree, spin
    1): U is a dense isometry with U (R (x) 1) = A (the state is unchanged),
    and the spawned directions span the same space as the plain dense
    selector's on the embedded tensors with the same number of directions;
    the leading singular value of the projected core agrees.  With bond spins
    capped at 1 the split spawns spins the bond did not carry.  (An edge into
    a leaf has a zero core here: the leaf's bond spans its whole space.)"""
    from graft_amd import backend as B, contractions as C, evolution as EV, networks as N, su2
    from graft_amd import trees as Tr
    topo = Tr.binary_topology(2)
    _say("su2 CBE split", u, "->", v, "maxspin", maxspin)
    rng = np.random.default_rng(53 + u)
    H = su2.heisenberg_ttno(topo, 1.0, 2)
    H.ishermitian = True
    psi = su2.random_ttns(rng, topo, 2, 1, maxspin)
    N.move_center_(psi, u)
    psid, Hd = _dense_copies(psi, H)
    t = topo
    k = psi.tensors[u].rank - 1 if t.parent[u] == v else Tr.childslot(t, u, v)
    old = psi.tensors[u].blocks.legs[k]
    ev = EV.TDVP1_CBE(cbe=EV.PredictorCBE(max_add=24, neigs=24, krylovdim=24),
                      trunc=B.TruncationScheme(maxdim=10 ** 4))
    ev.cache = C.EnvCache(topo)
    U, R, new = su2.cbe_link_split(ev, psi, H, u, v, -0.05j)
    info = dict(ev.last_cbe_info)
    _say("selected", info["selected"], "sigma_max", info["sigma_max"], "old", old.irreps, "new", new.irreps)
    Ud = U.to_numpy()
    Um = Ud.reshape(-1, Ud.shape[-1], order="F")
    assert np.abs(Um.conj().T @ Um - np.eye(Um.shape[1])).max() < 1e-12
    Ad = np.moveaxis(psi.tensors[u].to_numpy(), k, -1)
    assert _rel(np.tensordot(Ud, _dense_link(R, new, old), axes=([Ud.ndim - 1], [0])), Ad) < 1e-12
    s = info["selected"]
    assert s > 0 and new.dim == old.dim + s
    if maxspin is not None:
        assert {tj for tj, _ in new.irreps} - {tj for tj, _ in old.irreps}
    evd = EV.TDVP1_CBE(cbe=EV.PredictorCBE(max_add=s, neigs=s, krylovdim=max(s, 4)),
                       trunc=B.TruncationScheme(maxdim=10 ** 4))
    evd.cache = C.EnvCache(topo)
    Ut, _ = EV._cbe_split(evd, psid, Hd, u, v, -0.05j)
    assert evd.last_cbe_info["selected"] == s
    assert abs(evd.last_cbe_info["sigma_max"] - info["sigma_max"]) <= 1e-10 * info["sigma_max"]
    Dt = np.moveaxis(Ut.to_numpy(), k, -1)
    Dm = Dt.reshape(-1, Dt.shape[-1], order="F")
    assert Dm.shape == Um.shape
    assert np.abs(Um @ Um.conj().T - Dm @ Dm.conj().T).max() < 1e-9


def test_su2_tdvp1_cbe_matches_dense_tdvp1_cbe(g):
    """TDVP1_CBE on SU(2) states from small bonds (depth-2 binary tree, spin 1,
    degeneracies 1, spins up to 1 on the bonds) and the plain dense TDVP1_CBE
    on the embedded tensors, both spawning every direction above the
    threshold: the bonds grow alike (dense dimensions) and the states agree to
    1e-9; norm and energy are conserved."""
    from graft_amd import backend as B, contractions as C, evolution as EV, su2, trees as Tr
    topo = Tr.binary_topology(2)
    t = _tree(topo)
    _say("su2 vs dense TDVP1_CBE binary2 spin 1 maxdeg 1 maxspin 2")
    rng = np.random.default_rng(59)
    psi = su2.random_ttns(rng, topo, 2, 1, 2)
    H = su2.heisenberg_ttno(topo, 1.0, 2)
    H.ishermitian = True
    psid, Hd = _dense_copies(psi, H)
    e0 = C.expect(psi, H)
    dims0 = [a.to_numpy().shape for a in psi.tensors]
    for p, h in ((psi, H), (psid, Hd)):
        ev = EV.TDVP1_CBE(cbe=EV.PredictorCBE(max_add=500, neigs=500, krylovdim=500,
                                              spawn_threshold=1e-3),
                          trunc=B.TruncationScheme(maxdim=500))
        for _ in range(2):
            EV.step_(ev, p, h, -0.05j)
    dims = [a.to_numpy().shape for a in psi.tensors]
    assert dims == [a.to_numpy().shape for a in psid.tensors]
    assert dims != dims0
    v = ref.ttns_to_dense(t, [a.to_numpy() for a in psi.tensors], list(psi.hasphys))
    vd = ref.ttns_to_dense(t, [a.to_numpy() for a in psid.tensors], list(psid.hasphys))
    assert np.linalg.norm(v - vd) < 1e-9, np.linalg.norm(v - vd)
    assert abs(np.linalg.norm(v) - 1.0) < 1e-10
    assert abs(C.expect(psi, H) - e0) < 1e-9


def test_su2_dmrg2_physless_binary_root_bootstrap(g):
    """test/runtests.jl:781-804 on SU(2) blocks: depth-2 binary tree whose root
    carries no site (spin-1/2 sites on the other six nodes, Heisenberg on the
    tree's edges and between the root's two children through the root), from
    bonds of spin 1/2 only (degeneracy 1: dense 2 on each root bond).  The
    exact ground state (a singlet: the graph is bipartite with equal halves)
    has Schmidt rank 6 across each root edge, a spin-1/2 and a spin-3/2
    multiplet (8 is not reached: it is even under exchanging the two leaves
    of each child).  A root-edge split can only keep spins the other root
    bond carries, so spin 3/2 needs the per-spin completion: the bootstrap
    alone keeps the state (1e-12), opens both root bonds to 8 and leaves
    isometries; dmrg2_ then reaches the exact ground energy (1e-10) with root
    bonds at the exact Schmidt rank."""
    from graft_amd import backend as B, contractions as C, groundstate as GS, su2, trees as Tr
    topo = Tr.binary_topology(2)
    t = _tree(topo)
    _say("su2 dmrg2 physless binary(2) root bootstrap")
    flags = [n != topo.root for n in range(7)]
    H = su2.heisenberg_ttno(topo, 1.0, 1, root_phys=False)
    H.ishermitian = True
    w, V = np.linalg.eigh(ref.ttno_to_dense(t, [h.to_numpy() for h in H.tensors], flags))
    kids = topo.children[topo.root]
    sites = [n for n in range(7) if flags[n]]
    ranks = []
    for k in kids:                           # Schmidt rank of the ground state across edge k
        sub = [k] + list(topo.children[k])
        v = np.moveaxis(V[:, 0].reshape([2] * 6, order="F"), [sites.index(x) for x in sub], [0, 1, 2])
        sv = np.linalg.svd(v.reshape(8, 8, order="F"), compute_uv=False)
        ranks.append(int(np.sum(sv > 1e-10 * sv[0])))
    assert ranks == [6, 6]

    def fresh():
        return su2.random_ttns(np.random.default_rng(61), topo, 1, 1, 1, root_phys=False)

    boot = fresh()
    assert [boot.tensors[k].blocks.legs[-1].irreps for k in kids] == [((1, 1),)] * 2
    assert GS._physless_root_two_site_targets(boot, B.TruncationScheme()) == [(k, 8) for k in kids]
    v0 = ref.ttns_to_dense(t, [a.to_numpy() for a in boot.tensors], flags)
    GS._bootstrap_physless_root_(boot, C.EnvCache(topo),
                                 GS._physless_root_two_site_targets(boot, B.TruncationScheme(maxdim=8)))
    assert [boot.tensors[k].dims[-1] for k in kids] == [8, 8]
    boot.check_arrows()
    for k in kids:
        D = boot.tensors[k].to_numpy()
        M = D.reshape(-1, D.shape[-1], order="F")
        assert np.abs(M.conj().T @ M - np.eye(M.shape[1])).max() < 1e-12
    v1 = ref.ttns_to_dense(t, [a.to_numpy() for a in boot.tensors], flags)
    assert np.linalg.norm(v1 - v0) < 1e-12
    psi = fresh()
    _, Es = GS.dmrg2_(psi, H, trunc=B.TruncationScheme(maxdim=8), nsweeps=4)
    assert abs(Es[-1] - w[0]) < 1e-10, (Es, w[0])
    assert [psi.tensors[k].dims[-1] for k in kids] == ranks
