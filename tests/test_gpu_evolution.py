"""GPU checks of TDVP1 with controlled bond expansion (TDVP1_CBE).

Reference behaviour (test/runtests.jl:908-959): disabled enrichment reproduces
TDVP1 exactly; from bond dimension 1 on the 3x2 star TFI, TDVP1 cannot leave the
product manifold while TDVP1_CBE grows the bonds and tracks exact propagation.
The spawned directions are pinned against the oracle's dense projected core.
"""
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


def _setup(bond, seed):
    from graft_amd import networks as N, operators as O, trees as Tr
    topo = Tr.star_topology(3, 2)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    rng = np.random.default_rng(seed)
    psi = N.random_ttns(rng, topo, 2, bond)
    H = O.tfi_ttno(topo, 1.0, 0.9)
    Hd = ref.dense_hamiltonian(t.n, ref.tfi_terms(t, 1.0, 0.9))
    return topo, t, psi, H, Hd


def _dense(t, psi):
    return ref.ttns_to_dense(t, [a.to_numpy() for a in psi.tensors], [True] * t.n)


def test_disabled_cbe_is_tdvp1(g):
    from graft_amd import evolution as EV
    topo, t, psi, H, _ = _setup(1, 5)
    _say("TDVP1_CBE(enabled=False) vs TDVP1")
    pa, pb = psi.copy(), psi.copy()
    eva, evb = EV.TDVP1(order=2), EV.TDVP1_CBE(order=2, enabled=False)
    for _ in range(3):
        EV.step_(eva, pa, H, -0.02j)
        EV.step_(evb, pb, H, -0.02j)
    assert np.linalg.norm(_dense(t, pa) - _dense(t, pb)) < 1e-14


def test_cbe_grows_bonds_and_tracks_exact_evolution(g):
    from graft_amd import backend as B, evolution as EV
    topo, t, psi, H, Hd = _setup(1, 6)
    _say("TDVP1_CBE bond growth from D=1")
    v0 = _dense(t, psi)
    dt, nsteps = 0.02, 10
    vex = ref.exact_evolve(Hd, v0, -1j * dt * nsteps)
    p1, p3 = psi.copy(), psi.copy()
    ev1 = EV.TDVP1(order=2)
    evc = EV.TDVP1_CBE(trunc=B.TruncationScheme(maxdim=16, atol=1e-12),
                       cbe=EV.PredictorCBE(max_add=8, spawn_threshold=1e-10))
    for _ in range(nsteps):
        EV.step_(ev1, p1, H, -1j * dt)
        EV.step_(evc, p3, H, -1j * dt)
    f1 = abs(np.vdot(_dense(t, p1), vex))
    f3 = abs(np.vdot(_dense(t, p3), vex))
    assert max(p1.bond_dims()) == 1
    assert max(p3.bond_dims()) > 1
    assert abs(1 - f3) < 1e-6, (1 - f3, 1 - f1)
    assert abs(1 - f3) < abs(1 - f1)
    assert abs(np.linalg.norm(_dense(t, p3)) - 1) < 1e-10


@pytest.mark.parametrize("arm,D,active,solver", [
    (2, 2, "n", None), (5, 16, "n", None), (5, 16, "m", None), (5, 16, "n", "gram_u"),
    (5, 16, "m", "gram_v"), (5, 16, "m", "block_krylov"), (5, 16, "n", "exact_svd")])
def test_cbe_directions_match_dense_projected_core(g, arm, D, active, solver):
    """The spawned directions span the leading left singular space of the
    double-complement core computed densely by the oracle (two-site map
    applied to Theta, both projections), for either endpoint of the
    centre-arm edge active, through every solver (small-side Gram on either
    side, block Krylov, exact SVD); they are orthogonal to the active range Q."""
    import torch
    from graft_amd import backend as B, contractions as C, evolution as EV, networks as N
    from graft_amd import operators as O, trees as Tr
    topo = Tr.star_topology(3, arm)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    psi = N.random_ttns(np.random.default_rng(7), topo, 2, D)
    H = O.tfi_ttno(topo, 1.0, 0.9)
    _say("CBE directions vs dense core, star(3,%d) D=%d active %s solver %s" % (arm, D, active, solver))
    n = topo.children[topo.root][0]          # first arm node, child of the centre
    m = topo.root
    u = n if active == "n" else m
    cache = C.EnvCache(topo)
    N.move_center_(psi, u, cache)
    ev = EV.TDVP1_CBE(trunc=B.TruncationScheme(maxdim=10 * D),
                      cbe=EV.PredictorCBE(max_add=8, neigs=2, spawn_threshold=1e-12))
    ev.cache = cache
    host = [a.to_numpy() for a in psi.tensors]
    k = [c for c in topo.children[m]].index(n)
    An, Am = host[n], host[m]
    permm = [j for j in range(Am.ndim) if j != k] + [k]
    Mn = An.reshape(-1, An.shape[-1], order="F")
    Mm = np.transpose(Am, permm).reshape(-1, Am.shape[k], order="F")
    Qa, _ = np.linalg.qr(Mn if active == "n" else Mm)
    E = EV._cbe_directions(ev, psi, H, n, m, active == "n", torch.from_numpy(Qa).cuda(), -0.02j, 8,
                           solver=solver)
    # oracle: the two-site map on Theta, then both complements
    W = O.tfi_host(topo, 1.0, 0.9)
    envs = ref.all_envs(t, host, W)
    theta = ref.two_site_tensor(t, host, n, m)
    Y = ref.h2_apply(t, W, envs, n, m, theta)
    pn = An.ndim - 1
    Ym = Y.reshape(int(np.prod(Y.shape[:pn])), -1, order="F")
    Z, Pp = (Ym, Mm) if active == "n" else (Ym.T, Mn)
    Z = Z - (Z @ Pp.conj()) @ Pp.T
    Z = Z - Qa @ (Qa.conj().T @ Z)
    U, s, _ = np.linalg.svd(Z, full_matrices=False)
    r = E.shape[1]
    assert r == 2
    want = solver or ("gram_v" if Z.shape[1] <= Z.shape[0] else "gram_u")
    assert ev.last_cbe_info["solver"].startswith(want), (ev.last_cbe_info["solver"], want)
    assert abs(ev.last_cbe_info["sigma_max"] - s[0]) < 1e-10 * s[0]
    # same 2-dimensional subspace (the gap to the third value keeps it well defined)
    Eh = E.cpu().numpy()
    ov = np.linalg.svd(U[:, :2].conj().T @ Eh, compute_uv=False)
    assert s[1] - s[2] > 1e-6 * s[0]
    assert np.all(ov > 1 - 1e-9)
    assert np.abs(Qa.conj().T @ Eh).max() < 1e-12


def test_tdvp1_matches_oracle_at_truncated_bonds(g):
    """7-site binary TFI tree at D=3 (truncated bonds, where TDVP1 is not the
    exact propagator): the device TDVP1 of order 2 (Krylov exponentials to
    1e-12, tdvp.jl:171-301) and the oracle's sweep with exact local
    exponentials give the same state after 3 steps."""
    from graft_amd import evolution as EV, networks as N, operators as O, trees as Tr
    from graft_amd.tensors import DTensor
    topo = Tr.binary_topology(2)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    assert t.parent[0] < 0
    _say("TDVP1 vs oracle TDVP1, binary(2) TFI, D=3")
    host = ref.random_ttns(np.random.default_rng(9), t, 2, 3)
    psi = N.TTNS(topo, [DTensor.from_numpy(a, a.ndim - 1) for a in host], 0)
    H = O.tfi_ttno(topo, 1.0, 0.9)
    Hd = ref.dense_hamiltonian(t.n, ref.tfi_terms(t, 1.0, 0.9))
    ev = EV.TDVP1(order=2)
    p, c = list(host), 0
    for _ in range(3):
        EV.step_(ev, psi, H, -0.05j)
        p, c = ref.tdvp1_step(t, p, c, Hd, -0.05j, 2)
    want = ref.ttns_to_dense(t, p, [True] * t.n)
    v0 = ref.ttns_to_dense(t, host, [True] * t.n)
    assert np.linalg.norm(want - ref.exact_evolve(Hd, v0, -0.15j)) > 1e-3   # truncation matters
    assert np.linalg.norm(_dense(t, psi) - want) < 1e-9
