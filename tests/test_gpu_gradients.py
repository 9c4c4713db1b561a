"""GPU checks of the adjoint contractions (rrules) and the energy gradient.

Pinning: (1) a pullback cotangent sbar of y = F(x, s) satisfies
<sbar, delta> = <ybar, F(s -> delta)> exactly, because <ybar, y> is linear in
every operand; the right side runs the forward map (itself pinned against the
oracle) with the operand replaced.  (2) The energy gradient equals the
oracle's (which is pinned against dense vectors and finite differences on the
CPU).  (3) At the differentiation workload's shape (31-site binary tree, TFI
with random couplings) central differences with step 1e-6 of the GPU energy.
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


def _maps(D=8, seed=3):
    from graft_amd import contractions as C, networks as N, operators as O, trees as Tr
    topo = Tr.binary_topology(3)
    rng = np.random.default_rng(seed)
    psi = N.random_ttns(rng, topo, 2, D)
    H = O.heisenberg_ttno(topo, 1.0)
    cache = C.EnvCache(topo)
    N.move_center_(psi, 1, cache)
    return {"h1": (C.eff_h1(cache, psi, H, 1), psi.tensors[1]),
            "h2": (C.eff_h2(cache, psi, H, 1, 0), C.two_site_tensor(psi, 1, 0))}


@pytest.mark.parametrize("kind", ["h1", "h2"])
def test_rrule_apply_cotangents_satisfy_linearity(g, kind):
    from graft_amd import gradients as G, tensors as T
    from graft_amd.contractions import EffectiveMap
    from graft_amd.tensors import DTensor, TensorProto
    h, x = _maps()[kind]
    _say("rrule apply", kind, h.xdims)
    rng = np.random.default_rng(9)
    y, pb = G.rrule_apply(h, x)
    ybar = DTensor.from_numpy(ref.random_complex(rng, h.xdims), x.ncod)
    xbar, sbars = pb(ybar)
    # x: <xbar, dx> = <ybar, F dx>
    dx = DTensor.from_numpy(ref.random_complex(rng, h.xdims), x.ncod)
    lhs, rhs = T.inner(xbar, dx), T.inner(ybar, h(dx))
    assert abs(lhs - rhs) < 1e-12 * abs(rhs)
    # hermitian map: xbar = F ybar
    assert np.abs(xbar.to_numpy() - h(ybar).to_numpy()).max() < 1e-12 * np.abs(xbar.to_numpy()).max()
    # every static operand
    for k, s in enumerate(h.statics):
        assert sbars[k].dims == s.dims, k
        delta = DTensor.from_numpy(ref.random_complex(rng, s.dims), s.ncod)
        statics = list(h.statics)
        statics[k] = delta
        hd = EffectiveMap(h.native, statics, TensorProto(h.xdims, h.ncod, h.blocks), h.kind)
        lhs, rhs = T.inner(sbars[k], delta), T.inner(ybar, hd(x))
        assert abs(lhs - rhs) < 1e-11 * max(abs(rhs), 1e-300), (k, lhs, rhs)


def _tfi_random(topo, rng):
    from graft_amd import operators as O
    n = len(topo.ids)
    Js = {c: rng.uniform(0.5, 1.5) for c in range(n) if topo.parent[c] >= 0}
    hs = rng.uniform(0.5, 1.5, n)
    return Js, hs, O.tfi_random_host(topo, Js, hs)


def _upload_ttno(topo, Wh):
    from graft_amd import networks as N
    from graft_amd.tensors import DTensor
    return N.TTNO(topo, [DTensor.from_numpy(w, w.ndim - 2) for w in Wh], ishermitian=True)


def test_energy_gradient_matches_oracle(g):
    from graft_amd import gradients as G, networks as N, trees as Tr
    topo = Tr.binary_topology(2)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    _say("energy gradient vs oracle, 7 sites, non-canonical")
    rng = np.random.default_rng(31)
    host = []
    for n in range(t.n):
        shape = [3] * len(t.children[n]) + [2] + [1 if t.parent[n] < 0 else 3]
        host.append(ref.random_complex(rng, tuple(shape)))
    Js, hs, Wh = _tfi_random(topo, rng)
    psi = N.ttns_from_host(topo, host, 0)
    R = G.energy_and_gradient(psi, _upload_ttno(topo, Wh))
    E, Z, grads, Rs = ref.expectation_gradient(t, host, Wh)
    assert abs(R.energy - E) < 1e-12 * abs(E)
    assert abs(R.norm2 - Z.real) < 1e-12 * Z.real
    for n in range(t.n):
        assert np.abs(R.site[n].to_numpy() - grads[n]).max() < 1e-12 * np.abs(grads[n]).max(), n
        assert np.abs(R.ttno[n].to_numpy() - Rs[n]).max() < 1e-12 * np.abs(Rs[n]).max(), n


@pytest.mark.parametrize("D", [16, 128])
def test_energy_gradient_finite_differences_31_sites(g, D):
    """31-site binary tree, TFI with J_e, h_i ~ U[0.5, 1.5]: every coupling
    derivative and one random direction per site tensor against central
    differences (step 1e-6) of the GPU energy."""
    import torch
    from graft_amd import gradients as G, networks as N, operators as O, trees as Tr
    from graft_amd.tensors import DTensor
    topo = Tr.binary_topology(4)
    n = len(topo.ids)
    _say("energy gradient FD, 31 sites, D", D)
    rng = np.random.default_rng(11)
    psi = N.random_ttns(rng, topo, 2, D)
    Js, hs, Wh = _tfi_random(topo, rng)
    H = _upload_ttno(topo, Wh)
    R = G.energy_and_gradient(psi, H)
    assert abs(R.energy - G.energy(psi, H)) < 1e-12 * abs(R.energy)
    h = 1e-6
    # couplings (the builder is affine in them)
    zero = {c: 0.0 for c in Js}
    W0 = O.tfi_random_host(topo, zero, np.zeros(n))
    params = [("J", c) for c in Js] + [("h", i) for i in range(n)]
    if D > 16:
        params = params[:4] + params[-4:]
    for kind, i in params:
        Jp, Jm, hp, hm = dict(Js), dict(Js), hs.copy(), hs.copy()
        if kind == "J":
            Jp[i] += h
            Jm[i] -= h
            node = topo.parent[i]
            dW = O.tfi_random_host(topo, {**zero, i: 1.0}, np.zeros(n))[node] - W0[node]
        else:
            hp[i] += h
            hm[i] -= h
            node = i
            hv = np.zeros(n)
            hv[i] = 1.0
            dW = O.tfi_random_host(topo, zero, hv)[node] - W0[node]
        Ep = G.energy(psi, _upload_ttno(topo, O.tfi_random_host(topo, Jp, hp)))
        Em = G.energy(psi, _upload_ttno(topo, O.tfi_random_host(topo, Jm, hm)))
        fd = (Ep - Em) / (2 * h)
        an = R.coupling_gradient({node: dW})
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (kind, i, fd, an)
    sites = range(n) if D <= 16 else [0, 1, 3, 7, 15]
    for s in sites:
        A = psi.tensors[s]
        delta = ref.random_complex(rng, A.dims)
        dd = torch.from_numpy(np.ascontiguousarray(delta.reshape(-1, order="F"))).cuda()
        p1, p2 = psi.copy(), psi.copy()
        p1.tensors[s] = DTensor(A.dims, A.data + h * dd, A.ncod)
        p2.tensors[s] = DTensor(A.dims, A.data - h * dd, A.ncod)
        fd = (G.energy(p1, H) - G.energy(p2, H)) / (2 * h)
        an = 2 * np.real(np.vdot(R.site[s].to_numpy(), delta))
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an)), (s, fd, an)
