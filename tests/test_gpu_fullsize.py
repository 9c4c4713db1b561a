"""Parity at the measured sizes, through properties that hold at any size.

The oracle finishes in seconds only on small trees.  At the workloads the
benchmark measures -- the dense root-edge eff_h2 of the depth-5 binary tree at
D=256 (Theta = 6.7e7 complex entries, 4.97 TFLOP per apply) and the
U(1) x U(1) root-edge eff_h2 of the 16-site Hubbard tree at D=1024 (45 sectors
per bond) -- the checks are:
  * the Rayleigh quotient of the map on the two-site tensor of a state
    canonical towards the bond equals the full closure <psi|H|psi> built from
    the environment cache (the tangent-space identity of
    test/runtests.jl:669-671, expectation.jl:32-41);
  * hermiticity <a, H b> = <H a, b> and linearity H(a + 2b) = H a + 2 H b for
    random a, b, with bounds scaled by the norms of the vectors involved.
"""
import numpy as np
import pytest

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


def _random_like(x, seed):
    import torch
    from graft_amd.tensors import DTensor
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    data = torch.randn(x.numel, dtype=torch.complex128, device="cuda", generator=gen)
    return DTensor(x.dims, data, x.ncod, x.blocks)


def _check_map(h, x, E_closure, tol_rq):
    """Rayleigh quotient against the closure, hermiticity, linearity."""
    from graft_amd import tensors as T
    from graft_amd.tensors import DTensor
    nx = T.inner(x, x).real
    E_map = (T.inner(x, h(x)) / nx).real
    assert abs(E_map - E_closure) < tol_rq * abs(E_closure), (E_map, E_closure)
    a, b = _random_like(x, 1), _random_like(x, 2)
    ha, hb = h(a), h(b)
    lhs, rhs = T.inner(a, hb), T.inner(ha, b)
    assert abs(lhs - rhs) < 1e-12 * T.norm(a) * T.norm(hb), (lhs, rhs)
    c = DTensor(x.dims, a.data + 2.0 * b.data, x.ncod, x.blocks)
    hc = h(c).data
    want = ha.data + 2.0 * hb.data
    err = float((hc - want).abs().max().item())
    assert err < 1e-12 * float(want.abs().max().item()), err


def test_dense_root_edge_D256(g):
    from graft_amd import contractions as C, networks as N, operators as O, trees as Tr
    topo = Tr.binary_topology(5)
    _say("full size dense eff_h2, depth-5 binary tree, Heisenberg, D=256")
    rng = np.random.default_rng(7)
    psi = N.random_ttns(rng, topo, 2, 256)          # canonical, centre at the root
    H = O.heisenberg_ttno(topo, 1.0)
    cache = C.EnvCache(topo)
    h2 = C.eff_h2(cache, psi, H, 1, 0)
    x = C.two_site_tensor(psi, 1, 0)
    assert x.numel == 256 * 256 * 2 * 256 * 2
    E_closure = C.expect(psi, H, C.EnvCache(topo)).real
    _check_map(h2, x, E_closure, 1e-10)


def test_sector_root_edge_D1024(g):
    from graft_amd import contractions as C, operators as O, sectors as S, trees as Tr
    topo = Tr.TreeTopology("0", [(str((i - 1) // 2), str(i)) for i in range(1, 16)])
    _say("full size sector eff_h2, 16-site Hubbard tree, U(1)xU(1), D=1024")
    rng = np.random.default_rng(5)
    psi, _ = S.random_hubbard_ttns(rng, topo, 1024, sigma=1.6)
    assert len(psi.tensors[1].blocks.legs[-1].sectors) >= 30
    H = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    x = C.two_site_tensor(psi, 1, 0)
    assert x.blocks.total == sum(x.blocks.sizes)          # no gaps between blocks
    h2 = C.eff_h2(C.EnvCache(topo), psi, H, 1, 0, out_structure=x.blocks)
    E_closure = C.expect(psi, H, C.EnvCache(topo)).real
    _check_map(h2, x, E_closure, 1e-10)
