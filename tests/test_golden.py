"""Golden fixtures of the effective-Hamiltonian path (tests/golden/*.npz,
written by tests/golden/make_golden.py from the oracle).

CPU: the oracle still reproduces every stored array (the fixtures pin the
oracle's current outputs; the oracle itself is pinned in test_oracle.py).
GPU: the device environments and maps, run on the stored inputs, equal the
stored outputs to 1e-12 relative without running the oracle."""
import os

import numpy as np
import pytest

from oracle import heff_oracle as ref

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TREES = {"star31_tfi_D3": lambda: ref.star_tree(3, 1),
         "binary2_heis_D4": lambda: ref.binary_tree(2),
         "chain5_tfi_D4": lambda: ref.chain_tree(5)}


def _load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", sorted(TREES))
def test_oracle_reproduces_golden(name):
    t = TREES[name]()
    g = _load(name)
    psi = [g[f"psi_{i}"] for i in range(t.n)]
    W = [g[f"W_{i}"] for i in range(t.n)]
    envs = ref.all_envs(t, psi, W)
    for (u, v), E in envs.items():
        assert _rel(E, g[f"env_{u}_{v}"]) < 1e-13
    for n in range(t.n):
        assert _rel(ref.h1_apply(t, W, envs, n, psi[n]), g[f"h1_{n}"]) < 1e-13
    for n in range(1, t.n):
        m = t.parent[n]
        y = ref.h2_apply(t, W, envs, n, m, ref.two_site_tensor(t, psi, n, m))
        assert _rel(y, g[f"h2_{n}_{m}"]) < 1e-13
    assert abs(ref.env_closure(t, psi, W, envs, 0) - g["energy"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(TREES))
def test_device_maps_match_golden(name):
    import torch
    from graft_amd import _native, contractions as C, networks as N, trees as Tr
    from graft_amd.tensors import DTensor
    torch.cuda.set_device(0)
    _native.lib()
    print("[case] golden", name, flush=True)
    t = TREES[name]()
    g = _load(name)
    topo = Tr.TreeTopology(t.ids[0], [(t.ids[t.parent[i]], t.ids[i]) for i in range(1, t.n)])
    psi = N.ttns_from_host(topo, [g[f"psi_{i}"] for i in range(t.n)], 0)
    H = N.TTNO(topo, [DTensor.from_numpy(g[f"W_{i}"], g[f"W_{i}"].ndim - 2) for i in range(t.n)],
               ishermitian=True)
    cache = C.EnvCache(topo)
    for key in g:
        if key.startswith("env_"):
            u, v = map(int, key.split("_")[1:])
            assert _rel(C.env_(cache, psi, H, u, v).to_numpy(), g[key]) < 1e-12, key
    for n in range(t.n):
        h1 = C.eff_h1(cache, psi, H, n)
        assert _rel(h1(psi.tensors[n]).to_numpy(), g[f"h1_{n}"]) < 1e-12, n
    for n in range(1, t.n):
        m = t.parent[n]
        h2 = C.eff_h2(cache, psi, H, n, m)
        y = h2(C.two_site_tensor(psi, n, m)).to_numpy()
        assert _rel(y, g[f"h2_{n}_{m}"]) < 1e-12, (n, m)
    m = t.parent[1]
    h0 = C.eff_h0(cache, psi, H, 1, m)
    y = h0(DTensor.from_numpy(g["h0_C"], 1)).to_numpy()
    assert _rel(y, g[f"h0_1_{m}"]) < 1e-12
    assert abs(C.expect(psi, H, cache) - g["energy"]) < 1e-12
