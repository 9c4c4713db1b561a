"""graft_heff_apply_noalloc (the entry point whose workspace is a static
device array) against graft_heff_apply, through the C ABI.

Two calls in a row on a fresh context, the second with a smaller map; the
outputs must equal the allocating entry point's bit for bit (the same
kernels and workspace layout run), the fresh context must still hold no
workspace afterwards, and a plan whose workspace exceeds the static array
is refused with GRAFT_E_NOMEM before any launch.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _map_and_input(D, depth, seed):
    from graft_amd import contractions as C, networks as N, operators as O, trees as Tr
    topo = Tr.binary_topology(depth)
    psi = N.random_ttns(np.random.default_rng(seed), topo, 2, D)
    H = O.heisenberg_ttno(topo, 1.0)
    N.move_center_(psi, 1)
    h2 = C.eff_h2(C.EnvCache(topo), psi, H, 1, 0)
    return h2, C.two_site_tensor(psi, 1, 0)


def test_noalloc_apply_matches_allocating_apply():
    import torch
    from graft_amd import _native as nat
    from graft_amd.tensors import DTensor
    torch.cuda.set_device(0)
    L = nat.lib()
    ctx = ctypes.c_void_p()
    nat.check(L.graft_ctx_create(0, None, ctypes.byref(ctx)))
    try:
        for D, depth, seed in ((32, 4, 1), (8, 3, 2)):           # second call: a smaller map
            _say("noalloc apply, binary depth", depth, "D", D)
            h2, x = _map_and_input(D, depth, seed)
            assert h2.native.workspace_bytes <= 128 << 20
            st = nat.ptr_array([s.ptr for s in h2.statics])
            want = h2(x)
            y = DTensor.empty(x.dims, x.ncod)
            y.data.fill_(float("nan"))
            nat.check(L.graft_heff_apply_noalloc(ctx, h2.native.handle, ctypes.c_void_p(x.ptr), st,
                                                 ctypes.c_void_p(y.ptr), nat.PREC_FP64))
            torch.cuda.synchronize()
            assert torch.equal(y.data, want.data)
            ws = ctypes.c_uint64()
            nat.check(L.graft_ctx_workspace_bytes(ctx, ctypes.byref(ws)))
            assert ws.value == 0
        _say("noalloc apply refuses a plan above the static scratch")
        h2, x = _map_and_input(64, 4, 3)
        assert h2.native.workspace_bytes > 128 << 20
        st = nat.ptr_array([s.ptr for s in h2.statics])
        y = DTensor.empty(x.dims, x.ncod)
        rc = L.graft_heff_apply_noalloc(ctx, h2.native.handle, ctypes.c_void_p(x.ptr), st,
                                        ctypes.c_void_p(y.ptr), nat.PREC_FP64)
        assert rc == 2                                          # GRAFT_E_NOMEM
        assert b"static scratch" in L.graft_last_error()
    finally:
        torch.cuda.synchronize()
        L.graft_ctx_destroy(ctx)
