"""Device per-sector factorizations against the host block factorizations.

The host forms (sectors.qr_isolating, sectors.svd_split) are pinned on the
CPU against dense QR / SVD (tests/test_sectors_host.py).  Here the device
forms must give the same factorization on the dense embedding: the same
column space (projector, gauge-free) and reconstruction for the block QR of
every leg, and the same spectrum, bond sectors, discarded norm and truncated
tensor for the block SVD of a two-site tensor.
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


def _tree16():
    from graft_amd import trees as Tr
    return Tr.TreeTopology("0", [(str((i - 1) // 2), str(i)) for i in range(1, 16)])


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _matricize(T, k):
    """Dense embedding of T as (all legs but k) x (leg k)."""
    N = T.ndim
    perm = [j for j in range(N) if j != k] + [k]
    X = np.transpose(T, perm)
    return X.reshape(-1, X.shape[-1], order="F")


@pytest.mark.parametrize("D", [24, 96])
def test_block_qr_every_leg_matches_host(g, D):
    from graft_amd import sector_linalg as SL, sectors as S
    topo = _tree16()
    _say("device block QR, 16-site Hubbard tree, D", D)
    psi, host = S.random_hubbard_ttns(np.random.default_rng(D), topo, D)
    for n in (0, 1, 3):
        A = psi.tensors[n]
        Ad = A.to_numpy()
        for k in range(A.rank):
            Q, R = SL.qr_isolating(A, k)
            Qh, Rh = S.qr_isolating(host[n], k)
            assert [d for _, d in Q.blocks.legs[k].sectors] == [d for _, d in Qh.legs[k].sectors]
            Qd, Rd = Q.to_numpy(), R.to_numpy()
            # reconstruction: contract R (new <- old) back onto leg k of Q
            rec = np.moveaxis(np.tensordot(Qd, Rd, axes=([k], [0])), -1, k)
            assert _rel(rec, Ad) < 1e-12, (n, k)
            Qm = _matricize(Qd, k)
            assert np.abs(Qm.conj().T @ Qm - np.eye(Qm.shape[1])).max() < 1e-12
            Qhm = _matricize(Qh.dense(), k)
            # the same column space: Qm lies in span(Qhm) (both isometric, same rank)
            assert np.linalg.norm(Qm - Qhm @ (Qhm.conj().T @ Qm)) < 1e-11 * np.linalg.norm(Qm), (n, k)
            # R is block diagonal in the bond sectors and upper triangular per sector
            for b in R.blocks.blocks:
                assert R.blocks.legs[0].sectors[b[0]][0] == tuple(-x for x in R.blocks.legs[1].sectors[b[1]][0])


@pytest.mark.parametrize("D,maxdim", [(32, 24), (128, 100)])
def test_block_svd_matches_host(g, D, maxdim):
    from graft_amd import backend as B, contractions as C, sector_linalg as SL, sectors as S
    topo = _tree16()
    _say("device block SVD, two-site tensor on edge (1, 0), D", D, "maxdim", maxdim)
    psi, host = S.random_hubbard_ttns(np.random.default_rng(7 + D), topo, D)
    theta = C.two_site_tensor(psi, 1, 0)
    nrow = psi.tensors[1].ncod
    Th = S.HostBlocks.from_device(theta)
    # a cut inside the nonzero spectrum with a gap: the kept sets are then unique
    full = S.svd_split(Th, nrow, lambda s: len(s))[3]
    k = min(maxdim, int(np.sum(full > 1e-8 * full[0])) - 1)
    while k > 1 and full[k - 1] < full[k] * (1 + 1e-6):
        k -= 1
    ts = B.TruncationScheme(maxdim=k)
    rank = lambda s: B.truncation_rank(s, ts)
    U, Sd, Vh, spec, disc, bond = SL.svd_split(theta, nrow, rank)
    Uh, Sh, Vhh, spech, disch, bondh = S.svd_split(Th, nrow, rank)
    assert sum(d for _, d in bond.sectors) == k
    big = spech >= 1e-5 * spech[0]
    assert np.abs(spec[big] / spech[big] - 1).max() < 1e-10
    assert np.abs(spec - spech).max() < 1e-13 * spech[0]
    assert bond.sectors == bondh.sectors
    assert abs(disc - disch) <= 1e-10 * max(disch, 1e-300)
    # the truncated tensor U S Vh equals the host one
    s_all = np.concatenate([Sd[i].detach().cpu().numpy() for i in range(len(bond.sectors))])
    Ud, Vd = U.to_numpy(), Vh.to_numpy()
    low = np.tensordot(Ud * s_all, Vd, axes=([Ud.ndim - 1], [0]))
    sh_all = np.concatenate([Sh[i] for i in range(len(bondh.sectors))])
    Uhd, Vhd = Uh.dense(), Vhh.dense()
    low_h = np.tensordot(Uhd * sh_all, Vhd, axes=([Uhd.ndim - 1], [0]))
    assert _rel(low, low_h) < 1e-10
    # absorb and bond slot: the same factorization with S in Vh and the bond at slot 1
    U2, _, V2, _, _, _ = SL.svd_split(theta, nrow, rank, absorb="vh", bond_slot=1, vh_ncod=2)
    V2d = np.moveaxis(V2.to_numpy(), 1, 0)
    assert _rel(np.tensordot(U2.to_numpy(), V2d, axes=([Ud.ndim - 1], [0])), low_h) < 1e-10


def test_permute2_and_to_structure(g):
    from graft_amd import sector_linalg as SL, sectors as S
    topo = _tree16()
    _say("permute2 / to_structure")
    psi, host = S.random_hubbard_ttns(np.random.default_rng(4), topo, 24)
    Q, R = SL.qr_isolating(psi.tensors[3])
    Rt = SL.permute2(R)
    assert np.array_equal(Rt.to_numpy(), R.to_numpy().T)
    from graft_amd.tensors import BlockStructure
    full = BlockStructure.invariant(R.blocks.legs)
    Rf = SL.to_structure(R, full, 1)
    assert np.array_equal(Rf.to_numpy(), R.to_numpy())
