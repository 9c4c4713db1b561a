"""Host-side sector machinery (CPU): block QR/SVD, random symmetric states, charge sectors."""
import numpy as np

from oracle import heff_oracle as ref

from graft_amd import sectors as S
from graft_amd import trees as Tr
from graft_amd.planning import SectorLeg


def _tree(topo):
    return ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])


def test_block_qr_reproduces_and_is_isometric():
    rng = np.random.default_rng(0)
    L = lambda *s: SectorLeg([((q,), d) for q, d in s])
    legs = [L((0, 3), (1, 2)), L((0, 1), (1, 1)), L((0, 4), (-1, 5), (-2, 2)).dual().dual()]
    legs[2] = SectorLeg([((-q[0],), d) for q, d in [((0,), 4), ((1,), 5), ((2,), 2)]])
    T = S.random_blocks(rng, legs, 2)
    Q, R = S.qr_left(T)
    back = S.contract_leg(Q, 2, R, 0, 1)
    assert np.abs(back.dense() - T.dense()).max() < 1e-12
    q = Q.dense().reshape(-1, Q.dense().shape[-1], order="F")
    assert np.allclose(q.conj().T @ q, np.eye(q.shape[1]), atol=1e-12)


def test_block_qr_isolating_a_middle_leg_matches_dense_qr():
    """The QR of a link split that isolates a child leg (tdvp.jl:291-301,
    Backend.jl:602-609) on block tensors: Q's column space over the other legs
    is the dense QR's, Q . R rebuilds T, and a sector with fewer rows than
    columns keeps a bond degeneracy equal to its row count."""
    rng = np.random.default_rng(2)
    L = lambda *s: SectorLeg([((q,), d) for q, d in s])
    # flat charges: leg0 + leg1 + leg2 + leg3 = 0; the charge-2 sector of leg 1 (isolated)
    # meets (leg0, leg3) = (0, -2) and (1, -3): 2 + 1 = 3 rows against 4 columns
    legs = [L((0, 2), (1, 1)), L((0, 2), (1, 3), (2, 4)), L((0, 1)), L((-1, 3), (-2, 1), (-3, 1))]
    T = S.random_blocks(rng, legs, 3)
    k = 1
    Q, R = S.qr_isolating(T, k)
    assert [l.dim for l in Q.legs[:k] + Q.legs[k + 1:]] == [l.dim for l in T.legs[:k] + T.legs[k + 1:]]
    back = S.contract_leg(Q, k, R, 0, 1)                  # R :: new <- old on leg k
    assert np.abs(back.dense() - T.dense()).max() < 1e-12
    # rank bound per sector: rows over the other legs with the complementary charge
    rows = {}
    for key, blk in T.blocks.items():
        rows[key[k]] = rows.get(key[k], 0) + blk.size // blk.shape[k]
    degs = {q: d for q, d in Q.legs[k].sectors}
    for s_, (q, d) in enumerate(T.legs[k].sectors):
        assert degs.get(q, 0) == min(d, rows.get(s_, 0)), (q, d, rows.get(s_))
    assert Q.legs[k].dim < T.legs[k].dim                   # the deficient sector was reduced
    # dense comparison: isolate leg k as columns
    perm = [j for j in range(4) if j != k] + [k]
    Td = np.transpose(T.dense(), perm).reshape(-1, T.legs[k].dim, order="F")
    Qd = np.transpose(Q.dense(), perm).reshape(-1, Q.legs[k].dim, order="F")
    assert np.allclose(Qd.conj().T @ Qd, np.eye(Qd.shape[1]), atol=1e-12)
    U, sv, _ = np.linalg.svd(Td, full_matrices=False)
    r = int(np.sum(sv > 1e-10 * sv[0]))
    assert r == Qd.shape[1]
    assert np.abs(Qd @ Qd.conj().T - U[:, :r] @ U[:, :r].conj().T).max() < 1e-10


def test_block_svd_matches_dense_spectrum():
    rng = np.random.default_rng(1)
    topo = Tr.star_topology(2, 1)
    host = S.random_hubbard_ttns_host(rng, topo, 16)
    # two-site tensor of (child 1, root 0) on the host: A1 . A0
    A, B0 = host[1], host[0]
    blocks = {}
    for ka, va in A.blocks.items():
        for kb, vb in B0.blocks.items():
            if kb[0] == ka[-1]:
                key = ka[:-1] + kb[1:]
                v = np.tensordot(va, vb, axes=([A.dense().ndim - 1], [0]))
                blocks[key] = blocks.get(key, 0) + v
    theta = S.HostBlocks(A.legs[:-1] + B0.legs[1:], blocks, 4)
    U, Sv, Vh, spec, disc, bond = S.svd_split(theta, 1, lambda s: len(s))
    dense = theta.dense()
    sv = np.linalg.svd(dense.reshape(dense.shape[0], -1, order="F"), compute_uv=False)
    assert np.allclose(np.sort(spec)[::-1], sv[:len(spec)], atol=1e-12)
    assert disc < 1e-12


def test_random_hubbard_state_is_canonical_and_in_sector():
    rng = np.random.default_rng(2)
    topo = Tr.star_topology(3, 1)
    t = _tree(topo)
    host = S.random_hubbard_ttns_host(rng, topo, 12)
    dense = [h.dense() for h in host]
    v = ref.ttns_to_dense(t, dense, [True] * t.n)
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    for n in range(1, t.n):
        a = dense[n].reshape(-1, dense[n].shape[-1], order="F")
        assert np.allclose(a.conj().T @ a, np.eye(a.shape[1]), atol=1e-12)
    _, c = ref.hubbard_dense(t)
    N, Sz2 = ref.number_operators(c, t.n)
    assert np.linalg.norm(N @ v - t.n * v) < 1e-10          # half filling
    assert np.linalg.norm(Sz2 @ v) < 1e-10                 # S_z = 0


def test_hubbard_bond_sectors_follow_config():
    """D = 1024 split over a few dozen sectors, largest block ~ 100, smallest 1."""
    leg = S.hubbard_bond_leg(8, 16, 1024, 1.6, 16)
    degs = leg.degs()
    assert 30 <= len(degs) <= 70 and 60 <= max(degs) <= 160 and min(degs) == 1
    assert 900 <= sum(degs) <= 1100


def test_two_site_tensor_spans_the_full_invariant_space():
    """Theta = A_n A_m is stored on every charge-conserving block of its legs,
    and its dense embedding is the dense product (16-site tree, D=24)."""
    rng = np.random.default_rng(6)
    edges = [(str((i - 1) // 2), str(i)) for i in range(1, 16)]
    topo = Tr.TreeTopology("0", edges)
    host = S.random_hubbard_ttns_host(rng, topo, 24, sigma=1.6)
    n, m = 1, 0
    k = Tr.childslot(topo, m, n)
    th = S.two_site_host(host[n], host[m], k)
    assert set(th.blocks) == set(S.invariant_blocks(th.legs))
    An, Am = host[n].dense(), host[m].dense()
    want = np.moveaxis(np.tensordot(An, Am, axes=([An.ndim - 1], [k])), [], [])
    assert np.abs(th.dense() - want).max() < 1e-12
