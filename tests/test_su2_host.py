"""Host side of the SU(2) block path, checked on the CPU.

Every piece is pinned by what the mathematics fixes, independently of any
convention of the code: Clebsch-Gordan values against textbook entries, the
coupled states against the spin operators' action, fusion-tree tensors
against invariance under random SU(2) rotations, and every coupled
contraction against the dense contraction of the embedded tensors.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sl

from graft_amd import su2
from oracle import heff_oracle as ref


def _rotation(tj, theta, conj=False):
    jx, jy, jz = su2.spin_matrices(tj)
    u = sl.expm(-1j * (theta[0] * jx + theta[1] * jy + theta[2] * jz))
    return u.conj() if conj else u


def _rotate(X, tjs, duals, theta):
    for l, (tj, d) in enumerate(zip(tjs, duals)):
        X = np.moveaxis(np.tensordot(_rotation(tj, theta, d), X, axes=([1], [l])), 0, l)
    return X


@pytest.mark.parametrize("args,value", [
    ((1, 1, 1, -1, 0, 0), 1 / math.sqrt(2)),        # <1/2 1/2; 1/2 -1/2 | 0 0>
    ((1, -1, 1, 1, 0, 0), -1 / math.sqrt(2)),
    ((1, 1, 1, 1, 2, 2), 1.0),
    ((1, 1, 1, -1, 2, 0), 1 / math.sqrt(2)),
    ((2, 2, 2, -2, 0, 0), 1 / math.sqrt(3)),        # <1 1; 1 -1 | 0 0>
    ((2, 0, 2, 0, 0, 0), -1 / math.sqrt(3)),
    ((2, 2, 1, -1, 3, 1), math.sqrt(1 / 3)),        # <1 1; 1/2 -1/2 | 3/2 1/2>
    ((2, 0, 1, 1, 3, 1), math.sqrt(2 / 3)),
    ((2, 2, 1, -1, 1, 1), math.sqrt(2 / 3)),        # <1 1; 1/2 -1/2 | 1/2 1/2>
    ((2, 0, 1, 1, 1, 1), -math.sqrt(1 / 3)),
    ((2, 2, 2, 0, 2, 2), 1 / math.sqrt(2)),         # <1 1; 1 0 | 1 1>
    ((2, 0, 2, 2, 2, 2), -1 / math.sqrt(2)),
])
def test_clebsch_gordan_known_values(args, value):
    assert abs(su2.clebsch_gordan(*args) - value) < 1e-15


@pytest.mark.parametrize("tj1,tj2", [(1, 1), (2, 1), (2, 2), (3, 4), (5, 2)])
def test_coupled_states_carry_the_total_spin(tj1, tj2):
    """sum CG[m1, m2, M] |m1>|m2> is |J M>: the total spin operators act on it
    as spin J on M, and the coupled states of all J are orthonormal."""
    S1, S2 = su2.spin_matrices(tj1), su2.spin_matrices(tj2)
    cols = []
    for tJ in su2.fusion(tj1, tj2):
        C = su2.cg_tensor(tj1, tj2, tJ)
        SJ = su2.spin_matrices(tJ)
        for a in range(3):
            lhs = np.einsum("ai,ijM->ajM", S1[a], C) + np.einsum("bj,ijM->ibM", S2[a], C)
            rhs = np.einsum("ijN,NM->ijM", C, SJ[a])
            assert np.abs(lhs - rhs).max() < 1e-13
        cols.append(C.reshape(-1, tJ + 1))
    U = np.concatenate(cols, axis=1)
    assert U.shape[0] == U.shape[1]
    assert np.abs(U.T @ U - np.eye(U.shape[0])).max() < 1e-13


@pytest.mark.parametrize("tjs,duals", [
    ((1, 1), (False, True)),
    ((1, 2, 1), (False, False, True)),
    ((1, 1, 1, 1), (False, True, False, True)),
    ((2, 1, 3, 2, 0), (True, False, False, True, False)),
    ((1, 2, 2, 1, 2, 2), (False, False, True, True, False, True)),
])
def test_tree_basis_is_an_orthonormal_invariant_basis(tjs, duals):
    trees, M = su2.tree_basis(tjs, duals)
    assert len(trees) >= 1
    assert np.abs(M @ M.T - np.eye(len(trees))).max() < 1e-13
    rng = np.random.default_rng(7)
    for tr in trees:
        X = su2.tree_tensor(tjs, tr, duals)
        assert np.abs(_rotate(X.astype(complex), tjs, duals, rng.standard_normal(3)) - X).max() < 1e-12
    # the trees span every invariant: dimension = multiplicity of the trivial irrep
    mult = {0: 1}
    for tj in tjs:
        mult = su2.fuse_multiplicities(mult, {tj: 1})
    assert len(trees) == mult.get(0, 0)


def test_embedding_round_trip_and_invariance():
    L = su2.SU2Leg
    legs = [L([(1, 2), (3, 1)]), L([(1, 1)]), L([(0, 2), (2, 1)]), L([(0, 1), (2, 2)], True)]
    st = su2.SU2Structure.invariant(legs)
    rng = np.random.default_rng(1)
    f = su2.random_flat(rng, st)
    D = st.to_dense(f)
    assert np.abs(st.from_dense(D) - f).max() < 1e-13
    assert abs(np.linalg.norm(D) - np.linalg.norm(f)) < 1e-12       # orthonormal structural tensors
    th = rng.standard_normal(3)
    # rotate every leg: block-diagonal over its irreps (degeneracies untouched)
    R = D.copy()
    for l, leg in enumerate(legs):
        U = sl.block_diag(*[np.kron(np.eye(d), _rotation(tj, th, leg.isdual)) for tj, d in leg.irreps])
        R = np.moveaxis(np.tensordot(U, R, axes=([1], [l])), 0, l)
    assert np.abs(R - D).max() < 1e-12


def _leg(irreps, dual=False):
    return su2.SU2Leg(irreps, dual)


PAIR_CASES = [
    # A legs, B legs, cA, cB, perm, conjA
    ([_leg([(1, 2), (3, 1)]), _leg([(1, 1)]), _leg([(0, 2), (2, 1)]), _leg([(0, 1), (2, 2)], True)],
     [_leg([(1, 2)]), _leg([(0, 2), (2, 1)], True), _leg([(1, 1)], True), _leg([(0, 1), (2, 2)])],
     (1, 2), (2, 1), (2, 0, 3, 1), False),
    ([_leg([(1, 2), (3, 1)]), _leg([(1, 1)], True), _leg([(0, 2), (2, 1)], True), _leg([(0, 1), (2, 2)])],
     [_leg([(1, 2)]), _leg([(0, 2), (2, 1)], True), _leg([(1, 1)], True), _leg([(0, 1), (2, 2)])],
     (1, 2), (2, 1), (0, 1, 2, 3), True),
    ([_leg([(0, 1), (2, 2)]), _leg([(1, 2), (3, 1)], True), _leg([(1, 1)])],
     [_leg([(1, 2), (3, 1)]), _leg([(0, 2), (2, 1)]), _leg([(1, 3)], True)],
     (1,), (0,), (3, 1, 0, 2), False),
    ([_leg([(0, 2), (2, 1)]), _leg([(0, 1), (2, 2), (4, 1)])],
     [_leg([(0, 1), (2, 2), (4, 1)], True), _leg([(0, 2), (2, 1)], True)],
     (0, 1), (1, 0), (), False),
]


@pytest.mark.parametrize("case", range(len(PAIR_CASES)))
def test_coupled_step_equals_dense_contraction(case):
    legsA, legsB, cA, cB, perm, conjA = PAIR_CASES[case]
    SA = su2.SU2Structure.invariant(legsA)
    SB = su2.SU2Structure.invariant(legsB)
    rng = np.random.default_rng(case)
    fa, fb = su2.random_flat(rng, SA), su2.random_flat(rng, SB)
    oA = tuple(i for i in range(len(legsA)) if i not in cA)
    oB = tuple(i for i in range(len(legsB)) if i not in cB)
    stepA = [l.dual() for l in legsA] if conjA else legsA
    C, trip, cf = su2.compile_pair(SA, stepA, SB, legsB, oA, cA, oB, cB, perm)
    got = su2.host_execute_pair(C, trip, cf, SA, fa, conjA, SB, fb, False, oA, cA, oB, cB, perm)
    Ad = SA.to_dense(fa)
    Ad = Ad.conj() if conjA else Ad
    Bd = SB.to_dense(fb)
    nc = len(cA)
    want = np.tensordot(Ad.transpose(oA + cA), Bd.transpose(cB + oB),
                        axes=(list(range(len(oA), len(oA) + nc)), list(range(nc))))
    want = np.asarray(want).transpose(perm) if perm else np.asarray(want)
    assert np.abs(C.to_dense(got) - want).max() <= 1e-13 * max(1.0, np.abs(want).max())
    # every coupled block is reached, and the pairs are grouped by output block
    assert np.all(np.diff(trip[:, 0]) >= 0)


def test_heisenberg_ttno_is_the_reference_model():
    """The SU(2) TTNO's dense embedding is the reference's spin-1/2 Heisenberg
    TTNO (J/2 (S+S- + S-S+) + J/4 ZZ per edge, operators.heisenberg_host),
    and the spin-1 form is J sum S_i . S_j built from spin matrices."""
    from graft_amd import operators as O, trees as Tr
    for topo in (Tr.star_topology(3, 1), Tr.binary_topology(2)):
        t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
        Wd = [su2.heisenberg_dense_site(len(topo.children[n]), topo.parent[n] < 0, 0.7)
              for n in range(len(topo.ids))]
        H1 = ref.ttno_to_dense(t, Wd)
        H2 = ref.ttno_to_dense(t, O.heisenberg_host(topo, 0.7))
        assert np.abs(H1 - H2).max() < 1e-13
    topo = Tr.mps_topology(3)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    Wd = [su2.heisenberg_dense_site(len(topo.children[n]), topo.parent[n] < 0, 1.0, 2)
          for n in range(3)]
    S = su2.spin_matrices(2)
    H = np.zeros((27, 27), dtype=complex)
    for c, p in [(c, topo.parent[c]) for c in range(3) if topo.parent[c] >= 0]:
        for a in range(3):
            ops = [np.eye(3)] * 3
            ops[c], ops[p] = S[a], S[a]
            # site i of the dense vector is tensor axis i, first site fastest
            H += np.kron(np.kron(ops[2], ops[1]), ops[0])
    assert np.abs(ref.ttno_to_dense(t, Wd) - H).max() < 1e-13
    for n in range(3):
        K = len(topo.children[n])
        legs = [su2.CHANNEL] * K + [su2.SU2Leg([(2, 1)]), su2.SU2Leg([(2, 1)], True),
                                   (su2.SU2Leg([(0, 1)]) if topo.parent[n] < 0 else su2.CHANNEL).dual()]
        st = su2.SU2Structure.invariant(legs)
        assert np.abs(st.to_dense(st.from_dense(Wd[n])) - Wd[n]).max() < 1e-13


def test_left_orthonormalize_gives_a_dense_isometry():
    L = su2.SU2Leg
    legs = [L([(1, 1), (3, 1)]), L([(1, 1)]), L([(0, 1), (2, 2)], True)]
    st = su2.SU2Structure.invariant(legs)
    f = su2.left_orthonormalize(st, su2.random_flat(np.random.default_rng(2), st))
    D = st.to_dense(f).reshape(-1, st.dims[-1], order="F")
    assert np.abs(D.conj().T @ D - np.eye(st.dims[-1])).max() < 1e-12


def test_physless_root_heisenberg_is_the_spin_model():
    """With no site at the root, the TTNO's dense embedding is J sum S_i . S_j
    over the tree's edges between sites plus the pair of the root's children,
    built from spin matrices (site i of the dense vector is axis i, first
    site fastest), and every tensor is exactly SU(2) invariant."""
    from graft_amd import trees as Tr
    topo = Tr.binary_topology(2)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    flags = [n != topo.root for n in range(7)]
    Wd = [su2.heisenberg_dense_site(len(topo.children[n]), topo.parent[n] < 0, 0.8, 1, flags[n])
          for n in range(7)]
    for n in range(7):
        K = len(topo.children[n])
        site = [su2.SU2Leg([(1, 1)]), su2.SU2Leg([(1, 1)], True)] if flags[n] else []
        legs = [su2.CHANNEL] * K + site + [(su2.SU2Leg([(0, 1)]) if n == topo.root else su2.CHANNEL).dual()]
        st = su2.SU2Structure.invariant(legs)
        assert np.abs(st.to_dense(st.from_dense(Wd[n])) - Wd[n]).max() < 1e-13
    sites = [n for n in range(7) if flags[n]]
    S = su2.spin_matrices(1)

    def op(n, M):
        out = np.ones((1, 1))
        for s in reversed(sites):
            out = np.kron(out, M if s == n else np.eye(2))
        return out

    edges = [(c, topo.parent[c]) for c in sites if flags[topo.parent[c]]] + [tuple(topo.children[topo.root])]
    assert len(edges) == 5
    H = sum(0.8 * op(a, S[k]) @ op(b, S[k]) for a, b in edges for k in range(3))
    assert np.abs(ref.ttno_to_dense(t, Wd, flags) - H).max() < 1e-13
