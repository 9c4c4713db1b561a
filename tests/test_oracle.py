"""Pin the numpy oracle against closed forms, invariants and the reference's known answers.

Each check is chosen so that a plausible mistake (wrong Kronecker order,
wrong leg label, missing conjugation, wrong coefficient) fails it.  CPU only.
"""
import numpy as np
import pytest

from oracle import heff_oracle as ref

from graft_amd import operators as O
from graft_amd import trees as Tr


def _product_tree(t):
    """Oracle Tree with the same node order as a product TreeTopology."""
    return ref.Tree(list(t.ids), list(t.parent), [list(c) for c in t.children])


# --------------------------------------------------------------------------
# dense Hamiltonians
# --------------------------------------------------------------------------


@pytest.mark.parametrize("n,J,g", [(2, 1.0, 0.0), (3, 0.0, 0.7), (6, 1.0, 0.9), (7, 0.8, 1.3)])
def test_tfi_chain_matches_free_fermions(n, J, g):
    t = ref.chain_tree(n)
    # chain edges run through consecutive sites in any rooting
    H = ref.dense_hamiltonian(n, ref.tfi_terms(t, J, g))
    E0 = np.linalg.eigvalsh(H)[0]
    assert abs(E0 - ref.tfi_chain_exact_energy(n, J, g)) < 1e-10


def test_heisenberg_two_site_singlet_triplet():
    t = ref.chain_tree(2)
    H = ref.dense_hamiltonian(2, ref.heisenberg_terms(t, J=1.3))
    w = np.sort(np.linalg.eigvalsh(H))
    assert np.allclose(w, [-0.75 * 1.3, 0.25 * 1.3, 0.25 * 1.3, 0.25 * 1.3], atol=1e-13)


def test_kron_order_first_site_fastest():
    H = ref.dense_hamiltonian(3, [(1.0, [(0, ref.SZ)])])
    diag = np.real(np.diag(H))
    assert np.allclose(diag, [1 - 2 * (j % 2) for j in range(8)])
    H2 = ref.dense_hamiltonian(3, [(1.0, [(2, ref.SZ)])])
    assert np.allclose(np.real(np.diag(H2)), [1 - 2 * ((j // 4) % 2) for j in range(8)])


# --------------------------------------------------------------------------
# TTNO builders (product, host side) against the dense Hamiltonian
# --------------------------------------------------------------------------


@pytest.mark.parametrize("topo_fn", [lambda: Tr.star_topology(3, 1), lambda: Tr.binary_topology(2),
                                     lambda: Tr.mps_topology(4), lambda: Tr.star_topology(2, 2)])
def test_ttno_builders_equal_dense(topo_fn):
    topo = topo_fn()
    t = _product_tree(topo)
    n = t.n
    for W, terms in [(O.tfi_host(topo, 1.0, 0.9), ref.tfi_terms(t, 1.0, 0.9)),
                     (O.heisenberg_host(topo, 1.0), ref.heisenberg_terms(t, 1.0))]:
        Hd = ref.dense_hamiltonian(n, terms)
        Ht = ref.ttno_to_dense(t, W)
        assert np.abs(Ht - Hd).max() < 1e-12
    rng = np.random.default_rng(5)
    Js = rng.uniform(0.5, 1.5, n)
    hs = rng.uniform(0.5, 1.5, n)
    Hd = ref.dense_hamiltonian(n, ref.tfi_terms_random(t, Js, hs))
    assert np.abs(ref.ttno_to_dense(t, O.tfi_random_host(topo, Js, hs)) - Hd).max() < 1e-12


def _leaf_heisenberg_chain(t, sites, J=1.0):
    """heisenberg_chain(sites) (test/runtests.jl:32-41) on consecutive listed sites."""
    terms = []
    for a, b in zip(sites[:-1], sites[1:]):
        terms += [(J / 2, [(a, ref.SP), (b, ref.SM)]), (J / 2, [(a, ref.SM), (b, ref.SP)]),
                  (J / 4, [(a, ref.SZ), (b, ref.SZ)])]
    return terms


@pytest.mark.parametrize("topo_fn", [lambda: Tr.star_topology(3, 1), lambda: Tr.binary_topology(2),
                                     lambda: Tr.mps_topology(4)])
def test_terms_ttno_equals_dense(topo_fn):
    """The general term automaton on every node equals the dense Hamiltonian,
    for edge terms and for terms between distant sites."""
    topo = topo_fn()
    t = _product_tree(topo)
    n = t.n
    rng = np.random.default_rng(9)
    terms = ref.tfi_terms(t, 1.0, 0.9) + _leaf_heisenberg_chain(t, list(rng.permutation(n)))
    W = O.terms_ttno_host(topo, [2] * n, terms)
    assert np.abs(ref.ttno_to_dense(t, W) - ref.dense_hamiltonian(n, terms)).max() < 1e-12


def test_terms_ttno_physless_internal_nodes():
    """Sites only on the leaves of binary(2) (root and inner nodes have no
    physical leg), Heisenberg chain through the leaves as in
    test/runtests.jl:781-804: the TTNO equals the dense chain Hamiltonian."""
    topo = Tr.binary_topology(2)
    t = _product_tree(topo)
    leaves = [i for i in range(t.n) if not t.children[i]]
    phys = [2 if i in leaves else 0 for i in range(t.n)]
    pos = {s: k for k, s in enumerate(leaves)}
    terms = _leaf_heisenberg_chain(t, leaves)
    W = O.terms_ttno_host(topo, phys, terms)
    assert all(W[i].ndim == len(t.children[i]) + (3 if phys[i] else 1) for i in range(t.n))
    local = [(c, [(pos[s], o) for s, o in ops]) for c, ops in terms]
    Hd = ref.dense_hamiltonian(len(leaves), local)
    assert np.abs(ref.ttno_to_dense(t, W, [bool(p) for p in phys]) - Hd).max() < 1e-12
    # the chain's ground energy (4-site open Heisenberg, J=1: -(3 + 2 sqrt 3)/4)
    assert abs(np.linalg.eigvalsh(Hd)[0] + (3 + 2 * np.sqrt(3)) / 4) < 1e-12


def test_ttno_bond_dims():
    topo = Tr.binary_topology(2)
    assert O.heisenberg_host(topo)[1].shape[-1] == 5
    assert O.tfi_host(topo)[1].shape[-1] == 3


# --------------------------------------------------------------------------
# states and environments
# --------------------------------------------------------------------------


def _setup(topo_fn=lambda: ref.star_tree(3, 2), D=3, seed=3, center=0):
    t = topo_fn()
    rng = np.random.default_rng(seed)
    psi = ref.random_ttns(rng, t, 2, D, center=center)
    phys = [True] * t.n
    return t, psi, phys


def test_random_ttns_canonical():
    t, psi, phys = _setup()
    v = ref.ttns_to_dense(t, psi, phys)
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    for n in range(1, t.n):
        A = psi[n].reshape(-1, psi[n].shape[-1], order="F")
        assert np.allclose(A.conj().T @ A, np.eye(A.shape[1]), atol=1e-12)


def test_move_center_preserves_state():
    t, psi, phys = _setup()
    v = ref.ttns_to_dense(t, psi, phys)
    psi2 = ref.move_center(t, psi, 0, 5)
    assert np.linalg.norm(ref.ttns_to_dense(t, psi2, phys) - v) < 1e-12
    A = np.moveaxis(psi2[0], 2, -1)  # centre left node 0; node 0 must be isometric towards leg 2 (child b3_1)
    M = A.reshape(-1, A.shape[-1], order="F")
    assert np.allclose(M.conj().T @ M, np.eye(M.shape[1]), atol=1e-12)


def test_env_closure_is_expectation():
    t, psi, phys = _setup(D=4)
    W = O.tfi_host(Tr.star_topology(3, 2), 1.0, 0.9)
    Hd = ref.dense_hamiltonian(t.n, ref.tfi_terms(t, 1.0, 0.9))
    v = ref.ttns_to_dense(t, psi, phys)
    envs = ref.all_envs(t, psi, W)
    e_dense = np.vdot(v, Hd @ v)
    for u in range(t.n):
        assert abs(ref.env_closure(t, psi, W, envs, u) - e_dense) < 1e-12


def _matrix_of(apply, shape):
    size = int(np.prod(shape))
    cols = []
    for j in range(size):
        e = np.zeros(size, dtype=complex)
        e[j] = 1
        cols.append(apply(e.reshape(shape, order="F")).reshape(-1, order="F"))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("center", [0, 1, 2])
def test_h1_tangent_identity(center):
    t, psi, phys = _setup(D=3, center=center)
    W = O.heisenberg_host(Tr.star_topology(3, 2), 1.0)
    Hd = ref.dense_hamiltonian(t.n, ref.heisenberg_terms(t, 1.0))
    envs = ref.all_envs(t, psi, W)
    h = _matrix_of(lambda x: ref.h1_apply(t, W, envs, center, x), psi[center].shape)
    P = ref.embedding_h1(t, psi, phys, center)
    assert np.allclose(P.conj().T @ P, np.eye(P.shape[1]), atol=1e-12)
    assert np.abs(h - P.conj().T @ Hd @ P).max() < 1e-12
    assert np.abs(h - h.conj().T).max() < 1e-12


@pytest.mark.parametrize("n", [1, 2, 4])
def test_h2_tangent_identity(n):
    t, psi, phys = _setup(D=3, center=n)
    m = t.parent[n]
    W = O.tfi_host(Tr.star_topology(3, 2), 1.0, 0.9)
    Hd = ref.dense_hamiltonian(t.n, ref.tfi_terms(t, 1.0, 0.9))
    envs = ref.all_envs(t, psi, W)
    theta = ref.two_site_tensor(t, psi, n, m)
    h = _matrix_of(lambda x: ref.h2_apply(t, W, envs, n, m, x), theta.shape)
    P = ref.embedding_h2(t, psi, phys, n, m)
    assert np.allclose(P.conj().T @ P, np.eye(P.shape[1]), atol=1e-12)
    assert np.abs(h - P.conj().T @ Hd @ P).max() < 1e-12
    # the state itself: Theta is the embedding preimage of psi
    v = ref.ttns_to_dense(t, psi, phys)
    assert np.linalg.norm(P @ theta.reshape(-1, order="F") - v) < 1e-12


def test_h2_binary_root_edge_identity():
    t = ref.binary_tree(2)
    rng = np.random.default_rng(4)
    psi = ref.random_ttns(rng, t, 2, 3, center=1)
    phys = [True] * t.n
    W = O.heisenberg_host(Tr.binary_topology(2))
    Hd = ref.dense_hamiltonian(t.n, ref.heisenberg_terms(t))
    envs = ref.all_envs(t, psi, W)
    theta = ref.two_site_tensor(t, psi, 1, 0)
    h = _matrix_of(lambda x: ref.h2_apply(t, W, envs, 1, 0, x), theta.shape)
    P = ref.embedding_h2(t, psi, phys, 1, 0)
    assert np.abs(h - P.conj().T @ Hd @ P).max() < 1e-12


def test_h0_tangent_identity():
    t, psi, phys = _setup(D=3, center=1)
    n, m = 1, 0
    # split the centre at n: psi[n] = Q C, Q isometric towards its parent leg
    Q, C = ref.left_orth(psi[n])
    split = list(psi)
    split[n] = Q
    W = O.tfi_host(Tr.star_topology(3, 2), 1.0, 0.9)
    Hd = ref.dense_hamiltonian(t.n, ref.tfi_terms(t, 1.0, 0.9))
    envs = ref.all_envs(t, split, W)
    h = _matrix_of(lambda x: ref.h0_apply(envs, n, m, x), C.shape)
    P = ref.embedding_h0(t, split, phys, n, m)
    assert np.allclose(P.conj().T @ P, np.eye(P.shape[1]), atol=1e-12)
    assert np.abs(h - P.conj().T @ Hd @ P).max() < 1e-12


# --------------------------------------------------------------------------
# cost model known answers (lib/GraftPlanning/test/runtests.jl:121-204)
# --------------------------------------------------------------------------


def test_dense_cost_fixture():
    m = ref.dense_cost([-1, 1], [2, 3], [1, -2], [3, 4])
    assert m["oindA"] == [0] and m["cindA"] == [1] and m["oindB"] == [1] and m["cindB"] == [0]
    assert m["labels"] == [-1, -2] and m["dims"] == [2, 4] and m["peak_elements"] == 8
    assert m["flops"] > 0


def test_u1_chain_known_answers():
    left = ({0: 1, 1: 1}, {0: 1, 1: 2})
    middle = ({0: 1, 1: 2}, {0: 1, 1: 4})
    right = ({0: 1, 1: 4}, {0: 2, 1: 1})
    assert ref.u1_chain_plan_costs([left, middle, right], "left") == (60.0, 30.0)
    assert ref.u1_chain_plan_costs([left, middle, right], "right") == (63.0, 28.0)


def test_execute_pair_steps_equals_ncon():
    rng = np.random.default_rng(9)
    A = ref.random_complex(rng, (3, 4, 5))
    B = ref.random_complex(rng, (5, 6, 3))
    C = ref.random_complex(rng, (6, 2))
    want = ref.ncon([A, B, C], [[1, -1, 2], [2, 3, 1], [3, -2]], [False, True, False])
    # (A . conj B) then . C
    steps = [(0, 1, 3, (1,), (0, 2), (1,), (2, 0), False, True, (0, 1)),
             (3, 2, 4, (0,), (1,), (1,), (0,), False, False, (0, 1))]
    got = ref.execute_pair_steps(steps, [A, B, C], 5, 4)
    assert np.abs(got - want).max() < 1e-12


def test_exact_evolve_unitary_and_groundstate():
    t = ref.chain_tree(4)
    H = ref.dense_hamiltonian(4, ref.tfi_terms(t, 1.0, 0.9))
    E0, v0 = ref.exact_groundstate(H)
    assert np.linalg.norm(H @ v0 - E0 * v0) < 1e-10
    rng = np.random.default_rng(2)
    v = ref.random_complex(rng, 16)
    v /= np.linalg.norm(v)
    w = ref.exact_evolve(H, v, -0.3j)
    assert abs(np.linalg.norm(w) - 1) < 1e-12
    back = ref.exact_evolve(H, w, 0.3j)
    assert np.linalg.norm(back - v) < 1e-12


# --------------------------------------------------------------------------
# spinful fermions (Hubbard, U(1) x U(1))
# --------------------------------------------------------------------------


def _tfi_tdvp_setup(D, seed=1):
    t = ref.binary_tree(2)
    Hd = ref.dense_hamiltonian(t.n, ref.tfi_terms(t, 1.0, 0.9))
    psi = ref.random_ttns(np.random.default_rng(seed), t, 2, D)
    return t, Hd, psi, ref.ttns_to_dense(t, psi, [True] * t.n)


@pytest.mark.parametrize("order", [1, 2])
def test_oracle_tdvp1_is_exact_at_full_bonds(order):
    """7-site binary TFI tree at D=8 (every bond holds its subtree's whole
    space): TDVP1 of either order equals exp(dz H) (test/runtests.jl:910-925).
    A link step evolved forward, or an order-2 sweep without the full terminal
    step, misses this by more than 0.07."""
    t, Hd, psi, v0 = _tfi_tdvp_setup(8)
    p, c = list(psi), 0
    for _ in range(3):
        p, c = ref.tdvp1_step(t, p, c, Hd, -0.05j, order)
    v = ref.ttns_to_dense(t, p, [True] * t.n)
    assert np.linalg.norm(v - ref.exact_evolve(Hd, v0, -0.15j)) < 1e-12
    assert abs(np.linalg.norm(p[c]) - 1) < 1e-12          # the returned centre carries the norm


def test_oracle_tdvp1_conserves_and_is_reversible_at_truncated_bonds():
    """D=3 (truncated): the projected evolution leaves the exact one, yet every
    sub-step is the exponential of a hermitian map, so the norm and the energy
    are conserved, and the symmetric order-2 step is undone by a step of -dz."""
    t, Hd, psi, v0 = _tfi_tdvp_setup(3)
    E0 = np.vdot(v0, Hd @ v0).real
    p, c = list(psi), 0
    for _ in range(3):
        p, c = ref.tdvp1_step(t, p, c, Hd, -0.05j, 2)
    v = ref.ttns_to_dense(t, p, [True] * t.n)
    assert np.linalg.norm(v - ref.exact_evolve(Hd, v0, -0.15j)) > 1e-3
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    assert abs(np.vdot(v, Hd @ v).real - E0) < 1e-12 * abs(E0)
    p, c = ref.tdvp1_step(t, list(psi), 0, Hd, -0.05j, 2)
    p, c = ref.tdvp1_step(t, p, c, Hd, 0.05j, 2)
    assert np.linalg.norm(ref.ttns_to_dense(t, p, [True] * t.n) - v0) < 1e-12


def test_fermion_operators_anticommute():
    t = ref.star_tree(2, 1)
    _, c = ref.hubbard_dense(t)
    I = np.eye(4 ** t.n)
    for k1 in c:
        for k2 in c:
            assert np.allclose(c[k1] @ c[k2].conj().T + c[k2].conj().T @ c[k1], I * (k1 == k2))
            assert np.allclose(c[k1] @ c[k2] + c[k2] @ c[k1], 0)


@pytest.mark.parametrize("U", [0.0, 4.0, 9.0])
def test_hubbard_dimer_closed_form(U):
    t = ref.chain_tree(2)
    H, c = ref.hubbard_dense(t, 1.0, U)
    N, S = ref.number_operators(c, 2)
    assert abs(ref.sector_ground_energy(H, N, S, 2, 0) - (U - np.sqrt(U * U + 16)) / 2) < 1e-12
    assert np.allclose(H @ N, N @ H) and np.allclose(H @ S, S @ H)


@pytest.mark.parametrize("topo_fn", [lambda: Tr.star_topology(3, 1), lambda: Tr.binary_topology(1),
                                     lambda: Tr.mps_topology(3), lambda: Tr.star_topology(2, 2),
                                     lambda: Tr.fork_topology(2, 1)])
def test_hubbard_ttno_equals_fermionic_hamiltonian(topo_fn):
    """The parity-channel automaton reproduces the Jordan–Wigner fermion Hamiltonian."""
    topo = topo_fn()
    t = _product_tree(topo)
    Hd, _ = ref.hubbard_dense(t, 1.0, 4.0)
    assert np.abs(ref.ttno_to_dense(t, O.hubbard_host(topo, 1.0, 4.0)) - Hd).max() < 1e-12


def test_hubbard_ttno_blocks_conserve_charge():
    topo = Tr.binary_topology(2)
    ch, ph = O.hubbard_channel_leg(), O.hubbard_phys_leg()
    for n, a in enumerate(O.hubbard_host(topo)):
        K = len(topo.children[n])
        from graft_amd.planning import SectorLeg
        up = SectorLeg([((0, 0), 1)]) if topo.parent[n] < 0 else ch
        O.blockify(a, [ch] * K + [ph, ph.dual(), up.dual()], K + 1)   # raises if not


def test_oracle_energy_gradient_against_dense_and_finite_differences():
    """The oracle's energy gradient (site tensors and TTNO couplings) on a
    7-site binary tree in a non-canonical gauge: energy vs the dense vector,
    site gradients vs central differences of the dense energy, coupling
    derivatives vs dense expectation values of the coupled terms."""
    from graft_amd import operators as O, trees as Tr
    topo = Tr.binary_topology(2)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    rng = np.random.default_rng(21)
    psi = []
    for n in range(t.n):
        shape = [3] * len(t.children[n]) + [2] + [1 if t.parent[n] < 0 else 3]
        psi.append(ref.random_complex(rng, tuple(shape)))
    Js = {c: rng.uniform(0.5, 1.5) for c in range(t.n) if t.parent[c] >= 0}
    hs = rng.uniform(0.5, 1.5, t.n)
    W = O.tfi_random_host(topo, Js, hs)
    terms = ref.tfi_terms_random(t, Js, hs)
    E, Z, grads, Rs = ref.expectation_gradient(t, psi, W)
    assert abs(E - ref.dense_energy(t, psi, terms)) < 1e-12 * abs(E)
    h = 1e-6
    for n in range(t.n):
        delta = ref.random_complex(rng, psi[n].shape)
        p1 = list(psi); p1[n] = psi[n] + h * delta
        p2 = list(psi); p2[n] = psi[n] - h * delta
        fd = (ref.dense_energy(t, p1, terms) - ref.dense_energy(t, p2, terms)) / (2 * h)
        an = 2 * np.real(np.vdot(grads[n], delta))
        assert abs(fd - an) < 1e-7 * max(1.0, abs(an)), (n, fd, an)
    zero = {c: 0.0 for c in Js}
    W0 = O.tfi_random_host(topo, zero, np.zeros(t.n))
    v = ref.ttns_to_dense(t, psi, [True] * t.n)
    nv = np.vdot(v, v).real
    for c in Js:
        p = t.parent[c]
        Wc = O.tfi_random_host(topo, {**zero, c: 1.0}, np.zeros(t.n))
        an = np.real(np.sum((Wc[p] - W0[p]) * Rs[p])) / Z.real
        Hc = ref.dense_hamiltonian(t.n, [(-1.0, [(c, ref.SZ), (p, ref.SZ)])])
        assert abs(an - np.vdot(v, Hc @ v).real / nv) < 1e-12, c
    for i in range(t.n):
        hv = np.zeros(t.n); hv[i] = 1.0
        Wi = O.tfi_random_host(topo, zero, hv)
        an = np.real(np.sum((Wi[i] - W0[i]) * Rs[i])) / Z.real
        Hi = ref.dense_hamiltonian(t.n, [(-1.0, [(i, ref.SX)])])
        assert abs(an - np.vdot(v, Hi @ v).real / nv) < 1e-12, i


def test_oracle_dmrg2_reaches_exact_ground_energy():
    """The host dmrg2 sweeps (the CPU sweep baseline) against exact
    diagonalization: TFI on the 3x2 star, maxdim 32 (runtests.jl:687-700)."""
    t = ref.star_tree(3, 2)
    psi = ref.random_ttns(np.random.default_rng(1), t, 2, 2)
    W = O.tfi_host(Tr.star_topology(3, 2), 1.0, 0.9)
    E0 = np.linalg.eigvalsh(ref.dense_hamiltonian(t.n, ref.tfi_terms(t, 1.0, 0.9)))[0]
    _, _, Es = ref.dmrg2_sweeps(t, psi, W, 0, 32, nsweeps=4)
    assert abs(Es[-1] - E0) < 1e-10, (Es, E0)
