"""Adjoint (reverse-mode) contractions of the effective-Hamiltonian network.

The reference exposes no derivative of its maps; the rrules here sit on its
effective-map seam (lib/GraftContractions/src/Contractions/effective.jl:709-730
for eff_h1, :986-1034 for eff_h2) and on the expectation value
(lib/GraftNetworks/src/Networks/expectation.jl:32-41).

Every adjoint is itself a tensor network executed by the same planner and
grouped GEMM kernels as the forward map: the derivative of a closed network
with respect to one operand is the network with that operand removed and its
labels left open.  For y = F(x, s_1..s_k) (multilinear, no operand
conjugated) and a cotangent ybar, the pullback is

    xbar   = F^H ybar          = conj(d <ybar, y> / d x)
    s_kbar = (dF/ds_k)^H ybar  = conj(d <ybar, y> / d s_k)

and conj(network) = network with every conjugation flag flipped, so each
cotangent is one contraction of (x, the other statics, ybar) with flipped
flags.  Environments are recomputed from the site tensors (EnvCache) rather
than stored from a forward pass.
"""
from __future__ import annotations

import numpy as np

from . import tensors as T
from .contractions import (EnvCache, _build_env_spec, eff_h1, env_, expect)
from .networks import TTNO
from .planning import ContractionSpec
from .tensors import DTensor
from .trees import neighbors, nchildren, nnodes


def derivative_spec(spec, j, ncod_j, flip_conj=False):
    """Network of d(closed network)/d(operand j): operand j removed, its labels
    become the open legs -1..-r in its leg order (``flip_conj`` conjugates the
    result)."""
    if spec.nopen != 0:
        raise ValueError("derivative_spec: the network must be closed")
    lab = spec.labels[j]
    if len(set(lab)) != len(lab):
        raise ValueError("derivative_spec: operand has a trace")
    rename = {L: -(i + 1) for i, L in enumerate(lab)}
    labels, conjs = [], []
    for i, (ls, c) in enumerate(zip(spec.labels, spec.conjs)):
        if i == j:
            continue
        labels.append([rename.get(L, L) for L in ls])
        conjs.append((not c) if flip_conj else c)
    r = len(lab)
    return ContractionSpec(labels, conjs, r, (ncod_j, r - ncod_j))


def closure_spec(map_spec):
    """<ybar, F(x, ...)> as a closed network: ybar joins as a conjugated operand."""
    top = max(abs(L) for ls in map_spec.labels for L in ls)
    rename = {-(i + 1): top + 1 + i for i in range(map_spec.nopen)}
    labels = [[rename.get(L, L) for L in ls] for ls in map_spec.labels]
    labels.append([top + 1 + i for i in range(map_spec.nopen)])
    conjs = list(map_spec.conjs) + [True]
    return ContractionSpec(labels, conjs, 0, (0, 0))


def _execute_spec(cache, kind, spec, operands, ncod):
    plan = cache.get_plan(kind, spec, operands)
    out = DTensor.empty(plan.out_dims, ncod, plan.out_structure)
    plan.execute(operands, out)
    return out


def rrule_apply(h, x):
    """rrule for y = h(x) with h an effective map: returns (y, pullback) where
    pullback(ybar) = (xbar, [cotangent of every static operand])."""
    if getattr(h, "spec", None) is None or getattr(h, "cache", None) is None:
        raise ValueError("rrule_apply needs a map built by eff_h1 / eff_h2 / eff_h0")
    y = h(x)
    cspec = closure_spec(h.spec)
    operands = [x] + list(h.statics)

    def pullback(ybar):
        ops = operands + [ybar]
        out = []
        for j, op in enumerate(operands):
            dspec = derivative_spec(cspec, j, op.ncod, flip_conj=True)
            rest = [o for i, o in enumerate(ops) if i != j]
            out.append(_execute_spec(h.cache, f"adjoint_{h.kind}_{j}", dspec, rest, op.ncod))
        return out[0], out[1:]

    return y, pullback


def identity_ttno(topo, d=2):
    """TTNO of the identity (channel dimension 1), for norms <psi|psi>."""
    tens = []
    for n in range(nnodes(topo)):
        K = nchildren(topo, n)
        a = np.zeros([1] * K + [d, d, 1], dtype=complex)
        a[(0,) * K + (slice(None), slice(None), 0)] = np.eye(d)
        tens.append(DTensor.from_numpy(a, K + 1))
    return TTNO(topo, tens, ishermitian=True)


class EnergyGradient:
    """E = <psi|H|psi>/<psi|psi>, its gradient with respect to every site
    tensor (Wirtinger: dE = 2 Re sum_n <g_n, dA_n>) and the derivative of
    <psi|H|psi> with respect to every TTNO tensor (dW-linear)."""

    def __init__(self, energy, norm2, site_grads, ttno_derivs):
        self.energy = energy
        self.norm2 = norm2
        self.site = site_grads
        self.ttno = ttno_derivs

    def coupling_gradient(self, dW):
        """dE/dp for a real parameter p with dW = {node: dW_n/dp (host array)}."""
        tot = 0.0
        for n, a in dW.items():
            R = self.ttno[n].to_numpy()
            tot += float(np.real(np.sum(np.asarray(a) * R)))
        return tot / self.norm2


def energy_and_gradient(psi, H, d=2):
    """Value and gradient of the energy functional in any gauge.

    Site gradient: g_n = (H_n A_n - E N_n A_n) / <psi|psi>, the one-site maps of
    H and of the identity applied to A_n (the adjoint of the closed network
    with respect to conj(A_n) is the forward one-site map).  TTNO derivative:
    the closure at n with W_n's legs left open."""
    topo = psi.topo
    cH, cI = EnvCache(topo), EnvCache(topo)
    I = identity_ttno(topo, d)
    num = expect(psi, H, cH)
    Z = expect(psi, I, cI).real
    E = (num / Z).real
    site = []
    for n in range(nnodes(topo)):
        A = psi.tensors[n]
        g = eff_h1(cH, psi, H, n)(A)
        gI = eff_h1(cI, psi, I, n)(A)
        T.axpby_(-E / Z, gI, 1.0 / Z, g)
        site.append(g)
    ttno = []
    for n in range(nnodes(topo)):
        for w in neighbors(topo, n):
            env_(cH, psi, H, w, n)
        spec, ops = _build_env_spec(cH, psi, H, psi, n, None)
        j = next(i for i, o in enumerate(ops) if o is H.tensors[n])
        dspec = derivative_spec(spec, j, H.tensors[n].ncod)
        rest = [o for i, o in enumerate(ops) if i != j]
        ttno.append(_execute_spec(cH, "dclosure_dW", dspec, rest, H.tensors[n].ncod))
    return EnergyGradient(E, Z, site, ttno)


def overlap(bra, ket, d=2):
    """<bra|ket> of two TTNS on the same tree (any gauge): the closure of the
    identity TTNO with ``bra`` on the conjugated side."""
    topo = ket.topo
    cache = EnvCache(topo)
    I = identity_ttno(topo, d)
    r = topo.root
    for w in neighbors(topo, r):
        env_(cache, ket, I, w, r, bra)
    spec, ops = _build_env_spec(cache, ket, I, bra, r, None)
    plan = cache.get_plan("overlap", spec, ops)
    out = DTensor.empty((), 0, plan.out_structure)
    plan.execute(ops, out)
    return complex(out.data.cpu().numpy()[0])


def energy(psi, H, d=2):
    """E = <psi|H|psi>/<psi|psi> in any gauge (closures of H and the identity)."""
    topo = psi.topo
    return float((expect(psi, H, EnvCache(topo)) / expect(psi, identity_ttno(topo, d), EnvCache(topo))).real)


def rrule_energy(psi, H, d=2):
    """rrule for E(psi, H): returns (E, pullback); pullback(Ebar) gives
    Ebar * dE/d conj(A_n) for every site (dE = 2 Re <g, dA>) and
    Ebar * dE/dW_n for every TTNO tensor (the holomorphic derivative
    <psi|.|psi>/<psi|psi>: dE/dp = sum (dW/dp) . dE/dW for a real coupling p
    that keeps H hermitian)."""
    G = energy_and_gradient(psi, H, d)

    def pullback(Ebar=1.0):
        sites = [g.clone() for g in G.site]
        for g in sites:
            T.scale_(g, float(Ebar))
        ttno = [R.clone() for R in G.ttno]
        for r in ttno:
            T.scale_(r, float(Ebar) / G.norm2)
        return sites, ttno

    return G.energy, pullback
