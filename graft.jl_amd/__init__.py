"""graft.jl_amd — MI355X-native effective-Hamiltonian engine for tree tensor networks.

The package directory name contains a dot, so it is imported under the module
name ``graft_amd`` (see ``load_package`` in ``__graft_entry__.py``).  Host
modules mirror the reference's API for the H_eff path:

* ``trees``        — Trees (topology.jl, geometries.jl)
* ``planning``     — Planning (ContractionSpec, PairStep, plan_contraction)
* ``tensors``      — device block store (TensorKit TensorMap blocks on HBM)
* ``networks``     — TTNS / TTNO, gauge moves, update_tensor_
* ``operators``    — closed-form nearest-neighbour TTNOs for the workloads
* ``contractions`` — EnvCache, env_, eff_h1 / eff_h2 / eff_h0, EffectiveMap
* ``krylov``       — Lanczos eigsolve / exponentiate on device vectors
* ``groundstate``  — dmrg2_
* ``evolution``    — TDVP1, step_, evolve_

Every contraction runs in libgraft_heff.so (hand-written gfx950 kernels);
there is no CPU fallback.
"""
__version__ = "0.1.0"
