"""Benchmark: eff_h2 applies/s on one MI355X (complex FP64), with roofline and CPU baseline.

Workload (the configuration the metric is quoted on): spin-1/2 Heisenberg
model (J = 1) on a balanced binary tree of depth 5 (63 sites), TTNO bond
dimension 5, random complex TTNS with bond dimension D = 256 on every bond
(capped by the subtree dimension after QR canonicalization), and the
two-site effective Hamiltonian on the root edge (node n1_1 and the root):
Theta has 256*256*2*256*2 = 6.7e7 complex entries.  One step = one apply
y = H_eff x with x resident in HBM.  Synthetic data: numpy seeded generator.

Prints one JSON line (the driver's contract).  ``--gpus N`` with N > 1 exits:
this engine runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_FP64_MFMA_TFLOPS = 78.6
THREE_M_CONFIGS = (7, 8, 9, 11, 12, 13, 14, 15, 16, 17, 18)     # tile configurations of the three-product complex form
MI355X_HBM_GBPS = 8000.0
# HBM bytes per launch of the dominant step, from the two counter runs
# (FETCH_SIZE and WRITE_SIZE, scaled by a 1 GiB copy) summarized in this file.
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "zgemm_heff_apply_D256.txt")


def profiled_traffic(step, D):
    """HBM bytes per launch of plan step ``step`` recorded by the counter runs, or None."""
    import re
    if D != 256 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    m = re.search(r"^traffic step (\d+): ([0-9.e+]+) bytes", open(TRAFFIC_PROFILE).read(), re.M)
    if m is None or int(m.group(1)) != step:
        return None
    return float(m.group(2))


def build_workload(D, depth=5, seed=7):
    import __graft_entry__
    g = __graft_entry__.load_package()
    from graft_amd import contractions as C, networks as N, operators as O, trees as Tr
    topo = Tr.binary_topology(depth)
    rng = np.random.default_rng(seed)
    psi = N.random_ttns(rng, topo, 2, D)
    H = O.heisenberg_ttno(topo, 1.0)
    cache = C.EnvCache(topo)
    n, m = 1, 0
    h2 = C.eff_h2(cache, psi, H, n, m)
    x = C.two_site_tensor(psi, n, m)
    return g, topo, psi, H, cache, h2, x


def cpu_baseline(D_cpu, depth=5, seed=7, budget_s=20.0):
    """The oracle executes the same pairwise plan (permute + zgemm per step) on host cores.

    The apply's cost does not depend on the operand values, so the sample uses
    random environments of the exact shapes (D x 5 x D) instead of building
    them from a canonical state."""
    from oracle import heff_oracle as ref
    import __graft_entry__
    __graft_entry__.load_package()
    from graft_amd import operators as O, planning as P, trees as Tr
    topo = Tr.binary_topology(depth)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    rng = np.random.default_rng(seed)
    W = O.heisenberg_host(topo, 1.0)
    n, m = 1, 0
    envs = {(3, 1): ref.random_complex(rng, (D_cpu, 5, D_cpu)),
            (4, 1): ref.random_complex(rng, (D_cpu, 5, D_cpu)),
            (2, 0): ref.random_complex(rng, (D_cpu, 5, D_cpu))}
    shape = (D_cpu, D_cpu, 2, D_cpu, 2, 1)
    theta = ref.random_complex(rng, shape)
    labels, statics = ref.h2_network(t, W, envs, n, m, shape)
    spec = P.ContractionSpec(labels, [False] * len(labels), len(shape), (len(shape), 0), 0,
                             [3, 4, 1, 5, 2, 6])
    dims = [list(shape)] + [list(s.shape) for s in statics]
    plan = P.plan_contraction(spec, dims, fused_permutes=True)
    steps = [(s.a, s.b, s.dst, s.oindA, s.cindA, s.oindB, s.cindB, s.conjA, s.conjB,
              tuple(s.out[0]) + tuple(s.out[1])) for s in plan.steps]
    ops = [theta] + statics
    ref.execute_pair_steps(steps, ops, plan.nslots, plan.output_slot)  # warm-up
    times = []
    t0 = time.perf_counter()
    while True:
        a = time.perf_counter()
        ref.execute_pair_steps(steps, ops, plan.nslots, plan.output_slot)
        times.append(time.perf_counter() - a)
        if time.perf_counter() - t0 >= budget_s or len(times) >= 5:
            break
    sec = float(np.median(times))
    return dict(seconds_per_apply=sec, flops=8.0 * plan.flops, D=D_cpu, samples=len(times))


def cpu_gpu_sweep(maxdim=32, nsweeps=4, seed=1):
    """The reference's CPU-runnable workload, measured on both sides: dmrg2 on
    the transverse-field Ising model (J=1, h=0.9) on the star of a centre and
    3 arms of 2 sites, random bond-2 start, maxdim 32, ``nsweeps`` sweeps, on
    the host (the oracle's sweep: the same pairwise maps in numpy, Lanczos to
    1e-12, LAPACK SVD) and on the GPU (dmrg2_), against exact diagonalization."""
    import torch
    from oracle import heff_oracle as ref
    import __graft_entry__
    __graft_entry__.load_package()
    from graft_amd import backend as B, groundstate as GS, networks as N, operators as O, trees as Tr
    topo = Tr.star_topology(3, 2)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    host, _ = N.random_ttns_host(np.random.default_rng(seed), topo, 2, 2)
    W = O.tfi_host(topo, 1.0, 0.9)
    E0 = float(np.linalg.eigvalsh(ref.dense_hamiltonian(t.n, ref.tfi_terms(t, 1.0, 0.9)))[0])
    t0 = time.perf_counter()
    _, _, Ec = ref.dmrg2_sweeps(t, host, W, topo.root, maxdim, nsweeps=nsweeps)
    cpu_s = time.perf_counter() - t0
    psi = N.ttns_from_host(topo, host, topo.root)
    H = O.tfi_ttno(topo, 1.0, 0.9)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, Eg = GS.dmrg2_(psi, H, trunc=B.TruncationScheme(maxdim=maxdim), nsweeps=nsweeps, tol=0.0)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    return {"workload": f"dmrg2, TFI J=1 h=0.9 on star 3x2 (7 sites), bond-2 start, maxdim {maxdim}, "
                        f"{nsweeps} sweeps",
            "cpu_s": cpu_s, "gpu_s": gpu_s, "cpu_energy_err": abs(Ec[-1] - E0),
            "gpu_energy_err": abs(Eg[-1] - E0), "cpu_threads": blas_threads()}


def cpu_protocol(h2, x, D_cpu, D_one, flops_apply, samples=5):
    """The CPU baseline protocol: one GPU apply and one host apply of the same
    pairwise plan alternate (median of ``samples`` each), with all host cores at
    D_cpu and, separately, one core at D_one; host rates are FLOP-normalized to
    the GPU workload's apply."""
    import torch
    from threadpoolctl import threadpool_limits
    gpu_ms, cpu = [], {}
    y = h2(x)
    for label, D, threads in (("all_cores", D_cpu, None), ("one_core", D_one, 1)):
        times = []
        for i in range(samples):
            torch.cuda.synchronize()
            a = time.perf_counter()
            y = h2(x, out=y)
            torch.cuda.synchronize()
            gpu_ms.append((time.perf_counter() - a) * 1e3)
            if threads is None:
                c = cpu_baseline(D, budget_s=0.0)
            else:
                with threadpool_limits(limits=threads, user_api="blas"):
                    c = cpu_baseline(D, budget_s=0.0)
            times.append(c["seconds_per_apply"])
        sec = float(np.median(times))
        rate = c["flops"] / sec
        cpu[label] = {"D": D, "threads": threads or blas_threads(), "seconds_per_apply": sec,
                      "gflops": rate / 1e9, "applies_per_s_normalized": rate / flops_apply,
                      "samples": samples}
    cpu["gpu_ms_interleaved_median"] = float(np.median(gpu_ms))
    return cpu


def precision_lockstep(D, nsteps, dt=0.05, seed=3, arms=3, armlen=10,
                       precisions=("fp32", "bf16x3"), cbe=True):
    """The precision study over many steps: the FP64 run and each reduced-product
    run advance in lockstep from the same state, so the per-step energy drift
    and fidelity against FP64 need no stored trajectory (the Krylov-vector
    check of ``precision_bench`` is left out).  A progress line per step goes
    to stderr."""
    import torch
    import __graft_entry__
    __graft_entry__.load_package()
    from graft_amd import backend as B, evolution as EV, gradients as G, networks as N
    from graft_amd import operators as O, trees as Tr
    topo = Tr.star_topology(arms, armlen)
    psi0 = N.random_ttns(np.random.default_rng(seed), topo, 2, D, center=topo.root)
    H = O.tfi_ttno(topo, 1.0, 0.9)
    dz = -1j * dt
    precs = ("fp64",) + tuple(precisions)

    def make(prec):
        if cbe:
            return EV.TDVP1_CBE(order=2, precision=prec, trunc=B.TruncationScheme(maxdim=D),
                                cbe=EV.PredictorCBE(max_add=32))
        return EV.TDVP1(order=2, precision=prec)

    evs = {p: make(p) for p in precs}
    psis = {p: psi0.copy() for p in precs}
    E0 = G.energy(psi0, H)
    times = {p: [] for p in precs}
    drift = {p: 0.0 for p in precs}
    infid = {p: [] for p in precisions}
    ediff = {p: 0.0 for p in precisions}
    for s in range(nsteps):
        E = {}
        for p in precs:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            EV.step_(evs[p], psis[p], H, dz)
            torch.cuda.synchronize()
            times[p].append(time.perf_counter() - t0)
            E[p] = G.energy(psis[p], H)
            drift[p] = max(drift[p], abs(E[p] - E0))
        for p in precisions:
            infid[p].append(1.0 - abs(G.overlap(psis[p], psis["fp64"])))
            ediff[p] = max(ediff[p], abs(E[p] - E["fp64"]))
        print(f"[step {s + 1}/{nsteps}] " + " ".join(f"{p} {times[p][-1]:.2f}s" for p in precs)
              + " infidelity " + " ".join(f"{p} {infid[p][-1]:.1e}" for p in precisions),
              file=sys.stderr, flush=True)
    res = {}
    for p in precs:
        r = {"step_s_mean": float(np.mean(times[p])), "energy_drift_max": float(drift[p]),
             "bond_dims_final": max(psis[p].bond_dims())}
        if p != "fp64":
            r.update(infidelity_vs_fp64_max=float(max(infid[p])),
                     infidelity_vs_fp64_final=float(infid[p][-1]),
                     energy_diff_vs_fp64_max=float(ediff[p]))
        res[p] = r
    return {"D": D, "steps": nsteps, "dt": dt, "sites": len(topo.ids), "lockstep": True,
            "evolver": "TDVP1_CBE(PredictorCBE(max_add=32))" if cbe else "TDVP1", "results": res}


def cpu_edge_rate(n, m, D, depth=5, seed=7, reps=3):
    """The oracle runs the FLOP-optimal pairwise plan (explicit permute + numpy/OpenBLAS
    zgemm per step) of the eff_h2 map on edge (n, m) of the depth-``depth`` binary
    Heisenberg tree with every bond at D and random environments; returns
    (GFLOP/s on the host cores, seconds per apply, FLOP per apply)."""
    from oracle import heff_oracle as ref
    from graft_amd import operators as O, planning as P, trees as Tr
    topo = Tr.binary_topology(depth)
    t = ref.Tree(list(topo.ids), list(topo.parent), [list(c) for c in topo.children])
    rng = np.random.default_rng(seed)
    W = O.heisenberg_host(topo, 1.0)
    envs = {(c, n): ref.random_complex(rng, (D, 5, D)) for c in t.children[n]}
    envs.update({(c, m): ref.random_complex(rng, (D, 5, D)) for c in t.children[m] if c != n})
    if t.parent[m] >= 0:
        envs[(t.parent[m], m)] = ref.random_complex(rng, (D, 5, D))
    shape = tuple([D] * len(t.children[n]) + [2] + [D] * (len(t.children[m]) - 1) + [2]
                  + [D if t.parent[m] >= 0 else 1])
    theta = ref.random_complex(rng, shape)
    labels, statics = ref.h2_network(t, W, envs, n, m, shape)
    spec = P.ContractionSpec(labels, [False] * len(labels), len(shape), (len(shape), 0), 0)
    plan = P.plan_contraction(spec, [list(shape)] + [list(x.shape) for x in statics])
    steps = [(q.a, q.b, q.dst, q.oindA, q.cindA, q.oindB, q.cindB, q.conjA, q.conjB,
              tuple(q.out[0]) + tuple(q.out[1])) for q in plan.steps]
    ops = [theta] + statics
    ref.execute_pair_steps(steps, ops, plan.nslots, plan.output_slot)      # warm-up
    times = []
    for _ in range(reps):
        a = time.perf_counter()
        ref.execute_pair_steps(steps, ops, plan.nslots, plan.output_slot)
        times.append(time.perf_counter() - a)
    sec = float(np.median(times))
    return 8.0 * plan.flops / sec / 1e9, sec, 8.0 * plan.flops


def sweep_bench(D, precision, nsweeps=1, depth=5, seed=7, krylovdim=20, phase_times=False,
                cpu_estimate=False):
    """dmrg2_ sweeps on the depth-5 binary-tree Heisenberg model with maxdim D.

    ``cpu_estimate``: the FLOP-optimal plan's FLOPs of every apply the sweep ran,
    divided by the oracle's measured rate on the heaviest edge (a lower bound
    for the host's sweep time: it leaves out environments, SVDs and the lower
    rate of the smaller applies)."""
    import torch
    import __graft_entry__
    __graft_entry__.load_package()
    from graft_amd import backend as B, groundstate as GS, networks as N, operators as O, trees as Tr
    from graft_amd import planning as P
    topo = Tr.binary_topology(depth)
    rng = np.random.default_rng(seed)
    psi = N.random_ttns(rng, topo, 2, D)
    H = O.heisenberg_ttno(topo, 1.0)
    times, counts = [], {"fp64": 0, "lowp": 0}
    solves = []                        # (spec, operand dims, applies) of every local problem
    orig_backend = GS.groundstate_eigsolve_backend

    def counted(effective, x, howmany, which, precision="fp64", **kw):
        r = orig_backend(effective, x, howmany, which, precision, **kw)
        solves.append((effective.spec, [list(x.dims)] + [list(st.dims) for st in effective.statics],
                       r[2].numops + getattr(r[2], "lowp_ops", 0)))
        return r

    def cb(sweep, n, m, E, info):
        counts["fp64"] += info.numops
        counts["lowp"] += getattr(info, "lowp_ops", 0)

    energies, phases = [], []
    if cpu_estimate:
        GS.groundstate_eigsolve_backend = counted
    try:
        for s in range(nsweeps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tm = {} if phase_times else None
            _, Es = GS.dmrg2_(psi, H, trunc=B.TruncationScheme(maxdim=D), nsweeps=1,
                              krylovdim=krylovdim, precision=precision, callback=cb, timings=tm)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            energies.append(Es[-1])
            if tm is not None:
                tm["eigsolve_tflops_fp64_equiv"] = (tm.get("flops_fp64", 0) + tm.get("flops_lowp", 0)) \
                    / max(tm.get("eigsolve", 1e-9), 1e-9) / 1e12
                phases.append(tm)
    finally:
        GS.groundstate_eigsolve_backend = orig_backend
    out = dict(D=D, precision=precision, sweep_s=times, energies=energies, applies=counts,
               phases=phases)
    if cpu_estimate:
        flops, cache = 0.0, {}
        for spec, dims, ops in solves:        # after the timed sweeps: planning is host work
            key = (spec.key(), tuple(tuple(d) for d in dims))
            if key not in cache:
                cache[key] = 8.0 * P.plan_contraction(spec, dims).flops
            flops += cache[key] * ops
        rate, sec, f = cpu_edge_rate(3, 1, D, depth, seed)
        out["cpu_estimate"] = {
            "seconds_per_sweep_lower_bound": flops / nsweeps / (rate * 1e9),
            "apply_flops_per_sweep": flops / nsweeps, "cpu_gflops": rate,
            "threads": blas_threads(),
            "sample": f"oracle eff_h2 on edge (3, 1) at D={D}: {sec:.3f} s per apply "
                      f"({f / 1e9:.1f} GFLOP), median of 3"}
    return out


def hubbard_tree():
    """16-site binary tree (heap order: node i's parent is (i - 1) // 2)."""
    import __graft_entry__
    __graft_entry__.load_package()
    from graft_amd import trees as Tr
    return Tr.TreeTopology("0", [(str((i - 1) // 2), str(i)) for i in range(1, 16)])


def hubbard_bench(D, nsweeps=2, precision="fp64", sigma=1.6, seed=5, dense=True, krylovdim=20,
                  tune_tiles=False):
    """Spinful Hubbard (t=1, U=4, half filling) on the 16-site tree: dmrg2_ sweeps
    with U(1) x U(1) sector-batched maps against the plain dense FP64 maps on the
    dense embedding of the same initial state, plus one root-edge apply of each."""
    import torch
    topo = hubbard_tree()
    from graft_amd import backend as B, contractions as C, groundstate as GS, networks as N
    from graft_amd import operators as O, sectors as S
    from graft_amd.tensors import DTensor
    rng = np.random.default_rng(seed)
    psi, host = S.random_hubbard_ttns(rng, topo, D, sigma=sigma)
    Hs = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    out = {"D": D, "sigma": sigma, "bond_sectors": {}, "precision": precision, "tune_tiles": tune_tiles}
    for c in range(len(topo.ids)):
        if topo.parent[c] >= 0:
            leg = psi.tensors[c].blocks.legs[-1]
            out["bond_sectors"][topo.ids[c]] = [len(leg.sectors), leg.dim, max(leg.degs())]
    psi_d = None
    if dense:
        Hd = N.TTNO(topo, [DTensor.from_numpy(w, w.ndim - 2) for w in O.hubbard_host(topo, 1.0, 4.0)],
                    ishermitian=True)
        psi_d = N.TTNS(topo, [DTensor.from_numpy(h.dense(), len(h.legs) - 1) for h in host], psi.center)

    def one_apply(state, H, blocks):
        cache = C.EnvCache(topo)
        x = C.two_site_tensor(state, 1, 0)
        h2 = C.eff_h2(cache, state, H, 1, 0, out_structure=x.blocks if blocks else None)
        y = h2(x)
        torch.cuda.synchronize()
        reps = 5
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            y = h2(x, out=y)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        info = h2.native.step_info()
        return {"ms": ms, "flops": h2.flops, "tflops": h2.flops / ms / 1e9,
                "flops_reference_model": h2.reference_flops,
                "tflops_reference_model": h2.reference_flops / ms / 1e9,
                "theta_elements": int(x.numel), "max_blocks_per_step": max(s["blocks"] for s in info)}

    out["apply_sector"] = one_apply(psi, Hs, True)
    if dense:
        out["apply_dense"] = one_apply(psi_d, Hd, False)

    def sweeps(state, H, prec):
        times, energies, tm_all = [], [], []
        for _ in range(nsweeps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            tm = {}
            _, Es = GS.dmrg2_(state, H, trunc=B.TruncationScheme(maxdim=D), nsweeps=1,
                              krylovdim=krylovdim, precision=prec, timings=tm,
                              tune_tiles=tune_tiles and state.tensors[0].blocks is not None)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            energies.append(Es[-1])
            tm_all.append(tm)
        return times, energies, tm_all

    out["sector_sweep_s"], out["sector_energies"], out["sector_phases"] = sweeps(psi, Hs, precision)
    if dense:
        out["dense_sweep_s"], out["dense_energies"], out["dense_phases"] = sweeps(psi_d, Hd, "fp64")
        out["speedup"] = [d / s for d, s in zip(out["dense_sweep_s"], out["sector_sweep_s"])]
        out["energy_rel_diff"] = [abs(a - b) / abs(b) for a, b in zip(out["sector_energies"],
                                                                      out["dense_energies"])]
    return out


def hubbard_tdvp_bench(D, nsteps=2, dt=0.05, sigma=1.6, seed=5, dense=True):
    """Real-time TDVP1 (order 2) of the spinful Hubbard model (t=1, U=4, half
    filling) on the 16-site tree: U(1) x U(1) sector maps with per-sector link
    QRs against the plain dense FP64 maps and dense QRs on the dense embedding
    of the same initial state.  Seconds per time step, energy drift, norm."""
    import torch
    topo = hubbard_tree()
    from graft_amd import contractions as C, evolution as EV, networks as N, operators as O
    from graft_amd import sectors as S
    from graft_amd.tensors import DTensor
    rng = np.random.default_rng(seed)
    psi, host = S.random_hubbard_ttns(rng, topo, D, sigma=sigma)
    Hs = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    out = {"D": D, "dt": dt, "order": 2, "bond_dims": psi.bond_dims()}
    if dense:          # the same initial state, centre at the root, before the sector run moves it
        Hd = N.TTNO(topo, [DTensor.from_numpy(w, w.ndim - 2) for w in O.hubbard_host(topo, 1.0, 4.0)],
                    ishermitian=True)
        psi_d = N.TTNS(topo, [DTensor.from_numpy(h.dense(), len(h.legs) - 1) for h in host], psi.center)

    def run(state, H):
        E0 = C.expect(state, H, C.EnvCache(topo)).real
        ev = EV.TDVP1(order=2)
        times = []
        for _ in range(nsteps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            EV.step_(ev, state, H, -1j * dt)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        E1 = C.expect(state, H, C.EnvCache(topo)).real
        return {"step_s": times, "E0": E0, "E": E1, "energy_drift_rel": abs(E1 - E0) / abs(E0),
                "norm_error": abs(N.norm(state) - 1.0)}

    out["sector"] = run(psi, Hs)
    if dense:
        out["dense"] = run(psi_d, Hd)
        out["speedup"] = [d / s for d, s in zip(out["dense"]["step_s"], out["sector"]["step_s"])]
        out["energy_rel_diff"] = abs(out["sector"]["E"] - out["dense"]["E"]) / abs(out["dense"]["E"])
    return out


def precision_bench(D, nsteps, dt=0.05, seed=3, arms=3, armlen=10, precisions=("fp32", "bf16x3"),
                    cbe=True):
    """TDVP1_CBE (order 2, PredictorCBE(max_add=32), maxdim D) of the
    transverse-field Ising model (J=1, h=0.9) on a star tree (TDVP1 with
    ``cbe=False``), one run per product precision from the same initial state:
    per-step energy drift, fidelity against the FP64 run, and the error of
    every reduced-precision Krylov exponential against the FP64 exponential of
    the same map and vector."""
    import torch
    import __graft_entry__
    __graft_entry__.load_package()
    from graft_amd import evolution as EV, gradients as G, krylov as K, networks as N
    from graft_amd import operators as O, tensors as T, trees as Tr
    topo = Tr.star_topology(arms, armlen)
    rng = np.random.default_rng(seed)
    psi0 = N.random_ttns(rng, topo, 2, D, center=topo.root)
    H = O.tfi_ttno(topo, 1.0, 0.9)
    dz = -1j * dt
    runs = {}
    kry_err = {}
    check_s = [0.0]
    orig = EV.evolution_exponentiate_backend

    def checked(effective, z, x, **kw):
        y, info = K.exponentiate(effective, z, x, **kw)
        if effective.precision != "fp64":
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            y64, _ = K.exponentiate(effective.with_precision("fp64"), z, x, **kw)
            e = T.norm(T.axpby_(1.0, y, -1.0, y64.clone())) / max(T.norm(y64), 1e-300)
            kry_err.setdefault(effective.precision, []).append(e)
            check_s[0] += time.perf_counter() - t0
        return y, info

    EV.evolution_exponentiate_backend = checked
    try:
        for prec in ("fp64",) + tuple(precisions):
            psi = psi0.copy()
            if cbe:
                from graft_amd import backend as B
                ev = EV.TDVP1_CBE(order=2, precision=prec, trunc=B.TruncationScheme(maxdim=D),
                                  cbe=EV.PredictorCBE(max_add=32))
            else:
                ev = EV.TDVP1(order=2, precision=prec)
            states, energies, times = [], [G.energy(psi, H)], []
            for s in range(nsteps):
                torch.cuda.synchronize()
                check_s[0] = 0.0
                t0 = time.perf_counter()
                EV.step_(ev, psi, H, dz)
                torch.cuda.synchronize()
                times.append(time.perf_counter() - t0 - check_s[0])   # without the FP64 checks
                energies.append(G.energy(psi, H))
                states.append(psi.copy())
            runs[prec] = dict(states=states, energies=energies, times=times)
    finally:
        EV.evolution_exponentiate_backend = orig
    out = {"D": D, "steps": nsteps, "dt": dt, "sites": len(topo.ids), "evolver":
           "TDVP1_CBE(PredictorCBE(max_add=32))" if cbe else "TDVP1", "results": {}}
    ref_states = runs["fp64"]["states"]
    for prec, r in runs.items():
        E0 = r["energies"][0]
        res = {"step_s_mean": float(np.mean(r["times"])),
               "energy_drift_max": float(max(abs(e - E0) for e in r["energies"])),
               "energy_final": r["energies"][-1]}
        if prec != "fp64":
            fid = [abs(G.overlap(a, b)) for a, b in zip(r["states"], ref_states)]
            res["infidelity_vs_fp64_max"] = float(max(1 - f for f in fid))
            res["infidelity_vs_fp64_final"] = float(1 - fid[-1])
            res["energy_diff_vs_fp64_max"] = float(max(abs(a - b) for a, b in
                                                       zip(r["energies"], runs["fp64"]["energies"])))
            errs = kry_err.get(prec, [])
            res["krylov_rel_error_max"] = float(max(errs)) if errs else None
            res["krylov_rel_error_median"] = float(np.median(errs)) if errs else None
        res["bond_dims_final"] = max(r["states"][-1].bond_dims())
        out["results"][prec] = res
    return out


def hubbard_apply_bench(D, reps=20, sigma=1.6, seed=5, autotune=True, lanczos=True):
    """One root-edge U(1) x U(1) eff_h2 on the 16-site Hubbard tree: per-step
    and whole-apply timing of the sector-batched map (the profiled workload)."""
    import torch
    topo = hubbard_tree()
    from graft_amd import contractions as C, operators as O, sectors as S
    rng = np.random.default_rng(seed)
    psi, _ = S.random_hubbard_ttns(rng, topo, D, sigma=sigma)
    H = O.hubbard_sector_ttno(topo, 1.0, 4.0)
    cache = C.EnvCache(topo)
    x = C.two_site_tensor(psi, 1, 0)
    h2 = C.eff_h2(cache, psi, H, 1, 0, out_structure=x.blocks)
    y = h2(x)
    step_ms = h2.native.execute_timed(h2.operands(x), y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y = h2(x, out=y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    # the tile autotuner on this plan (every configuration and orientation per step)
    t_at, step_ms_tuned, ms_tuned = None, step_ms, ms
    if autotune:
        torch.cuda.synchronize()
        t_at = time.perf_counter()
        h2.autotune(x)
        torch.cuda.synchronize()
        t_at = time.perf_counter() - t_at
        step_ms_tuned = h2.native.execute_timed(h2.operands(x), y)
        e0.record()
        for _ in range(reps):
            y = h2(x, out=y)
        e1.record()
        torch.cuda.synchronize()
        ms_tuned = e0.elapsed_time(e1) / reps
    # two Lanczos cycles of 20 applies: device driver (expansions replayed as
    # HIP graphs) against the per-call driver, wall time per apply
    from graft_amd import krylov as K
    lanczos_runs = {}
    for device in ((True, False, True, False) if lanczos else ()):
        torch.cuda.synchronize()
        t = time.perf_counter()
        vals, _, info = K.eigsolve(h2, x, 1, "SR", krylovdim=20, tol=1e-14, maxiter=2, device=device)
        torch.cuda.synchronize()
        lanczos_runs["device_graph" if device else "per_call"] = {
            "ms_per_apply": (time.perf_counter() - t) * 1e3 / info.numops, "applies": info.numops,
            "ritz_value": float(vals[0])}
    tflops_ref = h2.reference_flops / ms_tuned / 1e9
    return {"D": D, "ms_per_apply": ms_tuned, "ms_per_apply_untuned": ms, "autotune_s": t_at,
            "flops": h2.flops, "tflops": h2.flops / ms_tuned / 1e9,
            "flops_reference_model": h2.reference_flops, "tflops_reference_model": tflops_ref,
            "fp64_mfma_fraction_reference_model": tflops_ref / MI355X_FP64_MFMA_TFLOPS,
            "step_ms": step_ms_tuned, "step_ms_untuned": step_ms, "steps": h2.native.step_info(),
            "theta_elements": int(x.numel),
            "blocks": len(x.blocks.blocks) if hasattr(x.blocks, "blocks") else None,
            "lanczos": lanczos_runs}


def su2_apply_bench(tphys=2, maxdeg=24, maxspin=10, reps=20, dense=True, seed=0):
    """Root-edge eff_h2 of the spin-(tphys/2) Heisenberg model on the 16-site
    tree with SU(2) blocks (bond multiplet degeneracies capped at ``maxdeg``,
    spins at ``maxspin`` / 2), against the plain dense FP64 map of the same
    tensors' dense embeddings."""
    import torch
    topo = hubbard_tree()
    from graft_amd import contractions as C, networks as N, su2, tensors as T
    rng = np.random.default_rng(seed)
    psi = su2.random_ttns(rng, topo, tphys, maxdeg, maxspin)
    H = su2.heisenberg_ttno(topo, 1.0, tphys)
    cache = C.EnvCache(topo)
    x = C.two_site_tensor(psi, 1, 0)
    for w, v in ((3, 1), (4, 1), (2, 0)):
        C.env_(cache, psi, H, w, v)
    torch.cuda.synchronize()
    t = time.perf_counter()
    h2 = C.eff_h2(cache, psi, H, 1, 0, out_structure=x.blocks)
    t_plan = time.perf_counter() - t
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(m, v):
        y = m(v)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            y = m(v, out=y)
        e1.record()
        torch.cuda.synchronize()
        return y, e0.elapsed_time(e1) / reps

    y, ms = timed(h2, x)
    step_ms = h2.native.execute_timed(h2.operands(x), y)
    r = {"tphys": tphys, "maxdeg": maxdeg, "maxspin": maxspin,
         "bond_dims_dense": [psi.tensors[c].dims[-1] for c in range(1, 16)],
         "theta_elements_su2": int(x.numel), "theta_elements_dense": int(np.prod(x.dims)),
         "theta_blocks": len(x.blocks.blocks), "block_pairs": h2.native.npairs,
         "plan_s": t_plan, "ms_per_apply": ms, "step_ms": step_ms,
         "steps": [(s["m"], s["k"], s["n"], s["blocks"], s["config"]) for s in h2.native.step_info()]}
    if dense:
        psid = N.TTNS(topo, [T.DTensor.from_numpy(a.to_numpy(), a.ncod) for a in psi.tensors],
                      psi.center, psi.hasphys)
        Hd = N.TTNO(topo, [T.DTensor.from_numpy(w.to_numpy(), w.ncod) for w in H.tensors])
        cd = C.EnvCache(topo)
        xd = C.two_site_tensor(psid, 1, 0)
        h2d = C.eff_h2(cd, psid, Hd, 1, 0)
        yd, msd = timed(h2d, xd)
        ref = yd.to_numpy()
        r.update({"dense_ms_per_apply": msd, "dense_tflops": h2d.flops / msd / 1e9,
                  "speedup_vs_dense": msd / ms,
                  "rel_diff_vs_dense": float(np.linalg.norm(y.to_numpy() - ref) / np.linalg.norm(ref))})
    return r


def su2_sweep_bench(L=32, D=256, nsweeps=3, dense=True, seed=1):
    """dmrg2_ sweeps of the spin-1/2 Heisenberg chain of L sites (singlet
    ground state) at maxdim D: SU(2) blocks against plain dense FP64 maps,
    the same number of sweeps from random states; seconds per sweep."""
    import torch
    import __graft_entry__
    __graft_entry__.load_package()
    from graft_amd import backend as B, groundstate as GS, networks as N, operators as O, su2
    from graft_amd import trees as Tr
    topo = Tr.mps_topology(L)
    out = {"L": L, "D": D}
    for kind in (("su2", "dense") if dense else ("su2",)):
        rng = np.random.default_rng(seed)
        if kind == "su2":
            psi = su2.random_ttns(rng, topo, 1, 1, maxspin=4)
            H = su2.heisenberg_ttno(topo)
        else:
            psi = N.random_ttns(rng, topo, 2, 16, center=topo.root)
            H = O.heisenberg_ttno(topo)
            H.ishermitian = True
        secs, es = [], []
        for _ in range(nsweeps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            _, Es = GS.dmrg2_(psi, H, trunc=B.TruncationScheme(maxdim=D), nsweeps=1)
            torch.cuda.synchronize()
            secs.append(time.perf_counter() - t)
            es.append(Es[-1])
        out[kind] = {"sweep_s": secs, "energies": es,
                     "max_bond": max(psi.tensors[c].dims[-1] for c in range(L) if topo.parent[c] >= 0)}
    if dense:
        out["speedup_last_sweep"] = out["dense"]["sweep_s"][-1] / out["su2"]["sweep_s"][-1]
        out["energy_diff"] = abs(out["dense"]["energies"][-1] - out["su2"]["energies"][-1])
    return out


def su2_tdvp_bench(L=32, D=128, nsteps=4, dt=0.05, dense=True, seed=3):
    """Real-time TDVP1_CBE (order 2) of the spin-1/2 Heisenberg chain of L sites
    from a random singlet state with small bonds (spins up to 2, degeneracy 1),
    bond cap D: SU(2) blocks against the plain dense FP64 maps on the state's
    dense embedding, same PredictorCBE settings; seconds per step."""
    import torch
    import __graft_entry__
    __graft_entry__.load_package()
    from graft_amd import backend as B, contractions as C, evolution as EV, networks as N, su2
    from graft_amd import tensors as T, trees as Tr
    topo = Tr.mps_topology(L)
    psi = su2.random_ttns(np.random.default_rng(seed), topo, 1, 1, maxspin=4)
    H = su2.heisenberg_ttno(topo)
    H.ishermitian = True
    runs = [("su2", psi, H)]
    if dense:
        runs.append(("dense", N.TTNS(topo, [T.DTensor.from_numpy(a.to_numpy(), a.ncod) for a in psi.tensors],
                                     psi.center, psi.hasphys),
                     N.TTNO(topo, [T.DTensor.from_numpy(w.to_numpy(), w.ncod) for w in H.tensors],
                            ishermitian=True)))
    out = {"L": L, "D": D, "dt": dt}
    for kind, p, h in runs:
        ev = EV.TDVP1_CBE(cbe=EV.PredictorCBE(max_add=16, neigs=16, krylovdim=16),
                          trunc=B.TruncationScheme(maxdim=D))
        e0 = C.expect(p, h)
        secs = []
        for _ in range(nsteps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            EV.step_(ev, p, h, -1j * dt)
            torch.cuda.synchronize()
            secs.append(time.perf_counter() - t)
        out[kind] = {"step_s": secs, "energy_drift": abs(C.expect(p, h) - e0),
                     "max_bond": max(p.tensors[c].dims[-1] for c in range(L) if topo.parent[c] >= 0)}
    if dense:
        out["speedup_last_step"] = out["dense"]["step_s"][-1] / out["su2"]["step_s"][-1]
    return out


def blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max((i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"),
                   default=os.cpu_count())
    except Exception:
        return os.cpu_count()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--D", type=int, default=256)
    ap.add_argument("--cpu-D", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-one-D", type=int, default=64,
                    help="bond dimension of the one-core host apply")
    ap.add_argument("--no-su2-apply", action="store_true",
                    help="apply mode: skip the SU(2) block apply figure")
    ap.add_argument("--su2-maxdeg", type=int, default=24)
    ap.add_argument("--no-sector-apply", action="store_true",
                    help="apply mode: skip the D=1024 sector-batched apply figure")
    ap.add_argument("--no-sweep", action="store_true",
                    help="apply mode: skip the sweep figures (Hubbard sector vs dense, host vs GPU)")
    ap.add_argument("--sweep-hub-D", type=int, default=128,
                    help="apply mode: maxdim of the Hubbard sector vs dense FP64 sweep")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the heuristic tile configurations (default: autotune every step)")
    ap.add_argument("--mode", choices=["apply", "sweep", "hubbard", "hubbard-apply", "hubbard-tdvp",
                                       "precision", "su2-apply", "su2-sweep", "su2-tdvp"],
                    default="apply")
    ap.add_argument("--prec-D", type=int, default=64)
    ap.add_argument("--prec-steps", type=int, default=10)
    ap.add_argument("--no-cbe", action="store_true", help="precision mode: TDVP1 without expansion")
    ap.add_argument("--prec-lockstep", action="store_true",
                    help="precision mode: advance every precision in lockstep (long runs)")
    ap.add_argument("--prec-lowp", default="fp32,bf16x3",
                    help="precision mode: reduced precisions run beside FP64 (comma list, may be empty)")
    ap.add_argument("--hub-D", type=int, default=64)
    ap.add_argument("--no-lanczos", action="store_true",
                    help="hubbard-apply mode: skip the Lanczos cycles (short profiled runs)")
    ap.add_argument("--no-dense", action="store_true", help="hubbard mode: sector form only")
    ap.add_argument("--sweep-D", type=int, default=32)
    ap.add_argument("--sweeps", type=int, default=2)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--apply-precision", default="fp64", choices=["fp64", "fp32", "bf16x3"],
                    help="apply mode: product arithmetic of the map (reduced precision is not the "
                         "metric's FP64 form and is reported separately)")
    ap.add_argument("--tune-tiles", action="store_true",
                    help="hubbard mode: autotune the tiles of heavy sector maps once per step shape "
                         "class inside the sweeps")
    ap.add_argument("--phase-times", action="store_true",
                    help="sweep mode: synchronized wall time per dmrg2_ phase")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one plain 1 GiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    torch.cuda.set_device(0)
    if args.mode == "sweep":
        r = sweep_bench(args.sweep_D, args.precision, args.sweeps, phase_times=args.phase_times,
                        cpu_estimate=not args.no_cpu)
        print(json.dumps({"metric": "dmrg2 sweep seconds", "value": r["sweep_s"][-1], "unit": "s",
                          "higher_is_better": False, "n_gpus": 1, "config": {
                              "workload": f"dmrg2_ sweep, depth-5 binary tree Heisenberg, maxdim {r['D']}",
                              "precision": r["precision"]}, "detail": r}), flush=True)
        return
    if args.mode == "hubbard-apply":
        r = hubbard_apply_bench(args.hub_D, autotune=not args.no_autotune, lanczos=not args.no_lanczos)
        print(json.dumps({"metric": "sector-batched eff_h2 ms per apply", "value": r["ms_per_apply"],
                          "unit": "ms", "higher_is_better": False, "n_gpus": 1,
                          "config": {"workload": f"eff_h2 root edge, Hubbard 16-site tree, U(1)xU(1), "
                                                 f"D={r['D']}"}, "detail": r}), flush=True)
        return
    if args.mode == "su2-apply":
        r = su2_apply_bench(maxdeg=args.su2_maxdeg, dense=not args.no_dense)
        print(json.dumps({"metric": "SU(2)-block eff_h2 ms per apply", "value": r["ms_per_apply"],
                          "unit": "ms", "higher_is_better": False, "n_gpus": 1,
                          "config": {"workload": f"eff_h2 root edge, spin-1 Heisenberg 16-site tree, "
                                                 f"SU(2), multiplet degeneracy <= {r['maxdeg']}"},
                          "detail": r}), flush=True)
        return
    if args.mode == "su2-tdvp":
        r = su2_tdvp_bench(D=args.sweep_D, nsteps=args.prec_steps, dense=not args.no_dense)
        print(json.dumps({"metric": "TDVP1_CBE seconds per time step, SU(2) blocks vs plain dense FP64",
                          "value": r["su2"]["step_s"][-1], "unit": "s", "higher_is_better": False,
                          "n_gpus": 1, "config": {"workload": f"TDVP1_CBE, spin-1/2 Heisenberg chain, "
                                                              f"L={r['L']}, maxdim {r['D']}"},
                          "detail": r}), flush=True)
        return
    if args.mode == "su2-sweep":
        r = su2_sweep_bench(D=args.sweep_D, nsweeps=args.sweeps, dense=not args.no_dense)
        print(json.dumps({"metric": "dmrg2 seconds per sweep, SU(2) blocks vs plain dense FP64",
                          "value": r["su2"]["sweep_s"][-1], "unit": "s", "higher_is_better": False,
                          "n_gpus": 1, "config": {"workload": f"dmrg2_ sweeps, spin-1/2 Heisenberg chain, "
                                                              f"L={r['L']}, maxdim {r['D']}"},
                          "detail": r}), flush=True)
        return
    if args.mode == "hubbard-tdvp":
        r = hubbard_tdvp_bench(args.hub_D, args.prec_steps, dense=not args.no_dense)
        print(json.dumps({"metric": "TDVP1 seconds per time step, sector-batched vs plain dense FP64",
                          "value": r["sector"]["step_s"][-1], "unit": "s", "higher_is_better": False,
                          "n_gpus": 1, "config": {
                              "workload": f"TDVP1 order 2, Hubbard t=1 U=4 half filling, 16-site tree, "
                                          f"U(1)xU(1), D={r['D']}, dt={r['dt']}"},
                          "detail": r}), flush=True)
        return
    if args.mode == "precision":
        lowp = tuple(p for p in args.prec_lowp.split(",") if p)
        if args.prec_lockstep:
            r = precision_lockstep(args.prec_D, args.prec_steps, precisions=lowp, cbe=not args.no_cbe)
        else:
            r = precision_bench(args.prec_D, args.prec_steps, precisions=lowp, cbe=not args.no_cbe)
        print(json.dumps({"metric": "TDVP1 reduced-precision products: drift, fidelity, Krylov error",
                          "value": r["results"].get("bf16x3", {}).get("infidelity_vs_fp64_max"),
                          "unit": "1 - |<psi_lowp|psi_fp64>|", "higher_is_better": False, "n_gpus": 1,
                          "config": {"workload": f"{r['evolver']} order 2, TFI J=1 h=0.9, star 3x10 "
                                                 f"({r['sites']} sites), D={r['D']}, dt={r['dt']}"},
                          "detail": r}), flush=True)
        return
    if args.mode == "hubbard":
        r = hubbard_bench(args.hub_D, args.sweeps, args.precision, dense=not args.no_dense,
                          tune_tiles=args.tune_tiles)
        print(json.dumps({"metric": "dmrg2 sweep seconds, sector-batched vs plain dense FP64",
                          "value": r["sector_sweep_s"][-1], "unit": "s", "higher_is_better": False,
                          "n_gpus": 1, "config": {
                              "workload": f"dmrg2_ sweep, Hubbard t=1 U=4 half filling, 16-site tree, "
                                          f"U(1)xU(1), maxdim {r['D']}", "precision": r["precision"]},
                          "detail": r}), flush=True)
        return
    t_setup = time.perf_counter()
    g, topo, psi, H, cache, h2, x = build_workload(args.D)
    if args.apply_precision != "fp64":
        h2 = h2.with_precision(args.apply_precision)
    from graft_amd import _native as nat
    prec_code = nat.PRECISIONS[args.apply_precision]
    native = h2.native
    if not args.no_autotune:
        t_at = time.perf_counter()
        h2.autotune(x)
        print(f"# autotune {time.perf_counter() - t_at:.1f} s", file=sys.stderr, flush=True)
    print(f"# setup {time.perf_counter() - t_setup:.1f} s; plan {native.plan.strategy}, "
          f"{len(native.plan.steps)} steps, workspace {native.workspace_bytes / 2**30:.1f} GiB",
          file=sys.stderr, flush=True)
    for i, s in enumerate(native.step_info()):
        print(f"#   step {i}: m={s['m']} k={s['k']} n={s['n']} MACs={s['macs']:.3e} tile={s['config']} "
              f"orient={s['orientation']} ksplit={s['ksplit']}", file=sys.stderr, flush=True)

    y = None
    for _ in range(args.warmup):
        y = h2(x, out=y)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y = h2(x, out=y)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = e0.elapsed_time(e1) / args.steps
    applies_per_s = 1000.0 / ms_per_step
    flops_apply = h2.flops
    # dominant kernel: per-step HIP-event timing on the library's stream
    ops = h2.operands(x)
    reps = max(1, min(3, args.steps))
    step_ms = np.zeros(len(native.plan.steps))
    for _ in range(reps):
        step_ms += np.array(native.execute_timed(ops, y, prec_code))
    step_ms /= reps
    info = native.step_info()
    dom = int(np.argmax([s["macs"] for s in info]))
    dom_flops = 8.0 * info[dom]["macs"]
    achieved = dom_flops / (step_ms[dom] * 1e-3) / 1e12
    apply_tflops = flops_apply / (ms_per_step * 1e-3) / 1e12
    print(f"# {ms_per_step:.2f} ms/apply, {apply_tflops:.2f} TFLOP/s whole apply; dominant step {dom}: "
          f"{step_ms[dom]:.2f} ms, {achieved:.2f} TFLOP/s; per-step ms {np.round(step_ms, 3).tolist()}",
          file=sys.stderr, flush=True)

    if args.calib_copy:
        src = torch.empty(2 ** 30 // 16, dtype=torch.complex128, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cpu = None
    if not args.no_cpu:
        t_cpu = time.perf_counter()
        proto = cpu_protocol(h2, x, args.cpu_D, args.cpu_one_D, flops_apply)
        allc = proto["all_cores"]
        cpu = {"value": allc["applies_per_s_normalized"], "unit": "applies/s",
               "cores": allc["threads"], "kind": "port",
               "sample": (f"oracle (numpy/OpenBLAS zgemm, same pairwise plan) eff_h2 at the root edge "
                          f"with D={allc['D']}: {allc['seconds_per_apply']:.3f} s/apply "
                          f"({allc['gflops']:.1f} GFLOP/s), FLOP-normalized to the D={args.D} apply; "
                          f"median of {allc['samples']}, alternating with GPU applies; one core at "
                          f"D={proto['one_core']['D']}: {proto['one_core']['gflops']:.1f} GFLOP/s; "
                          f"host {platform.processor() or platform.machine()}, {os.cpu_count()} CPUs"),
               "protocol": proto}
        if not args.no_sweep:
            cpu["sweep"] = cpu_gpu_sweep()
        print(f"# cpu baseline {time.perf_counter() - t_cpu:.1f} s", file=sys.stderr, flush=True)
    out = {
        "metric": "H_eff applies/s + % MFMA peak; DMRG sweep s at fixed D vs plain FP64 GPU & CPU",
        "value": applies_per_s,
        "unit": "applies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (fp64)" if args.apply_precision == "fp64" else
                 f"c128 storage, {args.apply_precision} products",
        "data": "synthetic",
        "config": {"workload": f"eff_h2 apply, root edge of a depth-5 binary tree (63 sites), "
                               f"Heisenberg TTNO w=5, D={args.D}, complex FP64",
                   "theta_elements": int(np.prod(x.dims)),
                   "flops_per_apply": flops_apply,
                   "plan": native.plan.strategy,
                   "whole_apply_tflops": apply_tflops,
                   "mfma_fraction_whole_apply": apply_tflops / MI355X_FP64_MFMA_TFLOPS},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MI355X_FP64_MFMA_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / MI355X_FP64_MFMA_TFLOPS,
                     # achieved/frac count the algorithmic 8 FLOP per complex MAC; the 3M
                     # tile configurations issue 6 real MFMA FLOP per complex MAC
                     "complex_form": "3M" if info[dom]["config"] in THREE_M_CONFIGS else "4M",
                     "executed_frac": (achieved / MI355X_FP64_MFMA_TFLOPS
                                       * (0.75 if info[dom]["config"] in THREE_M_CONFIGS else 1.0)),
                     "traffic": profiled_traffic(dom, args.D),
                     "algorithmic_bytes": 16.0 * (info[dom]["m"] * info[dom]["k"]
                                                  + info[dom]["k"] * info[dom]["n"]
                                                  + info[dom]["m"] * info[dom]["n"]),
                     "kernel": f"zgemm_f64 step {dom} (m={info[dom]['m']}, k={info[dom]['k']}, "
                               f"n={info[dom]['n']})"},
        "cpu_baseline": cpu,
    }
    sweep = None
    if not args.no_sweep:
        t_sw = time.perf_counter()
        hubbard_bench(8, 1)        # first use of the factorization libraries, outside the figure
        r = hubbard_bench(args.sweep_hub_D, 1)
        sweep = {"workload": f"one dmrg2_ sweep, Hubbard t=1 U=4 half filling, 16-site tree, "
                             f"maxdim {r['D']}",
                 "sector_batched_fp64_s": r["sector_sweep_s"][-1],
                 "plain_dense_fp64_s": r["dense_sweep_s"][-1], "speedup": r["speedup"][-1],
                 "sector_energy": r["sector_energies"][-1],
                 "sector_phases": r["sector_phases"][-1],
                 "apply_ms_sector": r["apply_sector"]["ms"], "apply_ms_dense": r["apply_dense"]["ms"]}
        print(f"# sweeps {time.perf_counter() - t_sw:.1f} s", file=sys.stderr, flush=True)

    out["sweep"] = sweep
    sector = None
    if not args.no_sector_apply:
        # the sector-batched apply at the batching configuration's size (D=1024,
        # U(1) x U(1)): heuristic tiles, then the library's autotuner;
        # reference-model FLOPs
        t_sa = time.perf_counter()
        r = hubbard_apply_bench(1024, autotune=True, lanczos=False)
        sector = {"workload": "eff_h2 root edge, Hubbard 16-site tree, U(1)xU(1), D=1024",
                  "ms_per_apply": r["ms_per_apply"], "ms_per_apply_heuristic_tiles": r["ms_per_apply_untuned"],
                  "autotune_s": r["autotune_s"], "step_ms": r["step_ms"],
                  "flops_reference_model": r["flops_reference_model"],
                  "tflops_reference_model": r["tflops_reference_model"],
                  "fp64_mfma_fraction_reference_model": r["fp64_mfma_fraction_reference_model"],
                  "flops_stored_blocks": r["flops"]}
        print(f"# sector apply {time.perf_counter() - t_sa:.1f} s", file=sys.stderr, flush=True)
    out["sector_apply"] = sector
    su2r = None
    if not args.no_su2_apply:
        # SU(2) blocks: coupled block pairs (recoupling coefficients) per step
        t_su = time.perf_counter()
        r = su2_apply_bench(maxdeg=args.su2_maxdeg)
        su2r = {"workload": "eff_h2 root edge, spin-1 Heisenberg 16-site tree, SU(2), "
                            f"multiplet degeneracy <= {r['maxdeg']}",
                "ms_per_apply": r["ms_per_apply"], "dense_ms_per_apply": r["dense_ms_per_apply"],
                "speedup_vs_dense": r["speedup_vs_dense"], "rel_diff_vs_dense": r["rel_diff_vs_dense"],
                "theta_elements_su2": r["theta_elements_su2"],
                "theta_elements_dense": r["theta_elements_dense"], "block_pairs": r["block_pairs"],
                "step_ms": r["step_ms"]}
        print(f"# su2 apply {time.perf_counter() - t_su:.1f} s", file=sys.stderr, flush=True)
    out["su2_apply"] = su2r
    su2s = None
    if not args.no_su2_apply:
        t_su = time.perf_counter()
        r = su2_sweep_bench(D=256, nsweeps=3)
        su2s = {"workload": "dmrg2_ sweeps, spin-1/2 Heisenberg chain, L=32, maxdim 256, 3 sweeps each",
                "su2_sweep_s": r["su2"]["sweep_s"], "plain_dense_fp64_sweep_s": r["dense"]["sweep_s"],
                "speedup_last_sweep": r["speedup_last_sweep"], "energy_su2": r["su2"]["energies"][-1],
                "energy_dense": r["dense"]["energies"][-1], "energy_diff": r["energy_diff"]}
        print(f"# su2 sweeps {time.perf_counter() - t_su:.1f} s", file=sys.stderr, flush=True)
    out["su2_sweep"] = su2s
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
