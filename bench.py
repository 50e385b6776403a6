"""Benchmark: second-order one-site TDVP sweeps/s at chi = 256 on one MI355X.

Workload (the configuration the headline metric is quoted on): ternary tree of
40 nodes (root with three children, each with three children, each with three
leaves), one spin-1/2 per node, random-coupling XYZ Hamiltonian
H = -sum_e (Jx XX + Jy YY + Jz ZZ) with J ~ U[-1, 1] per edge and axis
(seed 1), random complex TTNS (seed 0) at bond dim min(256, 2^|subtree|, ...),
QR(KEEP)-canonicalised and normalised; dt = 0.05, Krylov dim <= 10,
krylov_tol = 1e-10.  One step = forward + backward sweep.

Runs from the same state the FP64 path (complex128, FP64 MFMA), the relaxed
path (FP64 except the Krylov applications of H_eff whose error can change the
local result by at most RELAX_TOL relative to its input; those run as 3M FP32
MFMA on complex64 copies of the environments) and the mixed path (complex64
environments, 3M FP32 MFMA; Krylov/QR/SVD FP64).  Each reduced-precision path
reports its errors against FP64 at the same time (energy, <Z_i> of every site,
infidelity).  ``value`` is the sweeps/s of the fastest path within 1e-8 of the
FP64 run on all three (FP64 meets the bar by definition).  The relaxed path
with the bf16 level is run and reported beside them.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--chi 256] [--no-cpu]
                    [--c4-instances N] [--c4-group G] [--c4-single S] [--launch-log PATH]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "bf16": 2.5e15}
# measured HBM bytes per LARGE launch of each (backend, precision): rocprofv3
# --pmc FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (scales from --calibrate-copy) per
# dispatch of the relaxed path, averaged with the timed steps' shape mix (the
# per-shape table is under profiles/); None where not measured
TRAFFIC_BYTES_PER_LAUNCH = {
    # relaxed path's timed mix per step: 12 x (65536, 512, 512) accumulating at
    # 2.66 GB, 14 x (65536, 512, 512) at 2.13 GB, 16 x (131072, 256, 256) at 1.29 GB:
    # 1.96 GB per launch, 1.59x the 1.23 GB of operands read once, C written once
    # (and read once where the launch accumulates)
    ("gather", "fp64"): 1.96e9, ("blas", "fp64"): None,
    # 28 x (65536, 512, 512) accumulating at 1.63 GB and 14 non-accumulating at 1.27 GB
    ("gather", "fp32"): 1.51e9, ("blas", "fp32"): None,
    ("gather", "bf16"): None}
KERNEL_NAME = {("gather", "fp64"): "gemm_direct_f64_kernel<LARGE=1> (ptn_gemm_gather, FP64 3M)",
               ("blas", "fp64"): "rocblas_zgemm (ptn_gemm_strided)",
               ("gather", "fp32"): "gemm_gather_kernel<F32_3M, ..., LARGE=1> (ptn_gemm_gather)",
               ("blas", "fp32"): "rocblas_cgemm (ptn_gemm_strided)",
               ("gather", "bf16"): "gemm_bf16_kernel<..., LARGE=1> (ptn_gemm_gather, BF16 3M)"}
# real flops executed per complex multiply-add (8 = four-product form)
EXECUTED_PER_CMAC = {("gather", "fp64"): 6.0, ("blas", "fp64"): 8.0,
                     ("gather", "fp32"): 6.0, ("blas", "fp32"): 8.0, ("gather", "bf16"): 6.0}
METRIC = "TDVP sweeps/s & H_eff TFLOP/s at χ=256, mixed vs FP64 GPU at ΔE≤1e-8"
# relax_tol of the relaxed path: bound on the change of each local Krylov result
# (relative to its input's norm) from the applications run in 3M FP32
RELAX_TOL = 1e-7


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(chi: int, device):
    import pytreenet_amd as ptn
    from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
    ids, parent, children = complete_tree(3, 3)
    tensors = random_state_tensors(ids, parent, children, chi, seed=0)
    st = build_state(ids, parent, children, tensors, device=device)
    H = ptn.random_xyz_model(st, seed=1)
    return st, H, (ids, parent, children)


def run_tdvp(st, H, precision: str, warmup: int, steps: int, profile: bool,
             relax_tol: float = RELAX_TOL, relax_bf16: bool = False, relax_low: str = "fp32"):
    import torch
    import pytreenet_amd as ptn
    from pytreenet_amd import engine as E
    from pytreenet_amd import krylov as K
    from pytreenet_amd.contractions import multi_single_site_expectation_value
    from pytreenet_amd.ttno import PAULI_Z
    cfg = ptn.TDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, precision=precision,
                         relax_tol=relax_tol, relax_bf16=relax_bf16, relax_low=relax_low)
    tag = precision + ("-bf16" if relax_bf16 else "")
    E.PROFILER.phase = f"{tag}-init"
    t0 = time.time()
    alg = ptn.SecondOrderOneSiteTDVP(st, H, 0.05, 0.05 * (warmup + steps), [], config=cfg,
                                     solver_options={"max_krylov_dim": 10,
                                                     "krylov_tol": 1e-10})
    c = alg.state.orthogonality_center_id
    alg.state.tensors[c] = alg.state.tensors[c] / alg.state.norm()
    torch.cuda.synchronize()
    log(f"[{tag}] init {time.time() - t0:.2f}s")
    E.PROFILER.phase = f"{tag}-warmup"
    for i in range(warmup):
        t1 = time.time()
        alg.run_one_time_step()
        torch.cuda.synchronize()
        log(f"[{tag}] warmup step {i} {time.time() - t1:.2f}s")
    nlog = len(alg.apply_log)
    fl0, n0 = E.PROFILER.flops, E.PROFILER.launches
    low0, app0, lower0 = K.STATS.low_applies, K.STATS.applies, K.STATS.lower_applies
    E.PROFILER.phase = f"{tag}-timed"
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(steps):
        alg.run_one_time_step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    log(f"[{tag}] {steps} timed steps {elapsed:.3f}s")
    fl1, n1 = E.PROFILER.flops, E.PROFILER.launches
    low, apps = K.STATS.low_applies - low0, K.STATS.applies - app0
    E.PROFILER.phase = f"{tag}-eval"
    energy = ptn.expectation_value(alg.state, H).real
    z = multi_single_site_expectation_value(alg.state, {n: PAULI_Z for n in alg.state.nodes})
    return {"elapsed": elapsed, "energy": energy, "alg": alg,
            "z": {n: v[0].real for n, v in z.items()},
            "applies": alg.apply_log[nlog:],
            "gemm_flops": fl1 - fl0, "launches": n1 - n0,
            "low_fraction": low / max(apps, 1),
            "lower_fraction": (K.STATS.lower_applies - lower0) / max(apps, 1),
            "kernels": E.PROFILER.summary(phase=f"{tag}-timed") if profile else []}


def cpu_baseline(chi_sample: int, steps_applies, shapes_full, wshapes_full):
    """Reference-order (oracle) contraction rate on a bounded sample, scaled to chi=256.

    The reference contraction order needs a 67 GB intermediate per H_eff apply at
    the chi=256 ternary root, so the whole step cannot be run on the host.  The
    oracle's reference-order H_eff applies at the root and an L1 node, and its
    environment updates, are timed at chi = ``chi_sample``; their exact
    reference-order flop counts give the host's contraction rate, which is
    divided by the reference-order flops of one chi=256 step (Krylov applies per
    node taken from the GPU run).
    """
    import numpy as np
    from oracle import contract as OC
    from oracle.flops import bond_flops, env_flops, heff_flops
    from pytreenet_amd.random_states import bond_dims, complete_tree, random_state_tensors
    from pytreenet_amd.ttno import build_edge_ttno_tensors, pauli_matrices
    from tests.helpers import topology_from_spec
    ids, parent, children = complete_tree(3, 3)
    topo = topology_from_spec(ids, parent, children)
    tensors = random_state_tensors(ids, parent, children, chi_sample, seed=0)
    X, Y, Z = pauli_matrices()
    rng = np.random.default_rng(1)
    edges = []
    for n in ids:
        if parent[n] is not None:
            j = rng.uniform(-1, 1, 3)
            edges += [(parent[n], n, -j[0], X, X), (parent[n], n, -j[1], Y, Y),
                      (parent[n], n, -j[2], Z, Z)]

    class _S:  # minimal structure object for the TTNO builder
        def __init__(self):
            from pytreenet_amd.random_states import build_state
            self.st = build_state(ids, parent, children, tensors, device="cpu")
    W = build_edge_ttno_tensors(_S().st, {}, edges)
    shapes = {n: t.shape for n, t in tensors.items()}
    wshapes = {n: w.shape for n, w in W.items()}
    t0 = time.perf_counter()
    envs = OC.build_envs_to_root(topo, tensors, W)
    t_env = time.perf_counter() - t0
    f_env = sum(env_flops(topo, n, parent[n], shapes, wshapes) for n in ids if parent[n])
    f_apply, t_apply = 0.0, 0.0
    for nid in ["site0"]:
        t0 = time.perf_counter()
        OC.heff_apply(topo, nid, tensors[nid], W[nid], envs)
        t_apply += time.perf_counter() - t0
        f_apply += heff_flops(topo, nid, shapes, wshapes)
    rate = (f_env + f_apply) / (t_env + t_apply)
    # reference-order flops of one chi=256 step
    f_step = 0.0
    for kind, where, m in steps_applies:
        if kind == "site":
            f_step += m * heff_flops(topo, where, shapes_full, wshapes_full)
        elif kind == "bond":
            a, b = where
            d = shapes_full[a][topo.neighbour_index(a, b)]
            f_step += m * bond_flops(d, wshapes_full[a][topo.neighbour_index(a, b)])
    # environment updates: one per link update and per orthogonality move (two per edge)
    for n in ids:
        if parent[n]:
            f_step += env_flops(topo, n, parent[n], shapes_full, wshapes_full)
            f_step += env_flops(topo, parent[n], n, shapes_full, wshapes_full)
    return rate, f_step, t_env + t_apply


def c4_instances(count: int, seed0: int = 1000):
    """Instances of the batching configuration: 31-node binary tree, random-field
    Heisenberg (h_i ~ U[-2, 2], seed per instance), second-order two-site TDVP with
    truncation at 1e-10 and chi_max = 128, dt = 0.05, from bond dimension 8."""
    import pytreenet_amd as ptn
    from pytreenet_amd.random_states import build_state, complete_tree, random_state_tensors
    from pytreenet_amd.ttno import random_field_heisenberg
    ids, parent, children = complete_tree(2, 4)
    algs = []
    for b in range(count):
        st = build_state(ids, parent, children,
                         random_state_tensors(ids, parent, children, 8, seed=seed0 + b),
                         device="cuda")
        H = random_field_heisenberg(st, 2.0, seed=seed0 + 100000 + b)
        cfg = ptn.TwoSiteTDVPConfig(time_evo_mode=ptn.TimeEvoMode.KRYLOV, max_bond_dim=128,
                                    rel_tol=1e-10, total_tol=1e-10)
        a = ptn.SecondOrderTwoSiteTDVP(st, H, 0.05, 0.1, [], config=cfg,
                                       solver_options={"max_krylov_dim": 10,
                                                       "krylov_tol": 1e-10})
        c = a.state.orthogonality_center_id
        a.state.tensors[c] = a.state.tensors[c] / a.state.norm()
        algs.append(a)
    return algs


def run_c4(count: int, group=None, single=None):
    """Instance-steps/s of the batching configuration on ``count`` instances: the
    first step (bonds grow from 8 to 128) and the second (at chi_max), batched in
    lockstep (in lockstep groups of ``group`` instances when given) and one
    instance at a time from the same states (on the first ``single`` of them when
    given: that rate does not depend on the number of instances)."""
    import torch
    import pytreenet_amd as ptn
    out = {"instances": count}
    if group:
        out["lockstep_group"] = group
    for mode in ("batched", "one_at_a_time"):
        n = count if mode == "batched" or not single else min(single, count)
        algs = c4_instances(n)
        torch.cuda.synchronize()
        for step in range(2):
            t0 = time.perf_counter()
            if mode == "batched":
                bt = ptn.BatchedTDVP(algs)
                bt.lockstep_group = group
                bt.progress = lambda done, kind: (
                    log(f"[c4 batched] {done} lockstep points, {time.perf_counter() - t0:.0f}s")
                    if done % 50 == 0 else None)
                bt.run(1)
            else:
                for a in algs:
                    a.run_one_time_step()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            key = "growing" if step == 0 else "at_chi_max"
            out[f"{mode}_{key}_instance_steps_per_s"] = n / dt
            log(f"[c4 {mode}] step {step}: {dt:.2f}s for {n} instances")
        del algs
        torch.cuda.empty_cache()
    if single and single < count:
        out["one_at_a_time_instances"] = min(single, count)
    out["speedup_at_chi_max"] = (out["batched_at_chi_max_instance_steps_per_s"] /
                                 out["one_at_a_time_at_chi_max_instance_steps_per_s"])
    return out


def calibrate_copy():
    """8 contiguous copies of 1 GiB (complex128) by the ptn_permute kernel.

    Each launch reads 1 GiB and writes 1 GiB; the FETCH_SIZE / WRITE_SIZE
    counters of these launches give the scale factor for the GEMM counters.
    """
    import torch
    from pytreenet_amd import engine as E
    n = (1 << 30) // 16
    src = torch.ones(n, dtype=torch.complex128, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(8):
        E.strided_copy([n], src, [1], dst, [1])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 8
    print(json.dumps({"calibration": "ptn_permute contiguous copy", "bytes_read": 1 << 30,
                      "bytes_written": 1 << 30, "ms_per_copy": dt * 1e3,
                      "GBps_read_plus_write": 2 * (1 << 30) / dt / 1e9}), flush=True)


def host_facts():
    # read from /proc (no child process: the bench leaves nothing running behind it)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = ""
    try:
        import numpy as np
        cfg = np.show_config(mode="dicts")
        blas = cfg["Build Dependencies"]["blas"]["name"]
    except Exception:
        pass
    return {"cpu_count": os.cpu_count(), "cpu_model": model, "blas": blas,
            "python": platform.python_version()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--chi", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-chi", type=int, default=64)
    ap.add_argument("--c4-instances", type=int, default=16,
                    help="instances of the batching configuration measured beside the "
                         "bench (0: skip)")
    ap.add_argument("--c4-group", type=int, default=None,
                    help="lockstep group size of the batching configuration's batched run")
    ap.add_argument("--c4-single", type=int, default=None,
                    help="instances of the batching configuration run one at a time "
                         "(default: all)")
    ap.add_argument("--relax-low", choices=("fp32", "bf16x6"), default="fp32",
                    help="arithmetic of the relaxed paths' FP32-accurate level: the FP32 "
                         "MFMA, or six products of three-piece bf16 splits on the bf16 MFMA")
    ap.add_argument("--autotune", action="store_true",
                    help="launch the non-LARGE GEMMs on autotuned configurations")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 8 ptn_permute copies of 1 GiB (HBM counter calibration)")
    ap.add_argument("--launch-log", default=None,
                    help="write every LARGE gather-GEMM launch of the process in issue order "
                         "(M, N, K, precision, beta_one, c_mmajor, phase) to this JSON file, "
                         "to match per-dispatch profiler counters to shapes")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "bench.py runs on exactly one GPU: the TDVP sweep of one "
                                   "state is sequential and this build measures one GPU "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    dev = torch.device("cuda:0")
    from pytreenet_amd import engine as E
    E.AUTOTUNE = args.autotune
    # HIP events around every LARGE gather-GEMM launch of the process, tagged by phase
    E.PROFILER.start("setup")
    st, H, _ = make_problem(args.chi, dev)
    r64 = run_tdvp(st, H, "fp64", args.warmup, args.steps, profile=True)
    sps64 = args.steps / r64["elapsed"]
    res = {"fp64_sweeps_per_s": sps64, "fp64_energy": r64["energy"],
           "fp64_gemm_tflops": r64["gemm_flops"] / r64["elapsed"] / 1e12,
           "fp64_launches_per_step": r64["launches"] / args.steps,
           "fp64_applies_per_step": sum(m for _, _, m in r64["applies"]) / args.steps}
    # (path, run, sweeps/s) of every path within the accuracy bar against FP64 after the
    # same steps from the same state: |dE| <= 1e-8, max_i |d<Z_i>| <= 1e-8 over all
    # sites, and infidelity 1 - |<psi_fp64|psi>| <= 1e-8 (both states normalised)
    ok = [("fp64", r64, sps64)]
    from pytreenet_amd.contractions import contract_two_ttns
    if not args.no_mixed:
        for prec, key, bf16 in (("relaxed", "relaxed", False),
                                ("relaxed", "relaxed_bf16", True), ("fp32", "mixed", False)):
            r = run_tdvp(st, H, prec, args.warmup, args.steps, profile=True, relax_bf16=bf16,
                         relax_low=args.relax_low)
            sps = args.steps / r["elapsed"]
            de = abs(r["energy"] - r64["energy"])
            dz = max(abs(r["z"][n] - r64["z"][n]) for n in r64["z"])
            infid = 1.0 - abs(contract_two_ttns(r["alg"].state, r64["alg"].state,
                                                conj_second=True))
            res.update({f"{key}_sweeps_per_s": sps, f"{key}_energy": r["energy"],
                        f"{key}_abs_energy_error": de, f"{key}_max_abs_z_error": dz,
                        f"{key}_infidelity": infid, f"{key}_speedup": sps / sps64,
                        f"{key}_gemm_tflops": r["gemm_flops"] / r["elapsed"] / 1e12,
                        f"{key}_low_precision_applies": r["low_fraction"]
                        if prec == "relaxed" else 1.0})
            if bf16:
                res[f"{key}_bf16_applies"] = r["lower_fraction"]
            if de <= 1e-8 and dz <= 1e-8 and infid <= 1e-8:
                ok.append((key if prec == "relaxed" else prec, r, sps))
            del r["alg"]
    path, run, value = max(ok, key=lambda t: t[2])
    # the roofline follows the path whose rate is `value`
    kern = run["kernels"]
    roofline = None
    if kern:
        # dominant kernel: of the LARGE launches (gemm_kernel_class "large") in the
        # timed steps, the backend with the larger total time -- rocBLAS strided GEMM
        # ("blas") or the hand-written gather-GEMM's LARGE instantiation ("gather");
        # the other one is reported beside it.  `achieved` counts 8 M N K flops per
        # launch (a complex multiply-add as 8 real flops); `executed_*` count what
        # the MFMAs run (6 M N K for the three-product form)
        by = {}
        for k in kern:
            by.setdefault((k["backend"], k["compute"]), []).append(k)
        tot = {b: sum(k["total_ms"] for k in ks) for b, ks in by.items()}
        order = sorted(by, key=lambda b: -tot[b])

        def block(bp):
            ks = by[bp]
            b, prec = bp
            n = sum(k["launches"] for k in ks)
            fl = sum(k["flops_per_launch"] * k["launches"] for k in ks)
            ms = sum(k["total_ms"] for k in ks)
            ach = fl / (ms * 1e-3)
            ex = ach * EXECUTED_PER_CMAC[bp] / 8.0
            return {"bound": "mfma", "achieved": ach / 1e12, "peak": PEAK[prec] / 1e12,
                    "unit": "TFLOP/s", "frac": ach / PEAK[prec],
                    "executed_tflops": ex / 1e12, "executed_frac": ex / PEAK[prec],
                    # HBM bytes per launch: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
                    # (scales from --calibrate-copy), averaged over the launches
                    "traffic": TRAFFIC_BYTES_PER_LAUNCH.get(bp),
                    "kernel": KERNEL_NAME[bp], "flops_per_launch": fl / n,
                    "avg_ms": ms / n, "launches": n, "timed_steps": args.steps,
                    "share_of_step": ms * 1e-3 / run["elapsed"],
                    "shapes": [{"M": k["M"], "N": k["N"], "K": k["K"],
                                "launches": k["launches"], "avg_ms": round(k["avg_ms"], 3),
                                "tflops": round(k["flops_per_launch"] / k["avg_ms"] / 1e9, 2)}
                               for k in ks[:6]]}
        roofline = block(order[0])
        if len(order) > 1:
            roofline["other"] = block(order[1])
    cpu = None
    if not args.no_cpu:
        alg = r64["alg"]
        shapes = {n: tuple(t.shape) for n, t in alg.state.tensors.items()}
        wshapes = {n: tuple(t.shape) for n, t in H.tensors.items()}
        per_step = [(k, w, m / args.steps) for k, w, m in r64["applies"]]
        rate, f_step, t_s = cpu_baseline(args.cpu_chi, per_step, shapes, wshapes)
        facts = host_facts()
        cpu = {"value": rate / f_step, "unit": "sweeps/s", "cores": facts["cpu_count"],
               "kind": "port",
               "sample": (f"oracle (reference contraction order, NumPy/{facts['blas']}) env "
                          f"sweep + root H_eff apply on the same tree at chi={args.cpu_chi} "
                          f"({t_s:.1f}s, {rate / 1e9:.1f} GFLOP/s); scaled to chi=256 by the "
                          f"exact reference-order flops of one step ({f_step / 1e12:.1f} "
                          "TFLOP) -- an extrapolation"),
               "host": facts}
    line = {"metric": METRIC, "value": value, "unit": "sweeps/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": 1e3 / value, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None,
            "dtype": {"fp64": "c128", "relaxed": "c128 (+c64 3M on relaxed applies)",
                      "relaxed_bf16": "c128 (+c64 3M FP32 / bf16 on relaxed applies)",
                      "fp32": "c64-mixed"}[path],
            "data": "synthetic",
            "config": {"workload": "2nd-order 1-site TDVP, ternary tree N=40, chi=256, d=2, "
                                   "random XYZ TTNO D=5, dt=0.05, krylov m<=10 tol 1e-10",
                       "chi": args.chi, "nodes": 40, "precision": path,
                       "relax_low": args.relax_low},
            "roofline": roofline, "cpu_baseline": cpu}
    line.update(res)
    line["autotune"] = {"on": E.AUTOTUNE, "tuned_signatures": len(E.TUNE_LOG)}
    if args.c4_instances > 0:
        r64 = run = ok = alg = None   # free the chi=256 runs before the batching measurement
        E.clear_caches()
        torch.cuda.empty_cache()
        E.PROFILER.phase = "batching-config"
        line["batched_config"] = run_c4(args.c4_instances, args.c4_group, args.c4_single)
    if roofline is not None:
        # every LARGE launch of the dominant (backend, precision) in this process: the
        # set rocprofv3 --stats averages over for that kernel
        allk = E.PROFILER.summary(compute=order[0][1], backend=order[0][0])
        n_all = sum(k["launches"] for k in allk)
        roofline["all_phases"] = {"launches": n_all,
                                  "avg_ms": sum(k["total_ms"] for k in allk) / n_all,
                                  "note": f"every LARGE {order[0][1]} launch of this backend "
                                          "in this process (setup, warmup, timed, evaluation, "
                                          "the other runs' launches of this precision and the "
                                          "batching configuration's): the set rocprofv3 --stats "
                                          "averages over"}
    E.PROFILER.stop()
    if args.launch_log:
        with open(args.launch_log, "w") as f:
            json.dump(E.PROFILER.order, f)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
