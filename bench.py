#!/usr/bin/env python3
"""Benchmark of the propagator hot path on one MI355X.

Workload: 32^3x64 lattice, kappa=0.1632, mu=0.004, antiperiodic time
boundary (BCangle 1,0,0,0), seeded gauge U = exp(i 0.3 H) (seed 4), 12
right-hand sides (spin-colour point sources at the origin for the solves,
deterministic random fields for the operator timing).

A step is one application of the twisted-mass operator to all 12
right-hand sides (FP64, one launch, links shared between rhs).  The value is
its algorithmic HBM throughput: per site 12 x (192 B in + 192 B out) + 576 B
of links = 5184 B, divided by the average launch duration from HIP events on
the library's stream.  Beside it the line carries the single-rhs, FP32, FP16
and 48-rhs operator figures, the 12-rhs solve times to a relative residual of
1e-10 (squared 1e-20) in FP64 and mixed precision, unpreconditioned and
even/odd, the 48-rhs batched solve of four source positions with their pion
correlators, the time of one whole source (120 propagators and the
Pi_mu_nu(x) contraction), and the oracle's CPU operator on the host's threads
(a bounded sample: ten single-rhs applications), on one core, and one CPU
BiCGStab iteration extrapolated to the 12-rhs solve.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-solve] [--no-solves12] [--no-48] [--no-source]
                       [--no-cpu]
       python bench.py --copy-calibration   (reference copy kernel for the HBM counters)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dims", type=str, default="64,32,32,32", help="T,LX,LY,LZ")
    ap.add_argument("--nrhs", type=int, default=12)
    ap.add_argument("--no-solve", action="store_true", help="operator timing only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-source", action="store_true",
                    help="skip the whole-source run (120 propagators, FP16+norm inner solves, contraction)")
    ap.add_argument("--no-solves12", action="store_true",
                    help="skip the six 12-rhs solves (for profiling the whole-source run alone)")
    ap.add_argument("--no-48", action="store_true",
                    help="skip the 48-rhs batched solve (12 spin-colour x 4 source positions)")
    ap.add_argument("--delta", type=float, default=0.001, help="reliable-update factor of the FP32 mixed solver")
    ap.add_argument("--delta16", type=float, default=0.1, help="reliable-update factor of the FP16 mixed solver")
    ap.add_argument("--delta16-eo", type=float, default=0.02,
                    help="reliable-update factor of the FP16 mixed solver on the even/odd system")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device 3 times (the reference kernel for HBM counters)")
    return ap.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_dslash_f64_12rhs_64x32x32x32_hbm_traffic.txt")


def measured_traffic(dims, N):
    """HBM bytes per launch of the 12-rhs FP64 operator at 64x32^3, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (line
    'bytes_per_launch: B'), or None for any other workload."""
    if tuple(dims) != (64, 32, 32, 32) or N != 12 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            if line.startswith("bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    return None


def copy_calibration():
    import torch
    x = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")   # 1 GiB
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": "3 x (1 GiB read + 1 GiB write)"}), flush=True)


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the multi-GPU sharding is a design "
                                   "(source-level split) and is not executed", "n_gpus": a.gpus}))
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration()
        return
    import numpy as np
    import torch
    import cvc_amd

    dims = tuple(int(v) for v in a.dims.split(","))
    kappa, mu = 0.1632, 0.004
    N = a.nrhs
    V = int(np.prod(dims))
    log(f"workload {dims} V={V} nrhs={N}")
    t0 = time.time()
    g = cvc_amd.gauge_random(dims, 4, 0.3)
    log(f"gauge generated in {time.time() - t0:.1f}s")
    torch.cuda.init()
    ctx = cvc_amd.Context(dims, kappa, mu, (1.0, 0.0, 0.0, 0.0), gauge=g)

    fin, fout = ctx.field(N), ctx.field(N)
    fin.fill_random(1)
    for _ in range(a.warmup):
        ctx.dslash(fout, fin)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ms = ctx.dslash_timed(fout, fin, a.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    bytes_site = N * 2 * 192 + 576
    gbs = bytes_site * V / (ms * 1e-3) / 1e9
    log(f"{N}-rhs FP64 operator: {ms:.4f} ms/launch, {gbs:.0f} GB/s algorithmic (wall {wall:.3f}s)")

    extra = {}
    # single-rhs FP64 (recon 18 and 12), FP32 / FP16 N-rhs
    f1i, f1o = ctx.field(1), ctx.field(1)
    f1i.fill_random(2)
    for recon, bps in ((18, 960), (12, 768)):
        ctx.set_link_recon(recon)
        for _ in range(3):
            ctx.dslash(f1o, f1i)
        m1 = ctx.dslash_timed(f1o, f1i, max(10, a.steps))
        extra[f"dslash_1rhs_f64_r{recon}_ms"] = round(m1, 5)
        extra[f"dslash_1rhs_f64_r{recon}_GBs"] = round(bps * V / (m1 * 1e-3) / 1e9, 1)
    # the N-rhs FP64 operator with 2-row links (third row rebuilt; exact for unitary links only)
    ctx.set_link_recon(12)
    for _ in range(3):
        ctx.dslash(fout, fin)
    m12 = ctx.dslash_timed(fout, fin, max(10, a.steps))
    extra[f"dslash_{N}rhs_f64_r12_ms"] = round(m12, 5)
    extra[f"dslash_{N}rhs_f64_r12_GBs"] = round((N * 384 + 384) * V / (m12 * 1e-3) / 1e9, 1)
    ctx.set_link_recon(18)
    del f1i, f1o
    # the single-rhs operator through the host-array entry point the drop-in
    # shim's Q_phi_tbc calls (cvchip_Q_phi: host [V][24] in and out, both
    # directions over PCIe included, into an output array the caller owns as a
    # reference program does); reported beside, never as, the value
    phi_h = np.random.default_rng(5).standard_normal((V, 24))
    xi_h = ctx.Q_phi(phi_h)
    t0 = time.perf_counter()
    for _ in range(5):
        ctx.Q_phi(phi_h, out=xi_h)
    mh = (time.perf_counter() - t0) / 5 * 1e3
    extra["dslash_1rhs_host_arrays_ms"] = round(mh, 3)
    extra["dslash_1rhs_host_arrays_GBs"] = round(960 * V / (mh * 1e-3) / 1e9, 2)
    del phi_h, xi_h
    for prec, name, sb, lb in ((cvc_amd.PREC_F32, "f32", 96, 192), (cvc_amd.PREC_F16, "f16", 52, 192)):
        pi, po = ctx.field(N, prec), ctx.field(N, prec)
        pi.copy_from(fin)
        for _ in range(3):
            ctx.dslash(po, pi)
        mp = ctx.dslash_timed(po, pi, max(10, a.steps))
        extra[f"dslash_{N}rhs_{name}_ms"] = round(mp, 5)
        extra[f"dslash_{N}rhs_{name}_GBs"] = round((N * 2 * sb + lb) * V / (mp * 1e-3) / 1e9, 1)
        del pi, po
    # 48 rhs (12 spin-colour x 4 source positions) in one launch, links shared by all 48
    if N == 12:
        f48i, f48o = ctx.field(48), ctx.field(48)
        f48i.fill_random(3)
        for _ in range(2):
            ctx.dslash(f48o, f48i)
        m48 = ctx.dslash_timed(f48o, f48i, max(10, a.steps // 2))
        extra["dslash_48rhs_f64_ms"] = round(m48, 5)
        extra["dslash_48rhs_f64_GBs"] = round((48 * 384 + 576) * V / (m48 * 1e-3) / 1e9, 1)
        del f48i, f48o
    log(f"operator variants: {extra}")

    if not a.no_solve and not a.no_solves12:
        b = ctx.field(N)
        b.upload(np.stack([cvc_amd.point_source(V, 0, k // 3, k % 3) for k in range(N)]))
        # the reference's algorithm (unpreconditioned), then even/odd preconditioned
        for eo, prec, name in ((False, cvc_amd.PREC_F64, "f64"), (False, cvc_amd.PREC_F32, "mixed_f32"),
                               (False, cvc_amd.PREC_F16, "mixed_f16"), (True, cvc_amd.PREC_F64, "eo_f64"),
                               (True, cvc_amd.PREC_F32, "eo_mixed_f32"), (True, cvc_amd.PREC_F16, "eo_mixed_f16")):
            x = ctx.field(N)
            # each solve twice from the zero initial guess: the first call also fills the
            # library's pool of workspaces, the second is the steady-state cost
            # (solve_<name>_s) that the speed-ups compare
            for rep in (0, 1):
                x.zero()
                torch.cuda.synchronize()
                s0 = time.perf_counter()
                rc, st = ctx.invert_her(x, b, 1e-20, 20000 if prec == cvc_amd.PREC_F64 else 4000, prec=prec,
                                        delta=(a.delta16_eo if eo else a.delta16) if prec == cvc_amd.PREC_F16
                                        else a.delta,
                                        check=False, eo=eo)
                torch.cuda.synchronize()
                el = time.perf_counter() - s0
                extra[f"solve_{name}_{'first_' if rep == 0 else ''}s"] = round(el, 3)
                log(f"solve {name} call {rep}: rc={rc} {el:.2f}s {st}")
            extra[f"solve_{name}_iters"] = st["iterations"]
            extra[f"solve_{name}_true_res2"] = st["true_res2_max"]
            extra[f"solve_{name}_reliable_updates"] = st["reliable_updates"]
            extra[f"solve_{name}_rc"] = rc
            del x
        if extra.get("solve_mixed_f32_s"):
            extra["mixed_f32_speedup"] = round(extra["solve_f64_s"] / extra["solve_mixed_f32_s"], 3)
        if extra.get("solve_mixed_f16_s"):
            extra["mixed_f16_speedup"] = round(extra["solve_f64_s"] / extra["solve_mixed_f16_s"], 3)
        for k in ("eo_f64", "eo_mixed_f32", "eo_mixed_f16"):
            if extra.get(f"solve_{k}_s"):
                extra[f"{k}_speedup"] = round(extra["solve_f64_s"] / extra[f"solve_{k}_s"], 3)
        # the same comparison on the even/odd preconditioned system (FP32 or FP16 inner against FP64)
        for k in ("eo_mixed_f32", "eo_mixed_f16"):
            if extra.get(f"solve_{k}_s") and extra.get("solve_eo_f64_s"):
                extra[f"{k}_vs_eo_f64_speedup"] = round(extra["solve_eo_f64_s"] / extra[f"solve_{k}_s"], 3)

    if not a.no_solve and not a.no_48:
        # 48 rhs: 12 spin-colour x 4 source positions (seed 4) in one batched gamma5-hermitian
        # solve (every operator launch applies D to all 48, links shared), FP32 inner iteration
        # with FP64 reliable updates, then the four pion correlators from the batched propagators
        rng = np.random.default_rng(4)
        srcs = [tuple(int(rng.integers(0, L)) for L in dims) for _ in range(4)]
        b48, x48 = ctx.field(48), ctx.field(48)
        b48.point_sources(srcs)
        for eo, name in ((False, "48rhs_mixed_f32"), (True, "48rhs_eo_mixed_f32")):
            # called twice like the 12-rhs solves: the first call also allocates the
            # workspaces of its sizes, the second is the steady state
            for rep in range(2):
                x48.zero()
                torch.cuda.synchronize()
                s0 = time.perf_counter()
                rc, st = ctx.invert_her(x48, b48, 1e-20, 4000, prec=cvc_amd.PREC_F32, delta=a.delta, check=False,
                                        eo=eo)
                torch.cuda.synchronize()
                el = time.perf_counter() - s0
                if rep == 0:
                    extra[f"solve_{name}_first_s"] = round(el, 3)
            extra[f"solve_{name}_s"] = round(el, 3)
            extra[f"solve_{name}_iters"] = st["iterations"]
            extra[f"solve_{name}_true_res2"] = st["true_res2_max"]
            extra[f"solve_{name}_converged"] = st["converged"]
            log(f"solve {name} (sources {srcs}): rc={rc} {el:.2f}s {st}")
        one = ctx.field(12)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        for k in range(4):
            one.copy_columns(0, x48, 12 * k, 12)
            ctx.field_contract_twopoint(5, 5, one, one, src=srcs[k])
        extra["twopoint_48rhs_s"] = round(time.perf_counter() - s0, 4)
        log(f"four pion correlators from the 48 propagators: {extra['twopoint_48rhs_s']}s")
        # the same 48 columns as four sequential 12-rhs solves (one per source position)
        b12, x12 = ctx.field(12), ctx.field(12)
        for eo, name in ((False, "4x12rhs_mixed_f32"), (True, "4x12rhs_eo_mixed_f32")):
            tot, its = 0.0, []
            for k in range(4):
                b12.copy_columns(0, b48, 12 * k, 12)
                x12.zero()
                torch.cuda.synchronize()
                s0 = time.perf_counter()
                rc, st = ctx.invert_her(x12, b12, 1e-20, 4000, prec=cvc_amd.PREC_F32, delta=a.delta, check=False,
                                        eo=eo)
                torch.cuda.synchronize()
                tot += time.perf_counter() - s0
                its.append(st["iterations"])
            extra[f"solve_{name}_s"] = round(tot, 3)
            extra[f"solve_{name}_iters"] = its
            log(f"solve {name} (same sources, one after the other): {tot:.2f}s iterations {its}")
        del b48, x48, one, b12, x12

    if not a.no_solve and not a.no_source:
        # one source point: 60 up + 60 down propagators (10 batches of 12, even/odd
        # preconditioned, FP16+norm inner, FP64 reliable updates) and the Pi_mu_nu(x) contraction;
        # timed twice: the first call allocates the workspaces, a program looping over sources
        # pays the second call's time per source
        ctx.set_solver_flags(cvc_amd.SOLVE_EO)
        for rep, key in ((0, "source_first_solve_contract_s"), (1, "source_solve_contract_s")):
            torch.cuda.synchronize()
            s0 = time.perf_counter()
            conn, ct, st = ctx.cvc_source((0, 0, 0, 0), 1e-20, 4000, cvc_amd.PREC_F16, a.delta16_eo)
            el = time.perf_counter() - s0
            extra[key] = round(el, 3)
            log(f"whole source {rep} (120 solves FP16+norm inner + contraction): {el:.2f}s {st}")
            del conn
        ctx.set_solver_flags(0)
        extra["source_solves_s"] = round(st["seconds"], 3)
        extra["source_converged"] = st["converged"]
        extra["source_true_res2_max"] = st["true_res2_max"]

    cpu = None
    if not a.no_cpu:
        # the oracle restatement on the host cores: Q_phi_tbc with the reference's OpenMP site
        # distribution (Q_phi.c:190-206), then one core for comparison, then BiCGStab
        # iterations of invert_Qtm_her (operator threaded, vector operations serial as in the
        # reference) on one rhs, extrapolated to the 12-rhs FP64 solve's iteration count
        import oracle
        # every core this process may run on; the environment's OMP_NUM_THREADS caps it (on
        # the GPU hosts it is the CPU share that goes with one GPU)
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = min(omp, avail) if omp > 0 else avail
        extra["cpu_host_nproc"] = os.cpu_count()
        extra["cpu_host_affinity"] = avail
        extra["cpu_omp_num_threads_env"] = omp or None
        L = oracle.Lattice(dims, kappa, mu)
        phi = fin.download(0, 1)[0]
        L.set_threads(threads)
        L.Q_phi(g, phi)
        reps = 10
        c0 = time.perf_counter()
        for _ in range(reps):
            L.Q_phi(g, phi)
        ct = (time.perf_counter() - c0) / reps
        cpu = {"value": round(960 * V / ct / 1e9, 3), "unit": "GB/s", "cores": threads, "kind": "port",
               "sample": f"{reps} single-rhs Q_phi_tbc applications at {'x'.join(map(str, dims))} on {threads} "
                         f"threads ({ct:.3f} s each; oracle restatement, 960 B/site)"}
        L.set_threads(1)
        c0 = time.perf_counter()
        L.Q_phi(g, phi)
        ct1 = time.perf_counter() - c0
        extra["cpu_dslash_1core_GBs"] = round(960 * V / ct1 / 1e9, 3)
        L.set_threads(threads)
        b1 = cvc_amd.point_source(V, 0, 0, 0)
        its = []
        for n in (1, 3):
            c0 = time.perf_counter()
            L.invert_Qtm_her(g, b1, 1e-20, n)
            its.append(time.perf_counter() - c0)
        per_it = (its[1] - its[0]) / 2
        extra["cpu_solve_iteration_s"] = round(per_it, 3)
        if extra.get("solve_f64_iters"):
            extra["cpu_solve_f64_12rhs_extrapolated_s"] = round(per_it * extra["solve_f64_iters"] * N, 1)
        log(f"cpu baseline {cpu}; one core {extra['cpu_dslash_1core_GBs']} GB/s; "
            f"BiCGStab iteration {per_it:.2f} s ({threads} threads)")

    out = {
        "metric": "Dslash HBM GB/s; 12-RHS solve s at 32^3x64, mixed vs FP64 at resid 1e-10",
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"twisted-mass operator, {N} rhs, lattice {a.dims.replace(',', 'x')} "
                               f"(T,LX,LY,LZ), kappa={kappa}, mu={mu}, gauge exp(i 0.3 H) seed 4",
                   "lattice": list(dims), "nrhs": N, "link_recon": 18},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dims, N),
                     "traffic_from": ("committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload: "
                                      + os.path.relpath(TRAFFIC_FILE, ROOT))
                     if measured_traffic(dims, N) is not None else None},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
