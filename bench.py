#!/usr/bin/env python3
"""Benchmark: converged SCF throughput of the batched MNDO/AM1/PM3 SCF path on one MI355X.

Default workload (the one the headline metric is quoted on): 1024 AM1 linear alkanes CnH2n+2,
n ~ U{2..30} (seed 0), zig-zag geometry with N(0, 0.03 A) noise, packed padding-free layout,
FP64 batched diagonalisation density step, Pulay DIIS, convergence 1e-8 eV.

One "step" = one converged single point of the whole batch from coordinates already on the
device: pair integrals + core Hamiltonian, the SCF loop to convergence, the final Fock build and
energies (Hf, Etot, Eelec, Enuc).  value = molecules converged per second.

Also reported, in the same JSON line:
  roofline               the Fock build (algorithmic HBM bytes per launch / average launch time,
                         HIP events on the launch stream);
  sp2_mixed_same_inputs  the mixed-precision SP2 density step on the same batch: throughput,
                         speedup over the FP64 diagonalisation path and |dHf| per atom;
  roofline_sp2           the SP2 X^2 GEMMs on (H2O)_256 x 8 (MFMA-bound);
  serving, p50_ms/p99_ms an open-loop Poisson SCF request stream (half AM1 alkanes, half PM3 water
                         clusters) through the continuous-batching server beside 256 graph-stepped
                         XL-BOMD trajectories: arrival->result latency percentiles, SCF/s;
  cpu_baseline           the oracle's FP64 CPU time on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64


def fock_algorithmic_bytes(b):
    """8 * [sum_XX (100 + 3*16) + sum_XH (10 + 3*4) + sum_HH (1 + 3*1) + 30 * natoms]."""
    kinds = np.repeat(b.tile_kind, 64)
    valid = b.lane_valid
    nxx = int(np.sum(valid & (kinds == 0)))
    nxh = int(np.sum(valid & (kinds == 1)))
    nhh = int(np.sum(valid & (kinds == 2)))
    return 8 * (148 * nxx + 22 * nxh + 4 * nhh + 30 * b.natoms), (nxx, nxh, nhh)


def make_workload(cfg, seed):
    from pyseqm_amd import workloads as W
    c = W.CONFIGS[cfg]
    sp, xyz = c["make"](seed)
    return c["method"], sp, xyz


def _oracle_chunk(job):
    """Worker of the CPU baseline: the oracle's FP64 single point of a few molecules, one BLAS
    thread; returns (first index, Hf, converged, seconds)."""
    start, method, sp, xyz, eps, conv = job
    from threadpoolctl import threadpool_limits

    from oracle import seqm_ref as R
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        r = R.single_point(sp, xyz, method, eps=eps, converger=conv)
        return start, r["Hf"], bool(not r["notconverged"].any()), time.perf_counter() - t0


def make_cpu_pool():
    """Worker processes for the CPU baseline, started before the GPU is touched (a process that
    has opened the GPU does not fork or exec workers): one per host core, at most 16."""
    import multiprocessing as mp
    n = max(1, min(16, len(os.sched_getaffinity(0))))
    return mp.get_context("spawn").Pool(n), n


def cpu_baseline(pool, cores, method, sp, xyz, eps, conv, hf_gpu, chunk=4, target_s=15.0):
    """The oracle (FP64 numpy restatement of the reference algorithm) on the host cores: chunks of
    `chunk` molecules of the same batch, one per worker process, submitted until ~target_s of wall
    time; reports molecules/s over the wall time and the largest |dHf| between the oracle and the GPU
    result for every sampled molecule."""
    t0 = time.perf_counter()
    pending, results, nxt = [], [], 0
    while nxt < len(sp) and (time.perf_counter() - t0 < target_s or not results):
        while nxt < len(sp) and len(pending) < 2 * cores:
            K = int((sp[nxt:nxt + chunk] > 0).sum(1).max())
            pending.append(pool.apply_async(_oracle_chunk, ((nxt, method, sp[nxt:nxt + chunk, :K],
                                                             xyz[nxt:nxt + chunk, :K], eps, conv),)))
            nxt += chunk
        still = []
        for p in pending:
            if p.ready():
                results.append(p.get())
            else:
                still.append(p)
        pending = still
        time.sleep(0.01)
    for p in pending:
        results.append(p.get())
    dt = time.perf_counter() - t0
    done = sum(len(r[1]) for r in results)
    dhf = max(float(np.abs(r[1] - hf_gpu[r[0]:r[0] + len(r[1])]).max()) for r in results)
    return dict(value=done / dt, unit="molecules/s", cores=cores, kind="port",
                sample=f"{done} molecules of the same batch (chunks of {chunk}), oracle FP64 numpy single point "
                       f"(same method/eps/mixer, FP64 eigh), {cores} worker processes x 1 BLAS thread, "
                       f"{dt:.1f} s wall",
                converged=all(r[2] for r in results), max_abs_dHf_vs_gpu_eV=dhf)


def serve(args, dev, rate=None, xl_trajectories=None):
    """Open-loop Poisson SCF stream (half AM1 alkanes, half PM3 water clusters) through the
    continuous-batching server, beside 256 graph-stepped XL-BOMD trajectories."""
    import copy
    args = copy.copy(args)
    if rate is not None:
        args.serve_rate = rate
    if xl_trajectories is not None:
        args.xl_trajectories = xl_trajectories
    import threading

    import torch

    from pyseqm_amd import workloads as W
    from pyseqm_amd.serving import Request, SCFServer, XLService, percentiles
    stream = W.serving_stream(rate=args.serve_rate, window=args.serve_window, seed=args.seed,
                              water_max=args.serve_water_max)
    xl = None
    if args.xl_trajectories > 0:
        sp, xyz = W.alkanes(args.xl_trajectories, seed=args.seed + 7)
        xl = XLService(sp, xyz, "AM1", dev, k=6, dt=0.5)
        for _ in range(3):
            xl.step()
        xl.latencies.clear()
    srv = SCFServer(dev, lanes=args.serve_lanes, density="sp2", eps=args.serve_eps, base_lanes=args.serve_base_lanes,
                    backlog=args.serve_backlog)
    stop = threading.Event()
    th = None
    if xl is not None:
        th = threading.Thread(target=xl.run_until, args=(stop,), daemon=True)
        th.start()
    reqs = []
    t0 = time.perf_counter()
    for k, (ta, sp_, x_, method) in enumerate(stream):
        now = time.perf_counter() - t0
        if now < ta:
            time.sleep(ta - now)
        reqs.append(srv.submit(Request(k, sp_, x_, method, t_arrival=t0 + ta)))
    t_submit_end = time.perf_counter()
    deadline = t_submit_end + args.serve_drain
    for r in reqs:
        r.event.wait(timeout=max(0.0, deadline - time.perf_counter()))
    t_end = time.perf_counter()
    stop.set()
    if th is not None:
        th.join()
    done = [r for r in reqs if r.t_done is not None]
    lat = [r.latency * 1e3 for r in done]
    pending = len(reqs) - len(done)
    # a request still pending at the end counts with its elapsed time (a lower bound)
    lat_all = lat + [(t_end - r.t_arrival) * 1e3 for r in reqs if r.t_done is None]
    with srv.cv:
        srv.queue.clear()
    srv.close()
    out = {
        "offered_rate": args.serve_rate, "window_s": args.serve_window, "requests": len(reqs),
        "completed": len(done), "pending_at_end": pending,
        "converged": int(sum(1 for r in done if r.result and r.result["converged"])),
        "scf_per_s": len(done) / (t_end - t0), "wall_s": t_end - t0,
        "latency_ms": percentiles(lat_all), "latency_ms_completed": percentiles(lat),
        "batches": srv.batches, "lanes": args.serve_lanes, "base_lanes": srv.base_lanes, "backlog": srv.backlog,
        "hbm_budget_GiB": srv.hbm_budget / 2**30,
        "peak_in_flight_GiB": srv.peak_in_use / 2**30, "density": "sp2 (bf16x3 + f64 McWeeny)",
        "eps_eV": args.serve_eps, "mix": f"50% AM1 alkanes n~U[2,30], 50% PM3 (H2O)_N N~U[4,{args.serve_water_max}]",
    }
    if xl is not None:
        n = len(xl.latencies)
        out["xl_bomd"] = {"trajectories": int(xl.traj.eng.b.nmol), "steps": n, "k": 6, "dt_fs": 0.5,
                          "step_ms": percentiles(np.array(xl.latencies) * 1e3),
                          "trajectory_steps_per_s": n * xl.traj.eng.b.nmol / (t_end - t0),
                          "graph_captured": bool(xl.graphs)}
    return out


def sp2_roofline(eng, step):
    """Events around the SP2 GEMM loops of one SCF step: 2 n^3 flops per X^2 per active molecule."""
    import torch

    from pyseqm_amd import sp2 as sp2mod
    eng._sp2ws.timing = []
    step()
    torch.cuda.synchronize()
    t_ms = sum(e0.elapsed_time(e1) for e0, e1, *_ in eng._sp2ws.timing)
    flops = sum(sp2mod.gemm_flops(eng, ix, it) for _, _, ix, _, it in eng._sp2ws.timing)
    exe = sum(sp2mod.executed_mfma_flops(eng, ix, it) for _, _, ix, _, it in eng._sp2ws.timing)
    eng._sp2ws.timing = None
    ach = flops / (t_ms * 1e-3) / 1e12
    peak = 2500.0 / 3.0
    return {"kernel": "k_sp2_bf16x3 (SP2 X^2, bf16x3 MFMA, FP32 accumulate, fused branch/trace)",
            "bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
            "traffic": None, "algorithmic_flops": flops, "gemm_loop_ms": t_ms,
            "executed_mfma_flops": exe, "executed_mfma_frac_of_bf16_peak": exe / (t_ms * 1e-3) / 2.5e15,
            "note": "peak = dense bf16 2.5 PFLOP/s / 3 products (bf16x3 split); flops = 2 n^3 per SP2 iteration "
                    "per active molecule; time = SP2 iteration loops incl. the fused trace/branch epilogue; "
                    "executed = the bf16 MFMA work the kernels issue (upper-triangle tiles only, padded K, "
                    "3 products) over the dense bf16 peak"}


def measured_traffic(config):
    """Calibrated HBM bytes per Fock build from the committed rocprofv3 PMC summary, if present."""
    path = os.path.join(ROOT, "profiles", "fock_pmc_traffic.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        return d[config]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def fock_roofline(eng, P, reps):
    import torch
    F = torch.empty_like(P)
    s = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    eng.fock(P, F)
    ev0.record(s)
    for _ in range(reps):
        eng.fock(P, F)
    ev1.record(s)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    nbytes, (nxx, nxh, nhh) = fock_algorithmic_bytes(eng.b)
    ach = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "pyseqm_fock (k_fock_xx2 + k_fock_tiles<XH/HH> + k_fock_diag + k_mol_sum, fused Eelec)", "bound": "hbm",
            "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
            "algorithmic_bytes": nbytes, "launch_ms": ms, "pairs": {"XX": nxx, "XH": nxh, "HH": nhh}}


def eigh_roofline(eng, P, reps, step_ms, iters):
    """The FP64 density step, the dominant kernel of the FP64 SCF: one launch of the workgroup
    eigensolver (k_eigh_density) over every block, from the converged Fock matrices, timed with HIP
    events on its launch stream.  Algorithmic flops per block 4/3 n^3 + 2 n^2 nocc (Householder
    reduction; back-transformation of the occupied vectors and P = 2 C C^T)."""
    import torch
    b = eng.b
    F = torch.empty_like(P)
    eng.fock(P, F)
    idx = torch.arange(b.nmol, dtype=torch.int32, device=eng.device)
    Pout = torch.empty_like(P)
    eng._idx_host = np.arange(b.nmol)
    eng.density_eigh(F, idx, b.nmol, Pout)
    s = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(s)
    for _ in range(reps):
        eng.density_eigh(F, idx, b.nmol, Pout)
    ev1.record(s)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    n, o = b.norb.astype(np.float64), b.nocc.astype(np.float64)
    flops = float((4.0 / 3.0 * n ** 3 + 2.0 * n * n * o).sum())
    ach = flops / (ms * 1e-3) / 1e12
    return {"kernel": "k_eigh_density (FP64 workgroup eigensolver + density, one launch per density step)",
            "bound": "latency", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": ach / FP64_PEAK_TFLOPS, "algorithmic_flops": flops, "launch_ms": ms,
            "share_of_step": ms * iters / step_ms,
            "note": "the dominant kernel of the FP64 step; one 512-thread workgroup per molecule runs a chain of "
                    "n dependent Householder steps (two barriers each), so the FP64 rate is far below peak by "
                    "construction; its bound is the sum of the per-block latencies over 256 CUs "
                    "(profiles/c2_fp64_eigh_kernel_trace.md)"}


def kernel_only(args, method, sp, xyz, dev):
    """One kernel group on the workload, --steps times, for counter collection."""
    import torch

    from pyseqm_amd.engine import SCFEngine
    eng = SCFEngine(sp, xyz, method, dev)
    torch.cuda.synchronize()
    if args.kernel_only == "fock":
        # the Fock build's memory traffic does not depend on the density's values: the initial
        # guess keeps the profiled program to the few launches being counted
        P = eng.initial_density()
        F = torch.empty_like(P)
        for _ in range(args.steps):
            eng.fock(P, F)
        nbytes, _ = fock_algorithmic_bytes(eng.b)
        info = {"algorithmic_bytes_per_launch": nbytes}
    elif args.kernel_only == "pairs":
        c = torch.as_tensor(xyz, device=dev)
        for _ in range(args.steps):
            eng.set_coordinates(c)
        info = {}
    else:
        eng.scf(eps=args.sp2_eps, converger=(2,), density="sp2")
        for _ in range(args.steps):
            eng.set_coordinates(torch.as_tensor(xyz, device=dev))
            eng.scf(eps=args.sp2_eps, converger=(2,), density="sp2")
        info = {}
    torch.cuda.synchronize()
    print(json.dumps({"kernel_only": args.kernel_only, "config": args.config, "launches": args.steps, **info}),
          flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--density", default="eigh", choices=["eigh", "sp2"])
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--sp2-eps", type=float, default=1e-5,
                    help="SCF energy tolerance of the mixed-precision comparison run")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=4, help="molecules per CPU-baseline work chunk")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fock-reps", type=int, default=20)
    ap.add_argument("--no-compare", action="store_true", help="skip the SP2-mixed run on the same inputs")
    ap.add_argument("--sp2-roofline-config", default="C3", help="workload of the SP2 GEMM roofline ('' = skip)")
    ap.add_argument("--scale-config", default="C4",
                    help="single large system for the Fock / SP2 rooflines at scale ('' = skip)")
    ap.add_argument("--serve-window", type=float, default=10.0, help="seconds of Poisson arrivals (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-rate-light", type=float, default=300.0,
                    help="second, lighter Poisson rate whose latencies show the loop below saturation (0 = skip)")
    ap.add_argument("--serve-water-max", type=int, default=256)
    ap.add_argument("--serve-lanes", type=int, default=6)
    ap.add_argument("--serve-base-lanes", type=int, default=3,
                    help="lanes that always take work; the others only while the backlog is deep")
    ap.add_argument("--serve-backlog", type=int, default=256)
    ap.add_argument("--serve-eps", type=float, default=1e-5)
    ap.add_argument("--serve-drain", type=float, default=20.0)
    ap.add_argument("--xl-trajectories", type=int, default=256)
    ap.add_argument("--kernel-only", choices=["fock", "pairs", "sp2"], default=None,
                    help="profiling aid: run only this kernel group on the workload (--steps launches) and exit")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="profiling aid: copy this many MiB device-to-device --steps times and exit (the plain "
                         "copy kernel whose FETCH_SIZE / WRITE_SIZE calibrate the traffic counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported. The batched SCF is "
              "data-parallel over molecules and requests (one process per GPU, no exchange); only the "
              "one-GPU figure is measured here.", file=sys.stderr)
        return 2

    pool, cores = (None, 0) if (args.no_cpu_baseline or args.kernel_only or args.copy_calibration > 0) \
        else make_cpu_pool()

    import torch

    from pyseqm_amd.engine import SCFEngine

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration > 0:
        nb = int(args.copy_calibration * 2**20) // 8 * 8
        src = torch.ones(nb // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": nb, "launches": args.steps}), flush=True)
        return 0
    method, sp, xyz = make_workload(args.config, args.seed)
    conv = (2,)
    if args.kernel_only:
        return kernel_only(args, method, sp, xyz, dev)
    eng = SCFEngine(sp, None, method, dev)
    coords = torch.as_tensor(xyz, dtype=torch.float64, device=dev)
    nmol = eng.b.nmol
    eps = args.eps if args.density == "eigh" else args.sp2_eps

    def make_step(e, density, tol):
        def step():
            e.set_coordinates(coords)
            r = e.scf(eps=tol, converger=conv, density=density)
            return r, e.energies(r.P)
        return step

    step = make_step(eng, args.density, eps)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters, nconv = [], 0
    for _ in range(args.steps):
        r, en = step()
        iters.append(r.iterations)
        nconv += int((~r.notconverged).sum())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    value = nconv / dt
    hf_main = en["Hf"].clone()

    line = {
        "metric": "converged SCF/s", "value": value, "unit": "molecules/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.density == "eigh" else "bf16x3 (fp32 accumulate) + f64 McWeeny", "data": "synthetic",
        "config": {"workload": f"{args.config}: {nmol} {method} molecules "
                               f"({'alkanes CnH2n+2 n~U[2,30]' if args.config == 'C2' else args.config}), "
                               f"density step {args.density}, Pulay DIIS, eps {eps:g} eV",
                   "method": method, "molecules": nmol, "orbitals_total": int(eng.b.norb.sum()),
                   "norb_max": int(eng.b.nmax), "scf_iterations": iters, "seed": args.seed,
                   "parallelism": "single GPU, batch over molecules"},
        "roofline": fock_roofline(eng, r.P, args.fock_reps),
    }
    if args.density == "eigh":
        line["roofline_eigh"] = eigh_roofline(eng, r.P, max(3, args.fock_reps // 4), ms, max(iters))
    traffic = measured_traffic(args.config)
    if traffic is not None:
        line["roofline"]["traffic"] = traffic
        line["roofline"]["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per Fock build, calibrated "
                                              "by a 1 GiB copy (profiles/fock_hbm_traffic.md)")
    if not args.no_compare and args.density == "eigh":
        e2 = SCFEngine(sp, None, method, dev)
        step2 = make_step(e2, "sp2", args.sp2_eps)
        step2()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            r2, en2 = step2()
        torch.cuda.synchronize()
        ms2 = (time.perf_counter() - t1) / args.steps * 1e3
        dhf = (en2["Hf"] - hf_main).abs().cpu().numpy() / eng.b.nat
        # the FP64 path at the SP2 run's own tolerances (energy eps and DIIS test): like for like
        e5 = SCFEngine(sp, None, method, dev)
        step5 = lambda: (e5.set_coordinates(coords), e5.scf(eps=args.sp2_eps, converger=conv, density="eigh",
                                                          diis_tol=e2.diis_tol))[1]
        step5()
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        for _ in range(args.steps):
            r5 = step5()
        torch.cuda.synchronize()
        ms5 = (time.perf_counter() - t5) / args.steps * 1e3
        del e5
        line["sp2_mixed_same_inputs"] = {
            "value": int((~r2.notconverged).sum()) / (ms2 * 1e-3), "ms_per_step": ms2, "speedup_vs_fp64_eigh": ms / ms2,
            "max_abs_dHf_per_atom_eV": float(dhf.max()), "mean_abs_dHf_per_atom_eV": float(dhf.mean()),
            "scf_iterations": r2.iterations, "eps_eV": args.sp2_eps, "diis_tol_eV": e2.diis_tol,
            "fp64_eigh_same_tolerances": {"ms_per_step": ms5, "scf_iterations": r5.iterations,
                                          "speedup_sp2_vs_this": ms5 / ms2},
            "note": "speedup_vs_fp64_eigh compares with the headline FP64 run (eps %g, DIIS test %g); "
                    "fp64_eigh_same_tolerances reruns the FP64 path at the SP2 run's eps and DIIS test"
                    % (args.eps, eng.diis_tol)}
    c3_cpu = None
    if args.sp2_roofline_config:
        m3, sp3, xyz3 = make_workload(args.sp2_roofline_config, args.seed)
        if pool is not None:
            # CPU baseline of this workload: the oracle's FP64 single point of its first molecule on
            # one worker (one BLAS thread), running while the GPU phases below execute
            K3 = int((sp3[:1] > 0).sum())
            c3_cpu = pool.apply_async(_oracle_chunk, ((0, m3, sp3[:1, :K3], xyz3[:1, :K3], args.eps, conv),))
        e3 = SCFEngine(sp3, None, m3, dev)
        c3 = torch.as_tensor(xyz3, dtype=torch.float64, device=dev)

        def step3():
            e3.set_coordinates(c3)
            return e3.scf(eps=args.sp2_eps, converger=conv, density="sp2")
        step3()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        r3 = step3()
        torch.cuda.synchronize()
        t3 = time.perf_counter() - t3
        roof = sp2_roofline(e3, step3)
        roof["workload"] = f"{args.sp2_roofline_config}: {e3.b.nmol} x {int(e3.b.nmax)} orbitals, SCF {t3 * 1e3:.1f} ms, " \
                           f"{r3.iterations} SCF iterations"
        line["roofline_sp2"] = roof
        # the precision figure of this workload: SP2 mixed vs the FP64 diagonalisation SCF
        hf3 = e3.energies(r3.P)["Hf"]
        e6 = SCFEngine(sp3, None, m3, dev)
        e6.set_coordinates(c3)
        torch.cuda.synchronize()
        t6 = time.perf_counter()
        r6 = e6.scf(eps=args.eps, converger=conv, density="eigh")
        hf6 = e6.energies(r6.P)["Hf"]
        torch.cuda.synchronize()
        t6 = time.perf_counter() - t6
        d3 = ((hf3 - hf6).abs() / torch.as_tensor(e3.b.nat, device=dev)).cpu().numpy()
        line["sp2_vs_fp64_precision"] = {
            "workload": f"{args.sp2_roofline_config}: {e3.b.nmol} x {int(e3.b.nmax)} orbitals ({m3})",
            "max_abs_dHf_per_atom_eV": float(d3.max()), "mean_abs_dHf_per_atom_eV": float(d3.mean()),
            "fp64_eigh_scf_ms": t6 * 1e3, "fp64_scf_iterations": r6.iterations, "sp2_scf_ms": t3 * 1e3,
            "sp2_scf_iterations": r3.iterations, "bar_eV_per_atom": 1e-4}
        del e6
    if args.scale_config:
        # one large system: Fock HBM rate and SP2 MFMA use against peak, SP2 SCF time
        m4, sp4, xyz4 = make_workload(args.scale_config, args.seed)
        e4 = SCFEngine(sp4, None, m4, dev)
        c4 = torch.as_tensor(xyz4, dtype=torch.float64, device=dev)

        def step4():
            e4.set_coordinates(c4)
            return e4.scf(eps=args.sp2_eps, converger=conv, density="sp2")
        step4()
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        r4 = step4()
        torch.cuda.synchronize()
        t4 = time.perf_counter() - t4
        line["roofline_scale"] = {
            "workload": f"{args.scale_config}: {e4.b.nmol} x {int(e4.b.nmax)} orbitals, {e4.b.natoms} atoms, "
                        f"mixed SP2 SCF {t4 * 1e3:.1f} ms, {r4.iterations} SCF iterations, converged "
                        f"{int((~r4.notconverged).sum())}/{e4.b.nmol}",
            "fock": dict(fock_roofline(e4, r4.P, args.fock_reps), traffic=measured_traffic(args.scale_config)),
            "sp2": sp2_roofline(e4, step4)}
        del e4
    if c3_cpu is not None:
        _, hf_o, conv_o, sec = c3_cpu.get()
        line["cpu_baseline_c3"] = {
            "value": 1.0 / sec, "unit": "molecules/s", "cores": 1, "kind": "port",
            "sample": f"molecule 0 of {args.sp2_roofline_config} ({int(e3.b.norb[0])} orbitals, {m3}), oracle FP64 "
                      f"numpy single point at the FP64 run's eps / mixer, 1 worker x 1 BLAS thread, {sec:.1f} s",
            "converged": conv_o, "max_abs_dHf_vs_gpu_eV": float(abs(hf_o[0] - float(hf6[0]))),
            "gpu_fp64_molecules_per_s": e3.b.nmol / t6, "gpu_sp2_molecules_per_s": e3.b.nmol / t3}
    if args.serve_window > 0:
        sv = serve(args, dev)
        line["serving"] = sv
        line["p50_ms"] = sv["latency_ms"]["p50"]
        line["p99_ms"] = sv["latency_ms"]["p99"]
        if args.serve_rate_light > 0:
            line["serving_light"] = serve(args, dev, rate=args.serve_rate_light)
    if pool is not None:
        line["cpu_baseline"] = cpu_baseline(pool, cores, method, sp, xyz, args.eps, conv,
                                            hf_main.cpu().numpy(), args.cpu_sample, args.cpu_seconds)
        pool.terminate()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
