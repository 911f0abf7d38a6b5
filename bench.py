#!/usr/bin/env python3
"""Benchmark: converged SCF throughput of the batched MNDO/AM1/PM3 SCF path on one MI355X.

Default workload (the one the headline metric is quoted on): 1024 AM1 linear alkanes CnH2n+2,
n ~ U{2..30} (seed 0), zig-zag geometry with N(0, 0.03 A) noise, packed padding-free layout,
FP64 batched diagonalisation density step, Pulay DIIS, convergence 1e-8 eV.

One "step" = one converged single point of the whole batch from coordinates already on the
device: pair integrals + core Hamiltonian, the SCF loop to convergence, the final Fock build and
energies (Hf, Etot, Eelec, Enuc).  value = molecules converged per second.

Also reported: the Fock-build roofline (algorithmic HBM bytes per launch / average launch time
measured with HIP events on the launch stream) and the oracle's FP64 CPU time on a bounded
sample of the same workload.
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


def cpu_baseline(method, sp, xyz, nsample, eps, conv):
    from threadpoolctl import threadpool_limits

    from oracle import seqm_ref as R
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        r = R.single_point(sp[:nsample], xyz[:nsample], method, eps=eps, converger=conv)
        dt = time.perf_counter() - t0
    return dict(value=nsample / dt, unit="molecules/s", cores=1, kind="port",
                sample=f"first {nsample} molecules of the same batch, oracle FP64 numpy single point "
                       f"(same method/eps/mixer), 1 thread, {dt:.2f} s",
                converged=bool(not r["notconverged"].any()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--density", default="eigh", choices=["eigh", "sp2"])
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fock-reps", type=int, default=20)
    ap.add_argument("--compare", action="store_true",
                    help="also run the FP64 diagonalisation path once: speedup and |dHf|/atom")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this path is "
              "data-parallel over molecules; see the sharding notes in DESIGN.md)", file=sys.stderr)
        return 2

    import torch

    from pyseqm_amd.engine import SCFEngine

    method, sp, xyz = make_workload(args.config, args.seed)
    conv = (2,)
    if args.density == "sp2" and args.eps == 1e-8:
        args.eps = 1e-5  # energy tolerance of the mixed-precision path (its precision floor is ~1e-8 eV/atom)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    eng = SCFEngine(sp, None, method, dev)
    coords = torch.as_tensor(xyz, dtype=torch.float64, device=dev)
    nmol = eng.b.nmol

    def step():
        eng.set_coordinates(coords)
        r = eng.scf(eps=args.eps, converger=conv, density=args.density)
        en = eng.energies(r.P)
        return r, en

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

    # SP2 GEMM roofline (one more timed step with events around the SP2 GEMM loops)
    sp2_roof = None
    if args.density == "sp2":
        from pyseqm_amd import sp2 as sp2mod
        eng._sp2ws.timing = []
        step()
        torch.cuda.synchronize()
        t_ms = sum(e0.elapsed_time(e1) for e0, e1, *_ in eng._sp2ws.timing)
        flops = sum(sp2mod.gemm_flops(eng, ix, it) for _, _, ix, _, it in eng._sp2ws.timing)
        eng._sp2ws.timing = None
        ach = flops / (t_ms * 1e-3) / 1e12
        peak = 2500.0 / 3.0
        sp2_roof = {"kernel": "k_sp2_bf16x3 (SP2 X^2, bf16x3 MFMA, FP32 accumulate, fused branch/trace)",
                    "bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                    "traffic": None, "algorithmic_flops": flops, "gemm_loop_ms": t_ms,
                    "note": "peak = dense bf16 2.5 PFLOP/s / 3 products; flops = 2 n^3 per SP2 iteration"}
    # Fock-build roofline: HIP events on the stream the kernels are launched on
    P = r.P
    F = torch.empty_like(P)
    s = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    eng.fock(P, F)
    ev0.record(s)
    for _ in range(args.fock_reps):
        eng.fock(P, F)
    ev1.record(s)
    torch.cuda.synchronize()
    fock_ms = ev0.elapsed_time(ev1) / args.fock_reps
    nbytes, (nxx, nxh, nhh) = fock_algorithmic_bytes(eng.b)
    achieved = nbytes / (fock_ms * 1e-3) / 1e9

    line = {
        "metric": "converged SCF/s", "value": value, "unit": "molecules/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.density == "eigh" else "bf16x3 (fp32 accumulate) + f64 McWeeny", "data": "synthetic",
        "config": {"workload": f"{args.config}: {nmol} {method} molecules "
                               f"({'alkanes n~U[2,30]' if args.config == 'C2' else args.config}), "
                               f"density step {args.density}, Pulay DIIS, eps {args.eps:g} eV",
                   "method": method, "molecules": nmol, "orbitals_total": int(eng.b.norb.sum()),
                   "norb_max": int(eng.b.nmax), "scf_iterations": iters,
                   "parallelism": "single GPU, batch over molecules"},
        "roofline": {"kernel": "pyseqm_fock (Fock build + fused Eelec)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "algorithmic_bytes": nbytes, "launch_ms": fock_ms,
                     "pairs": {"XX": nxx, "XH": nxh, "HH": nhh}},
    }
    if sp2_roof is not None:
        line["roofline_fock"] = line["roofline"]
        line["roofline"] = sp2_roof
    if args.compare:
        e2 = SCFEngine(sp, None, method, dev)
        e2.set_coordinates(coords)
        torch.cuda.synchronize()
        tc = time.perf_counter()
        r2 = e2.scf(eps=1e-8, converger=conv, density="eigh")
        en2 = e2.energies(r2.P)
        torch.cuda.synchronize()
        tc = time.perf_counter() - tc
        dhf = (en["Hf"] - en2["Hf"]).abs().cpu().numpy() / eng.b.nat
        line["compare_fp64_eigh"] = {"ms_per_step": tc * 1e3, "speedup": tc * 1e3 / ms,
                                     "max_abs_dHf_per_atom_eV": float(dhf.max()),
                                     "scf_iterations": r2.iterations}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(method, sp, xyz, args.cpu_sample, args.eps, conv)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
