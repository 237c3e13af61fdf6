"""Benchmark: collision evals/s of the Strang-split collision + transport step at N=64, Nx=512.

One step = half transport, collision midpoint step (2 Cc+Cd evaluations per cell), half
transport, over all Nx=512 cells on one GPU: 1024 collision evals per step.  The fast path
(FP32 convolution kernels, FP64 state and time update) is timed for --steps steps after
--warmup steps; the FP64 plain path is timed on the same input for comparison, and the
error of the fast path against it is reported.  Inputs are synthetic (seeded generator of
qbf.workloads.periodic), resident in HBM before the timed region starts.

The CPU baseline (the oracle, a C restatement of the reference algorithm) is timed first,
before the GPU is touched: one full Cc+Cd evaluation at the same N, J, M per process, one
process per core.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "inhom-quantum-boltzmann-fourier-c_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "collision evals/s at N=64,Nx=512: FP32/surrogate vs FP64 plain, ≤1e-3 err"
PEAK = {"fp32": 157.3e12, "fp64": 78.6e12}

def cpu_baseline(N, J, M, cores, seed):
    """One full reference-formulation Cc+Cd evaluation (oracle) per process, `cores` processes."""
    script = (
        "import sys, time\n"
        f"sys.path.insert(0, {ROOT!r}); sys.path.insert(0, {PKG!r})\n"
        "import numpy as np\n"
        "import oracle as O\n"
        "from qbf import workloads as wl\n"
        f"c = int(sys.argv[1]); W = wl.periodic({N}, max(16, {cores}), seed={seed})[c]\n"
        f"q = O.Quad({N}, {J}, wl.L_BOX, wl.R_BALL, {M})\n"
        "WF = np.stack([O.r2c_fill(W[i]) for i in range(4)])\n"
        "t = time.time(); O.CcInt(q, WF); O.CdInt(q, WF); print(time.time() - t)\n"
    )
    env = dict(os.environ, OMP_NUM_THREADS="1", HIP_VISIBLE_DEVICES="")
    t0 = time.time()
    procs = [subprocess.Popen([sys.executable, "-c", script, str(c)], stdout=subprocess.PIPE, text=True, env=env)
             for c in range(cores)]
    per = []
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("oracle sample failed")
        per.append(float(out.strip().splitlines()[-1]))
    wall = time.time() - t0
    return {"value": cores / wall, "unit": "collision evals/s", "cores": cores, "kind": "port",
            "sample": f"{cores} full Cc+Cd evaluations (oracle, reference formulation, N={N}, J={J}, M={M}), "
                      f"one per single-threaded process; wall {wall:.1f} s incl. table setup, "
                      f"median eval {sorted(per)[len(per) // 2]:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--Nx", type=int, default=512)
    ap.add_argument("--J", type=int, default=32)
    ap.add_argument("--M", type=int, default=16)
    ap.add_argument("--fp64-steps", type=int, default=2, help="FP64 plain steps (first one compared)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cores", type=int, default=0)
    ap.add_argument("--seed", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the cells would shard across GPUs with a 2-cell halo exchange, not built here).")
        sys.exit(2)
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py is a single-process, single-GPU benchmark")
        sys.exit(2)

    N, Nx, J, M = args.N, args.Nx, args.J, args.M
    cpu = None
    if not args.no_cpu_baseline:
        cores = args.cpu_cores or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(N, J, M, cores, args.seed)

    import numpy as np
    import torch

    import qbf
    from qbf import workloads as wl

    dev = torch.device("cuda:0")
    W0 = torch.from_numpy(wl.periodic(N, Nx, seed=args.seed)).to(dev)
    h = 1.0 / Nx
    dt = wl.cfl_dt(N, Nx)
    tabs = qbf.quadrature(N, J, wl.L_BOX, wl.R_BALL, M)

    def make(prec):
        e = qbf.Engine(N, prec, max_cells=Nx)
        e.set_quadrature(*tabs)
        return e

    # ---------------------------------------------------------------- fast path (FP32 kernels)
    e32 = make(qbf.FP32)
    e32.set_moment_correction(qbf.CONS_ALL)     # FP64 conservation correction of every collision step
    # warmup: the first step from W0 is kept for the error comparison
    W32_1 = e32.inhom_step_periodic(W0, wl.L_BOX, h, dt)
    bufs = [W32_1.clone(), torch.empty_like(W0)]
    cur = 0
    for i in range(max(0, args.warmup - 1)):
        e32.inhom_step_periodic(bufs[cur], wl.L_BOX, h, dt, out=bufs[1 - cur])
        cur = 1 - cur
    torch.cuda.synchronize()
    e32.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        e32.inhom_step_periodic(bufs[cur], wl.L_BOX, h, dt, out=bufs[1 - cur])
        cur = 1 - cur
    torch.cuda.synchronize()
    t32 = time.perf_counter() - t0
    kms, nlaunch, kapps = e32.profile_read()
    e32.profile(False)
    evals = 2 * Nx * args.steps
    value = evals / t32

    # ---------------------------------------------------------------- FP64 plain path
    e64 = make(qbf.FP64)
    W64_1 = e64.inhom_step_periodic(W0, wl.L_BOX, h, dt)
    torch.cuda.synchronize()
    b64 = [W64_1.clone(), torch.empty_like(W0)]
    n64 = max(1, args.fp64_steps - 1)
    t0 = time.perf_counter()
    for i in range(n64):
        e64.inhom_step_periodic(b64[i % 2], wl.L_BOX, h, dt, out=b64[(i + 1) % 2])
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    value64 = 2 * Nx * n64 / t64

    a = W32_1.cpu().numpy()
    b = W64_1.cpu().numpy()
    m0 = wl.moments(W0.cpu().numpy()).sum(0)
    ma = wl.moments(a).sum(0)
    mb = wl.moments(b).sum(0)
    def drift(m):
        return {"mass": float(abs(m[0] - m0[0]) / abs(m0[0])),
                "momentum": float(np.hypot(m[4] - m0[4], m[5] - m0[5]) / abs(m0[0])),
                "energy": float(abs(m[6] - m0[6]) / abs(m0[6]))}

    err = {
        "step1_state_l1rel": float(np.abs(a - b).sum() / np.abs(b).sum()),
        "step1_increment_l1rel": float(np.abs(a - b).sum() / np.abs(b - W0.cpu().numpy()).sum()),
        "density_rel": float(abs(ma[0] - mb[0]) / abs(mb[0])),
        "energy_rel": float(abs(ma[6] - mb[6]) / abs(mb[6])),
        "step1_drift_fast_path": drift(ma),
        "step1_drift_fp64_plain": drift(mb),
    }

    P = 2 * N
    flop_per_k = 2 * 5 * P * P * np.log2(P * P)     # IFFT2 + FFT2 of a P x P complex grid
    achieved = kapps * flop_per_k / (kms / 1e3) if kms > 0 else 0.0
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "collision evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t32 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"Strang collision+transport step, N={N}, Nx={Nx}, J={J}, M={M}, "
                               f"L={wl.L_BOX}, R={wl.R_BALL}, dt={dt:.4g} (CFL 0.5); FP32 convolution kernels, "
                               "FP64 state, midpoint update and conservation correction",
                   "N": N, "Nx": Nx, "J": J, "M": M, "evals_per_step": 2 * Nx},
        "roofline": {"bound": "mfma", "kernel": "k_i3acc (I3 K-operator: IFFT2 -> psi2 -> FFT2 per term)",
                     "achieved": round(achieved / 1e12, 3), "peak": PEAK["fp32"] / 1e12, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK["fp32"], 4), "traffic": None,
                     "launches": nlaunch, "avg_launch_ms": round(kms / max(1, nlaunch), 4),
                     "kernel_share_of_step": round(kms / 1e3 / t32, 4)},
        "cpu_baseline": cpu,
        "fp64_plain": {"value": round(value64, 3), "unit": "collision evals/s",
                       "ms_per_step": round(1e3 * t64 / n64, 3), "steps": n64},
        "speedup_vs_fp64": round(value / value64, 3),
        "error_vs_fp64": err,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
