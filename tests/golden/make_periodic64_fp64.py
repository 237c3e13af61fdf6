"""Writes periodic64_fp64_ref.npz: the FP64 plain trajectory of the periodic workload at
N = 64, Nx = 512 (J = 32, M = 16, L = 6, R = 3.75, seed 3, h = 1/512, CFL 0.5), 200 Strang
steps of the FP64 engine without conservation correction (the reference's arithmetic; its
parity with the oracle to 1e-12 is tested in tests/test_gpu_parity.py).

Contents: moments[f, cell, 7] (mass, spin x/y/z, momentum x/y, energy) at steps 0, 20, ..., 200,
and the full final state of 16 cells (0, 32, ..., 480).  The fast paths (FP32 convolutions with
FP64 correction, CNN surrogate) are compared against it after 200 steps.

Runs on one MI355X in about 13 minutes:  python -u tests/golden/make_periodic64_fp64.py
"""
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "inhom-quantum-boltzmann-fourier-c_amd"))

import numpy as np  # noqa: E402
import torch  # noqa: E402

import qbf  # noqa: E402
from qbf import workloads as wl  # noqa: E402

N, NX, J, M, SEED, STEPS, EVERY = 64, 512, 32, 16, 3, 200, 20
KEEP = list(range(0, 512, 32))


def main(out=os.path.join(os.path.dirname(os.path.abspath(__file__)), "periodic64_fp64_ref.npz")):
    h, dt = 1.0 / NX, wl.cfl_dt(N, NX)
    e = qbf.Engine(N, qbf.FP64, max_cells=NX)
    e.set_quadrature(*qbf.quadrature(N, J, wl.L_BOX, wl.R_BALL, M))
    W = torch.from_numpy(wl.periodic(N, NX, seed=SEED)).cuda()
    buf = torch.empty_like(W)
    moms = [wl.moments(W.cpu().numpy())]
    t0 = time.time()
    for s in range(1, STEPS + 1):
        e.inhom_step_periodic(W, wl.L_BOX, h, dt, out=buf)
        W, buf = buf, W
        if s % EVERY == 0:
            moms.append(wl.moments(W.cpu().numpy()))
        if s % 10 == 0:
            torch.cuda.synchronize()
            print(f"step {s}/{STEPS}  {time.time() - t0:.0f} s", flush=True)
    final = W.cpu().numpy()[KEEP]
    np.savez(out, moments=np.stack(moms), every=EVERY, steps=STEPS, cells=np.array(KEEP), final=final,
             N=N, Nx=NX, J=J, M=M, seed=SEED, h=h, dt=dt)
    print("wrote", out, flush=True)


if __name__ == "__main__":
    main(*sys.argv[1:2])
