"""The allocation-free simulation loop (qbf_simulate_periodic_noalloc, through the C ABI) against
qbf_simulate_periodic (src/simulation.c:1288-1301) on the same inputs, bit for bit: two calls in a
row, the second on fewer cells, then the limits (a state above QBF_NOALLOC_STATE_BYTES, an engine
whose buffers are not yet sized) return -1 without computing."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frames(e, W0, L, h, dt, nsteps, noalloc):
    import qbf
    nf = nsteps
    out = np.empty((nf,) + tuple(W0.shape))
    fn = qbf.lib().qbf_simulate_periodic_noalloc if noalloc else qbf.lib().qbf_simulate_periodic
    e._bind()     # the engine launches on torch's current stream, where W0 was written
    rc = fn(e._h, C.c_void_p(W0.data_ptr()), W0.shape[0], L, h, dt, nsteps, 1, None,
            out.ctypes.data_as(C.c_void_p))
    return rc, out


@pytest.mark.parametrize("N,prec", [(32, "FP64"), (64, "FP32")])
def test_noalloc_loop_equals_allocating_loop(N, prec):
    import torch
    import oracle as O
    import qbf
    from qbf import workloads as wl
    J, M, L = 4, 2, 6.0
    q = O.Quad(N, J, L, 3.75, M)
    e = qbf.Engine(N, getattr(qbf, prec), max_cells=16)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    e.set_moment_correction(qbf.CONS_ALL)
    for ncell, nsteps in ((16, 4), (8, 3)):
        W0 = torch.from_numpy(wl.periodic(N, ncell, seed=ncell)).cuda()
        h, dt = 1.0 / ncell, wl.cfl_dt(N, ncell)
        print(f"N={N} {prec} ncell={ncell} steps={nsteps}: allocating loop, then the allocation-free one", flush=True)
        if ncell == 16:
            rc_a, ref = _frames(e, W0, L, h, dt, nsteps, noalloc=False)
            rc_b, got = _frames(e, W0, L, h, dt, nsteps, noalloc=True)
        else:
            # the second call on fewer cells comes straight after the first allocation-free one
            rc_b, got = _frames(e, W0, L, h, dt, nsteps, noalloc=True)
            rc_a, ref = _frames(e, W0, L, h, dt, nsteps, noalloc=False)
        assert rc_a == 0 and rc_b == 0, qbf.lib().qbf_last_error().decode()
        assert np.isfinite(ref).all()
        assert not np.array_equal(ref[-1], ref[0])
        assert np.array_equal(got, ref)
    e.close()


def test_noalloc_limits():
    import torch
    import oracle as O
    import qbf
    from qbf import workloads as wl
    N, L = 32, 6.0
    q = O.Quad(N, 4, L, 3.75, 2)
    e = qbf.Engine(N, qbf.FP64, max_cells=4)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    W0 = torch.from_numpy(wl.periodic(N, 4, seed=1)).cuda()
    # a state above the limit is rejected before anything runs (the pointer is never read)
    ncell_big = qbf.NOALLOC_STATE_BYTES // (4 * N * N * 8) + 1
    out = np.empty((1, 1))
    rc = qbf.lib().qbf_simulate_periodic_noalloc(e._h, C.c_void_p(W0.data_ptr()), ncell_big, L, 0.25, 1e-4, 1, 1,
                                                  None, out.ctypes.data_as(C.c_void_p))
    assert rc == -1 and "QBF_NOALLOC_STATE_BYTES" in qbf.lib().qbf_last_error().decode()
    # a fresh engine: its buffers would have to grow, so the loop stops at the first step with -1
    rc, frames = _frames(e, W0, L, 0.25, 1e-4, 2, noalloc=True)
    assert rc == -1 and "grow" in qbf.lib().qbf_last_error().decode(), qbf.lib().qbf_last_error().decode()
    assert np.array_equal(frames[0], W0.cpu().numpy())      # frame 0 was copied before the step failed
    # after one allocating step on these cells it runs
    e.inhom_step_periodic(W0, L, 0.25, 1e-4)
    rc, frames = _frames(e, W0, L, 0.25, 1e-4, 2, noalloc=True)
    assert rc == 0, qbf.lib().qbf_last_error().decode()
    e.close()
