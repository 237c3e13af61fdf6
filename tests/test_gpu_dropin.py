"""The reference's own programs, compiled unchanged against include/ and linked with libqbf.so
(recipe: oracle/dropin.mk), run on the GPU from the directory layout they expect (bin/ as the
working directory, fixtures in ../test/data and ../data).

Each reference test prints an error norm against its MATLAB fixture but asserts nothing; here the
printed norm is thresholded.  main_hom and main_inhom write full trajectories, compared frame by
frame with the trajectories the reference C code wrote (data/Wevolv_*.dat).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, l1rel, read_dat

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "oracle", "_ref", "dropin")
PROGS = ("main_hom", "main_inhom", "Cc_test", "Cd_test", "coll_step_test", "I1_test", "I2_test", "I2_test2", "I3_test",
         "I4_test", "lgwt_test", "sim_step_test", "slope_lim_test", "slope_lim_test2")
N = 32


@pytest.fixture(scope="module")
def layout(tmp_path_factory):
    missing = [p for p in PROGS if not os.access(os.path.join(BIN, p), os.X_OK)]
    assert not missing, (f"drop-in programs not built: {missing} (__graft_entry__.build() builds them with "
                         "oracle/dropin.mk where the reference sources are present)")
    base = tmp_path_factory.mktemp("dropin")
    for d in ("bin", "test/data", "data"):
        os.makedirs(base / d, exist_ok=True)
    for f in os.listdir(GOLDEN):
        if f.endswith(".dat"):
            dst = "data" if f.startswith(("W0_hom", "W0_inhom")) else "test/data"
            os.symlink(os.path.join(GOLDEN, f), base / dst / f)
    shutil.copy(os.path.join(GOLDEN, "params_periodic_example.txt"), base / "bin")
    return base


def run(layout, prog, *args):
    print(f"running {prog}", flush=True)
    out = subprocess.run([os.path.join(BIN, prog), *args], cwd=layout / "bin", capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "libqbf" not in out.stderr, out.stderr
    return out.stdout


def number(pattern, text):
    m = re.findall(pattern + r"\s*:?\s*([-+0-9.eE]+)", text)
    assert m, text
    return float(m[-1])


@pytest.mark.parametrize("prog,ref", [("I1_test", "I1_ref.dat"), ("I2_test", "I2_ref.dat"), ("I2_test2", "I2_ref2.dat"),
                                      ("I3_test", "I3_ref.dat"), ("I4_test", "I4_ref.dat")])
def test_integral_programs(layout, prog, ref):
    avg = number("average error", run(layout, prog))
    scale = np.abs(read_dat(ref, np.float64)).mean()
    assert avg / scale < 1e-11


@pytest.mark.parametrize("prog", ["Cc_test", "Cd_test", "coll_step_test", "sim_step_test"])
def test_operator_programs(layout, prog):
    assert number("relative error", run(layout, prog)) < 1e-11


def test_lgwt_and_limiters(layout):
    assert number("total error compared to reference", run(layout, "lgwt_test")) < 1e-13
    assert number("cumulative error", run(layout, "slope_lim_test")) == 0.0
    assert number("cumulative error", run(layout, "slope_lim_test2")) == 0.0


def test_main_hom_trajectory(layout):
    run(layout, "main_hom")
    out = np.fromfile(layout / "data" / "Wevolv_hom.dat").reshape(101, 4, N, N)
    ref = read_dat("Wevolv_hom.dat", shape=(101, 4, N, N))
    assert np.array_equal(out[0], ref[0])
    assert max(l1rel(out[i], ref[i]) for i in range(1, 101)) < 1e-11


def test_main_inhom_trajectory(layout):
    run(layout, "main_inhom", "params_periodic_example.txt")
    out = np.fromfile(layout / "data" / "Wevolv_inhom_periodic.dat").reshape(5, 10, 4, N, N)
    ref = read_dat("Wevolv_inhom_periodic.dat", shape=(5, 10, 4, N, N))
    assert max(l1rel(out[i], ref[i]) for i in range(1, 5)) < 1e-11


@pytest.mark.parametrize("btype", ["Dirichlet", "Maxwell"])
def test_main_inhom_boundaries(layout, btype):
    """main_inhom with Dirichlet and Maxwell boundaries (src/main_inhom.c:367-380), its parameter
    file written here, against the oracle's steps; J = M = 4 keeps the oracle at a few seconds."""
    import oracle as O
    from conftest import rand_state
    L, R, h, dt, nv, steps, lam = 12, 7.5, 0.1, 0.008, 10, 3, 0.4
    lines = ["J = 4", "M = 4", f"L = {L}", f"R = {R}", f"h = {h}", f"numVol = {nv}", f"dt = {dt}",
             f"numsteps = {steps}", f"boundaryType = {btype}", "filenameWinit = ../data/W0_inhom.dat",
             f"filenameWevolv = ../data/Wevolv_{btype}.dat"]
    WL, WR = rand_state(np.random.default_rng(5), 2, N, float(L))
    if btype == "Maxwell":
        WL.tofile(layout / "data" / "WMaxwL.dat")
        WR.tofile(layout / "data" / "WMaxwR.dat")
        lines += [f"lambda = {lam}", "filenameWMaxwL = ../data/WMaxwL.dat", "filenameWMaxwR = ../data/WMaxwR.dat"]
    (layout / "bin" / f"params_{btype}.txt").write_text("\n".join(lines) + "\n")
    run(layout, "main_inhom", f"params_{btype}.txt")
    out = np.fromfile(layout / "data" / f"Wevolv_{btype}.dat").reshape(steps, nv, 4, N, N)
    W = read_dat("W0_inhom.dat", shape=(nv, 4, N, N))
    assert np.array_equal(out[0], W)
    q = O.Quad(N, 4, float(L), R, 4)
    for it in range(1, steps):
        if btype == "Dirichlet":
            W = O.inhom_step_dirichlet(q, W, h, dt)
        else:
            W = O.inhom_step_maxwell(q, W, h, dt, lam, WL, WR)
        err = l1rel(out[it], W)
        print(f"main_inhom {btype} step {it}: {err:.2e}")
        assert err < 1e-11
