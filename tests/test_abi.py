"""CPU checks of the drop-in C ABI (no GPU compute calls)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "inhom-quantum-boltzmann-fourier-c_amd", "lib", "libqbf.so")


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INC)):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(INC, h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = "\n".join(l for l in txt.splitlines() if not l.lstrip().startswith("#"))
        for stmt in txt.split(";"):
            stmt = " ".join(stmt.split())
            if "static inline" in stmt or stmt.startswith("typedef") or "(" not in stmt or "{" in stmt:
                continue
            m = re.search(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", stmt)
            if m:
                names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def lib():
    assert os.path.exists(LIB), "libqbf.so not built"
    return C.CDLL(LIB)


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert {"CcInt", "CdInt", "I1integral", "CollisionStep", "InhomStepPeriodic", "SimulationHomogeneous",
            "SlopeLimiterStepMaxwell", "FourierI3", "ReadData", "fftw_execute", "qbf_collision"} <= names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_reference(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "simulation.h"\n#include <stdio.h>\nint main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n",'
                   'sizeof(grid_t),sizeof(grid2_t),sizeof(cgrid_t),sizeof(cgrid2_t),sizeof(wignerV_t),sizeof(wignerF_t),'
                   'sizeof(quadI1_t));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-std=gnu99", "-I" + INC, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    N = 32
    assert got[:6] == [8 * N * N, 32 * N * N, 16 * N * N, 64 * N * N, 32 * N * N, 64 * N * N]
    assert got[6] == 24  # unsigned num (padded) + two pointers, as include/quadrature.h:35-40 of the reference


def test_readdata_error_codes(lib, tmp_path):
    # src/util.c:33-63: -1 missing file, -2 size mismatch, 0 success
    lib.ReadData.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_size_t]
    buf = np.zeros(4)
    assert lib.ReadData(str(tmp_path / "nope.dat").encode(), buf.ctypes.data, 8, 4) == -1
    f = tmp_path / "x.dat"
    np.arange(3.0).tofile(f)
    assert lib.ReadData(str(f).encode(), buf.ctypes.data, 8, 4) == -2
    assert lib.ReadData(str(f).encode(), buf.ctypes.data, 8, 3) == 0 and list(buf[:3]) == [0.0, 1.0, 2.0]
    lib.WriteData.argtypes = [C.c_char_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_bool]
    g = tmp_path / "y.dat"
    assert lib.WriteData(str(g).encode(), buf.ctypes.data, 8, 3, False) == 0
    assert lib.WriteData(str(g).encode(), buf.ctypes.data, 8, 3, True) == 0
    assert np.fromfile(g).size == 6


def test_host_quadrature_matches_oracle():
    import oracle as O
    import qbf
    a = qbf.quadrature(16, 5, 6.0, 3.75, 3, J1=7)
    q = O.Quad(16, 5, 6.0, 3.75, 3, J1=7)
    for x, y in zip(a, (q.I1, q.I2, q.I3, q.I4)):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


def test_host_quadrature_matches_oracle_n128():
    import oracle as O
    import qbf
    a = qbf.quadrature(128, 4, 6.0, 3.75, 2)
    q = O.Quad(128, 4, 6.0, 3.75, 2)
    for x, y in zip(a, (q.I1, q.I2, q.I3, q.I4)):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


@pytest.mark.parametrize("ngrid", [64, 128])
def test_dropin_layer_compiles_for_other_grid_sizes(ngrid, tmp_path):
    """The drop-in layer follows the reference's compile-time N_GRID (include/grid.h)."""
    src = os.path.join(ROOT, "inhom-quantum-boltzmann-fourier-c_amd", "csrc", "qbf_compat.c")
    subprocess.run(["gcc", "-std=c11", "-D_GNU_SOURCE", "-Wall", "-Werror", "-Wno-unused-parameter", "-Wno-comment",
                    f"-DN_GRID={ngrid}", "-I" + INC, "-c", src, "-o", str(tmp_path / "compat.o")], check=True)


def test_product_fails_loudly_without_gpu():
    import torch
    import qbf
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(qbf.QBFError):
        qbf.Engine(32, qbf.FP64, 1)
