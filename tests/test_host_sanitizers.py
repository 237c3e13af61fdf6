"""AddressSanitizer / UBSan runs of the host code (CPU, gcc; no device).

* tests/host_asan/host_paths.c drives the host paths of the drop-in layer (csrc/qbf_compat.c,
  csrc/qbf_quadrature.c) and of the oracle (oracle/orc_*.c), built with -fsanitize=address,undefined
  and -fno-sanitize-recover; the device API comes from tests/host_asan/engine_stub.c (every call
  fails), so the device-backed entry points show their error path.
* The reference's own test programs (built by oracle/dropin.mk against the real library) meet a
  missing and a truncated fixture: ReadData's -1 / -2 end them with a non-zero status before any
  device work (test/*.c return through V_RETURN, include/util.h of the reference)."""
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

PKG = os.path.join(ROOT, "inhom-quantum-boltzmann-fourier-c_amd")
HOST = os.path.join(ROOT, "tests", "host_asan")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin")


def test_host_paths_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "host_paths"
    oracle = [os.path.join(ROOT, "oracle", f) for f in ("orc_fft.c", "orc_quadrature.c", "orc_collision.c",
                                                         "orc_transport.c")]
    cmd = ["gcc", "-std=gnu11", "-D_GNU_SOURCE", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "oracle"),
           os.path.join(HOST, "host_paths.c"), os.path.join(HOST, "engine_stub.c"),
           os.path.join(PKG, "csrc", "qbf_compat.c"), os.path.join(PKG, "csrc", "qbf_quadrature.c"),
           *oracle, "-lm", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="1")
    scratch = tmp_path / "io"
    scratch.mkdir()
    r = subprocess.run([str(exe), str(scratch)], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert "0 failed" in r.stdout


def _layout(base, drop=None, truncate=None):
    for d in ("bin", "test/data", "data"):
        os.makedirs(base / d, exist_ok=True)
    for f in os.listdir(GOLDEN):
        if not f.endswith(".dat") or f == drop:
            continue
        dst = base / ("data" if f.startswith(("W0_hom", "W0_inhom")) else "test/data") / f
        if f == truncate:
            with open(os.path.join(GOLDEN, f), "rb") as src:
                dst.write_bytes(src.read()[:-8])
        else:
            os.symlink(os.path.join(GOLDEN, f), dst)


@pytest.mark.parametrize("prog,fixture", [("Cc_test", "W2.dat"), ("I3_test", "h.dat"), ("slope_lim_test", "Un.dat")])
@pytest.mark.parametrize("kind", ["missing", "truncated"])
def test_dropin_programs_stop_on_bad_fixtures(tmp_path, prog, fixture, kind):
    exe = os.path.join(DROPIN, prog)
    assert os.access(exe, os.X_OK), f"{exe} not built (__graft_entry__.build() builds it with oracle/dropin.mk)"
    _layout(tmp_path, drop=fixture if kind == "missing" else None, truncate=fixture if kind == "truncated" else None)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    r = subprocess.run([exe], cwd=tmp_path / "bin", capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode != 0, r.stdout + r.stderr
    want = "cannot open" if kind == "missing" else "expected"
    assert f"ReadData: " in r.stderr and want in r.stderr and fixture in r.stderr, r.stderr


def test_main_inhom_stops_on_missing_parameter_file(tmp_path):
    exe = os.path.join(DROPIN, "main_inhom")
    assert os.access(exe, os.X_OK), f"{exe} not built"
    _layout(tmp_path)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="")
    r = subprocess.run([exe, "no_such_params.txt"], cwd=tmp_path / "bin", capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode != 0, r.stdout + r.stderr
    assert shutil.which("gcc")
