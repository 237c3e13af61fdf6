"""bench.py's command line on the CPU: more than one GPU is refused with a message (one GPU per box: the
cells would shard across GPUs with a halo exchange, which is not built), before any device work."""
import os
import subprocess
import sys

from conftest import ROOT


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2, r.stdout + r.stderr
    assert "one GPU" in r.stdout
