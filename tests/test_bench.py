"""bench.py's contract on the CPU: the one-GPU refusal and the CPU baseline leg (the
GPU workloads themselves run only on a GPU)."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_cpu_baseline_fields():
    # the oracle on a short sample of the torus workload, in a child process (the leg
    # sets OMP_NUM_THREADS for the oracle's threads)
    code = ("import json, bench; node, params, _ = bench.workload('torus-128'); "
            "print(json.dumps(bench.cpu_baseline(node, params, 0.2)))")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       cwd=str(ROOT))
    assert r.returncode == 0, r.stderr[-2000:]
    cb = json.loads(r.stdout.strip().splitlines()[-1])
    assert cb["unit"] == "Mvoxels/s" and cb["kind"] == "port"
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["nproc"] >= 1
    assert "10 chunks of" in cb["sample"] and "from z=64 of the 128^3 job" in cb["sample"]
