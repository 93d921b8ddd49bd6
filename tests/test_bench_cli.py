"""bench.py's contract on the CPU: it measures one GPU only and refuses more
before touching a device, and its CPU-baseline leg (the oracle's FP64
restatement of the reference CPU path, timed on a bounded sample) reports
the fields of the bench line's cpu_baseline object."""
import pathlib
import subprocess
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_cpu_baseline_leg_reports_its_sample():
    sys.path.insert(0, str(ROOT))
    import bench

    model = bench.headline_model()
    props, frames, parts = bench.workload(model, 4, np.random.default_rng(1), (200,))
    obs = np.zeros((model.n_k, model.n_de), np.uint32)
    r = bench.cpu_baseline(model, props, frames, parts, obs, 4)
    assert r["kind"] == "port" and r["unit"] == "sims/s" and r["cores"] >= 1
    assert r["value"] > 0 and r["particle_steps_per_s"] > 0
    assert r["sample"].startswith("first 4 proposals")
