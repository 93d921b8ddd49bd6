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


def test_cpu_baseline_legs_run_the_cpu_workload():
    """The CPU workload (100 proposals from the prior, 500 frames, H 500 x 200)
    and the oracle runs that time it, on a few proposals, one thread and two."""
    sys.path.insert(0, str(ROOT))
    import bench

    model, jobs, props, frames, parts = bench.cpu_workload()
    assert (model.n_k, model.n_de, model.n_species) == (500, 200, 2) and len(jobs) == len(props) == 100
    assert set(frames.tolist()) == {500}
    lam = [sum(c) * model.volume_factor for _, c in jobs]
    assert 100 <= min(lam) and max(lam) <= 24_000
    assert all(0.1 <= m[0] <= m[1] <= 5.0 for m, _ in jobs)
    obs = np.zeros((model.n_k, model.n_de), np.uint32)
    k = np.argsort(parts)[:3]
    sub = ([jobs[i] for i in k], [props[i] for i in k], frames[k], parts[k])
    for threads in (1, 2):
        sims, steps, dt = bench.cpu_run(model, *sub, obs, threads)
        assert sims > 0 and steps > 0 and dt > 0
    assert 1 <= bench.cpu_threads() <= 16
    hm = bench.headline_model()
    bp, bf, bn = bench.workload(hm, 4, np.random.default_rng(1), (200,))
    r = bench.cpu_batched_sample(hm, bp, bf, bn, np.zeros((hm.n_k, hm.n_de), np.uint32), 4, 2)
    assert r["proposals"] == 4 and r["sims_per_s"] > 0 and r["particle_steps_per_s"] > 0
