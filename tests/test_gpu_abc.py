"""End to end: the C++ ABC-PMC estimator (pcs_abc) on synthetic data with a
known truth, with the parameters of the reference's integration script
(test/test_part_conc_spt_abc.jl:23-77: 60x60x10 um box, 40x40 um field,
50 videos x 250 frames, dt 0.05 s, kmin 2, fixed depth 2 um, m = [1, 3]
um^2/s, c = [1e8, 1e8] /mL, 200 DE bins up to 12, 128 ABC samples,
adaptive gamma, inverse-distance-squared weights), with a shorter trial
budget so the run takes seconds.
"""
import math
import pathlib
import re
import subprocess

import numpy as np
import pytest

from abc_data import make_data, read_output

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parents[1]
PCS_ABC = ROOT / "part_conc_spt_abc_amd" / "lib" / "pcs_abc"


def run_abc(inp, out, mode, seed=7, steps=60, extra=()):
    r = subprocess.run([str(PCS_ABC), "-i", str(inp), "-o", str(out), "-F", str(seed), "-S", str(steps), "--mode",
                        mode, *extra], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return read_output(out)


@pytest.mark.parametrize("proposals", ["host", "device"])
def test_abc_recovers_truth(tmp_path, proposals):
    """Proposals drawn on the host, or on the GPU (pcs_batch_propose) in every
    round, with the launch shape autotuned on the first round."""
    inp, ntracks, _ = make_data(tmp_path)
    assert ntracks > 100
    extra = ("--device-propose-min", "1", "--autotune-min", "64", "-v") if proposals == "device" else (
        "--device-propose-min", "0")
    txt, vec, scal = run_abc(inp, tmp_path / "out.xml", "batched", extra=extra)
    m, c, w = vec("m").reshape(-1, 2), vec("c").reshape(-1, 2), vec("w")
    assert m.shape == (128, 2) and abs(w.sum() - 1.0) < 1e-9
    assert np.all(m[:, 0] <= m[:, 1])  # components ordered by m
    ctot = (c.sum(axis=1) * w).sum()
    print("posterior mean total c", ctot, "m", (m * w[:, None]).sum(axis=0), "gamma", scal("gamma"))
    assert abs(math.log10(ctot) - math.log10(2e8)) < 0.25
    assert int(re.search(r"<number_of_simulations>(\d+)<", txt).group(1)) >= 128


def test_plain_and_batched_forms_give_identical_estimates(tmp_path):
    """Same seed, plain (one launch pair per simulation) vs batched (one
    batch per round): every simulation gives the same distance, so the
    whole ABC run, and its output, is identical."""
    inp, _, _ = make_data(tmp_path, weighting="pmc-standard", trials=4, samples=48)
    for proposals in ("0", "1"):  # drawn on the host, drawn on the GPU
        a, va, sa = run_abc(inp, tmp_path / "a.xml", "batched", seed=3, steps=6, extra=("--device-propose-min", proposals))
        b, vb, sb = run_abc(inp, tmp_path / "b.xml", "plain", seed=3, steps=6, extra=("--device-propose-min", proposals))
        for k in ("m", "c", "az", "dist", "w"):
            assert np.array_equal(va(k), vb(k)), (proposals, k)
        assert sa("gamma") == sb("gamma")
