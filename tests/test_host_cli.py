"""The C++ host: unit checks, the XML contract against the reference's own
loader/writer output (tests/golden), and CLI error handling. No GPU."""
import json
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIB = ROOT / "part_conc_spt_abc_amd" / "lib"
GOLDEN = ROOT / "tests" / "golden"
INPUT = GOLDEN / "test_files" / "input.xml"
DATA = GOLDEN / "test_files" / "data_20171018_175nm_500nm.xml"


def run(args, **kw):
    return subprocess.run([str(LIB / "pcs_abc"), *args], capture_output=True, text=True, timeout=60, **kw)


def test_host_unit_checks(tmp_path):
    r = subprocess.run([str(LIB / "pcs_host_tests"), str(INPUT), str(DATA), str(tmp_path)], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ", 0 failed" in r.stdout


def test_parse_matches_reference_loader():
    import hashlib

    r = run(["--parse-only", "-i", str(INPUT), "-d", str(DATA)])
    assert r.returncode == 0, r.stderr
    ours = json.loads(r.stdout)
    ref = json.loads((GOLDEN / "ref_xml_parse.json").read_text())
    for k, v in ref.items():
        if isinstance(v, dict):
            vals = ours[k]
            assert len(vals) == v["n"] and float(sum(vals)) == v["sum"], k
            txt = ",".join("%.17g" % x for x in vals)
            assert hashlib.sha256(txt.encode()).hexdigest() == v["sha256"], k
        else:
            assert ours[k] == v, k


def test_output_writer_matches_reference_writer(tmp_path):
    out = tmp_path / "o.xml"
    r = run(["--write-sample", str(out)])
    assert r.returncode == 0, r.stderr
    assert out.read_text() == (GOLDEN / "ref_output.xml").read_text()


@pytest.mark.parametrize("args,msg", [
    (["--bogus"], "Unknown command line argument"),
    (["-S"], "requires"),
    (["-S", "x"], "requires"),
    (["--mode", "turbo"], "--mode"),
    (["-hs", "q"], "Scalar type"),
    (["-i", "/nonexistent/input.xml", "--parse-only"], "Error loading input"),
])
def test_cli_errors(args, msg):
    r = run(args)
    assert r.returncode == 1
    assert msg in r.stderr, r.stderr
