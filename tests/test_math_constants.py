"""The polynomial constants in the device header and in the oracle are the
ones the committed fitting script produces."""
import pathlib
import re
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]


def fitted():
    out = subprocess.run([sys.executable, str(ROOT / "part_conc_spt_abc_amd/csrc/fit_math_coeffs.py")],
                         capture_output=True, text=True, check=True).stdout
    vals = {}
    for line in out.splitlines():
        name, rest = line.split(" ", 1)
        vals[name] = [float.fromhex(v) for v in re.findall(r"'([^']+)'", rest)]
    return vals


def floats(text):
    return [float.fromhex(v.rstrip("f")) for v in re.findall(r"-?0x1(?:\.[0-9a-f]+)?p[+-]\d+f", text)]


def test_device_and_oracle_constants_match_fit():
    v = fitted()
    dev = (ROOT / "part_conc_spt_abc_amd/csrc/pcs_device.h").read_text()
    orc = (ROOT / "oracle/pcs_oracle.c").read_text()
    for name, dev_pat, orc_pat in [
        ("LOG_P", r"constexpr float kLogP0.*?;", r"LOG_P\[8\] = \{.*?\};"),
        ("SIN_S", r"constexpr float kSin0.*?;", r"SIN_S\[4\] = \{.*?\};"),
    ]:
        assert floats(re.search(dev_pat, dev, re.S).group(0)) == v[name], name
        assert floats(re.search(orc_pat, orc, re.S).group(0)) == v[name], name
    cos_dev = floats(re.search(r"constexpr float kCos1.*?;", dev, re.S).group(0))
    assert [1.0] + cos_dev == v["COS_C"]
    assert floats(re.search(r"COS_C\[5\] = \{.*?\};", orc, re.S).group(0)) == v["COS_C"]
