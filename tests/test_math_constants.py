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
    ker = (ROOT / "part_conc_spt_abc_amd/csrc/pcs_kernels.hip").read_text()
    orc = (ROOT / "oracle/pcs_oracle.c").read_text()
    assert floats(re.search(r"constexpr float kLogP0.*?;", dev, re.S).group(0)) == v["LOG_P"]
    assert floats(re.search(r"LOG_P\[8\] = \{.*?\};", orc, re.S).group(0)) == v["LOG_P"]
    # angle rotation constants K, C, -1/6, -1/2, 1/24 in that order
    ang_dev = floats(re.search(r"constexpr float kAngK.*?kAngC4 = [^;]*;", dev, re.S).group(0))
    ang_orc = floats(re.search(r"static const float ANG_K.*?ANG_C4 = [^;]*;", orc, re.S).group(0))
    assert ang_dev == v["ANGLE_K"] and ang_orc == v["ANGLE_K"]
    # the 256-entry (sin, cos) table, bit for bit in both copies
    tab_dev = floats(re.search(r"kSinCosTab\[kSinCosTabFloats\] = \{.*?\};", ker, re.S).group(0))
    tab_orc = floats(re.search(r"SINCOS_TAB\[512\] = \{.*?\};", orc, re.S).group(0))
    assert tab_dev == v["SINCOS_TAB"] and tab_orc == v["SINCOS_TAB"] and len(tab_dev) == 512


def test_angle_table_is_correctly_rounded():
    """Each table entry is the FP32 value nearest to sin/cos at 50 digits, so
    in particular within half an ulp of FP64 sin/cos."""
    import numpy as np

    v = fitted()["SINCOS_TAB"]
    i = np.arange(256)
    th = 2.0 * np.pi * (i + 0.5) / 256
    want = np.stack([np.sin(th), np.cos(th)], 1).ravel()
    got = np.array(v)
    sp = np.spacing(np.abs(got).astype(np.float32)).astype(np.float64)
    assert (np.abs(got - want) / sp).max() <= 0.5 + 1e-6
