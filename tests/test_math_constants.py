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
    assert floats(re.search(r"LOG_P\[7\] = \{.*?\};", orc, re.S).group(0)) == v["LOG_P"]
    assert floats(re.search(r"constexpr float kExp2Q0.*?;", dev, re.S).group(0)) == v["EXP2_Q"]
    assert floats(re.search(r"EXP2_Q\[6\] = \{.*?\};", orc, re.S).group(0)) == v["EXP2_Q"]
    # the two 256-entry angle tables, bit for bit in both copies
    for name_dev, name_orc, key in (("kSinCosTab", "SINCOS_TAB", "SINCOS_TAB"), ("kBetaTab", "BETA_TAB", "BETA_TAB")):
        tab_dev = floats(re.search(name_dev + r"\[kSinCosTabFloats\] = \{.*?\};", ker, re.S).group(0))
        tab_orc = floats(re.search(name_orc + r"\[512\] = \{.*?\};", orc, re.S).group(0))
        assert tab_dev == v[key] and tab_orc == v[key] and len(tab_dev) == 512
    # -2 ln 2 in both copies: the FP32 value nearest the exact one
    import math

    import numpy as np

    c = floats(re.search(r"constexpr float kNeg2Ln2 = .*?;", dev, re.S).group(0))
    assert c == floats(re.search(r"NEG2_LN2 = .*?;", orc, re.S).group(0)) and len(c) == 1
    x = -2.0 * math.log(2.0)
    assert np.float32(c[0]) == c[0] and abs(c[0] - x) <= 0.5 * float(np.spacing(np.float32(abs(c[0]))))


def _max_ulp(got, want):
    import numpy as np

    got = np.array(got)
    sp = np.spacing(np.abs(got).astype(np.float32)).astype(np.float64)
    return (np.abs(got - want) / sp).max()


def test_angle_table_is_correctly_rounded():
    """Each table entry is the FP32 value nearest to sin/cos at 50 digits, so
    in particular within half an ulp of FP64 sin/cos."""
    import numpy as np

    i = np.arange(256)
    th = 2.0 * np.pi * (i + 0.5) / 256
    assert _max_ulp(fitted()["SINCOS_TAB"], np.stack([np.sin(th), np.cos(th)], 1).ravel()) <= 0.5 + 1e-6


def test_beta_table_is_correctly_rounded():
    """sin beta and cos beta - 1 for beta = 2 pi (r - 127.5) / 2^16; cos beta - 1
    is taken as -2 sin^2(beta/2), which FP64 gives to full relative precision
    (the direct difference would cancel)."""
    import numpy as np

    b = 2.0 * np.pi * (np.arange(256) - 127.5) / 65536
    want = np.stack([np.sin(b), -2.0 * np.sin(b / 2) ** 2], 1).ravel()
    assert _max_ulp(fitted()["BETA_TAB"], want) <= 0.5 + 1e-6
