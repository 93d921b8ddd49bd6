"""The latency anatomy tool's reading of the tracing build's completion lines
(CPU only: synthetic lines with known stage times)."""
import pathlib
import re

import numpy as np
import pytest

from part_conc_spt_abc_amd import latency_anatomy as la

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_trace_line_fields_match_the_library_format():
    """The pcs-trace line printed by pcs_server.hip has the tag plus one
    conversion per field the tool reads."""
    src = (ROOT / "part_conc_spt_abc_amd" / "csrc" / "pcs_server.hip").read_text()
    m = re.search(r'"pcs-trace ([^"]*)\\n"', src)
    assert m, "pcs-trace format not found"
    conversions = re.findall(r"%[-0-9.]*(?:llu|u|f)", m.group(1))
    assert len(conversions) == 1 + len(la.TRACE_FIELDS)


def trace_line(tag, sub, pub, seen, admit, claim, dist, sweep, done, walked, records):
    return (f"pcs-trace {tag} {sub:.3f} {pub:.3f} {seen:.3f} {admit} {claim} {dist} {sweep} {done} "
            f"{walked} {records}")


def test_stages_of_synthetic_requests():
    requests = 5
    lines = ["case 64 1000 0"]
    for k in range(la.WARMUP + requests):
        # host clock in us; device ticks of 10 ns, request k shifted by k us
        t0 = 1000.0 * k
        a = 100 * (1000 * k + 7)          # admitted 7 us after submit
        lines.append(trace_line(k, t0, t0 + 0.5, t0 + 7 + 160 + 9, a, a + 300, a + 300 + 12000,
                                a + 300 + 12000 + 450, a + 300 + 12000 + 1800, k % 2, 180 + k))
    lines.append("unrelated output")
    cases, rows = la.parse("\n".join(lines))
    assert cases == [(64, 1000, 0)]
    assert len(rows) == la.WARMUP + requests
    rec, med, walked, count = la.summarize(rows, 0, requests)
    assert count == requests
    assert walked == sum(k % 2 for k in range(la.WARMUP, la.WARMUP + requests))
    assert rec == int(np.median([180 + k for k in range(la.WARMUP, la.WARMUP + requests)]))
    submit_publish, admit_claim, walk, distance, passes, sweep, ring, total = med
    assert submit_publish == pytest.approx(0.5)
    assert admit_claim == pytest.approx(3.0)
    assert walk == pytest.approx(120.0)
    assert distance == pytest.approx(18.0)
    assert passes == pytest.approx(4.5) and sweep == pytest.approx(13.5)
    # seen - publish = 175.5 us, of which the device held the request 141 us
    assert ring == pytest.approx(175.5 - 141.0)
    assert total == pytest.approx(176.0)


def test_a_line_with_another_field_count_is_refused():
    with pytest.raises(ValueError):
        la.parse("pcs-trace 1 0.0 0.5 10.0 1 2 3 4 5")
