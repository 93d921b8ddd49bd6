"""Where a single serving request's latency goes, on an idle GPU.

Runs lone requests (one video of 64 or more particles, one model: box
200 x 200 x 30 um, H 1000 x 500, no observed tracks) through the device
serving loop of the tracing build of the library (lib/trace/libpcs.so,
`make -C part_conc_spt_abc_amd/csrc trace`), which prints per completion the
host clock at submit, at publication to the request ring and when the result
is seen, and the device wall clock at admission, first tile claim, and the
start of the distance, of its row sweep and its end. Prints the median of each
stage and of the request's track records.

    python -m part_conc_spt_abc_amd.latency_anatomy [--particles 64 6400] [--frames 250 1000]
        [--requests 50] [--gap-ms 1] [--one-wave-tiles]
"""
from __future__ import annotations

import argparse
import os
import pathlib
import subprocess
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
TRACE_LIB = ROOT / "lib" / "trace" / "libpcs.so"
TICK_US = 0.01  # device wall clock: 100 MHz


def child(args):
    import time

    from part_conc_spt_abc_amd.gpu import Server
    from part_conc_spt_abc_amd.model import Model, proposal_record, stokes_einstein

    srv = Server(device=0, arena_bytes=1 << 30, max_inflight=256, device_loop=True,
                 one_wave_tiles=args.one_wave_tiles)
    m = Model(Lx=200.0, Ly=200.0, Lz=30.0, ax=160.0, ay=120.0, deltat=0.05, kmin=3, n_k=1000, n_de=500,
              ub_de=20.0, n_species=2)
    sid = srv.add_system(m)
    p = proposal_record(m, [stokes_einstein(175.0), stokes_einstein(500.0)], [5e8, 5e8], [2.0], 1, 0, 1)
    tag = 0
    for n in args.particles:
        for f in args.frames:
            print(f"case {n} {f} {tag}", file=sys.stderr, flush=True)
            for k in range(args.requests + WARMUP):
                assert srv.submit(tag, sid, p, np.array([f], np.uint32), np.array([n], np.uint32), 100 + k)
                tag += 1
                while not srv.poll(4, wait_us=5000):
                    pass
                time.sleep(args.gap_ms * 1e-3)
    srv.close()


# A pcs-trace line of the tracing build (pcs_server.hip): tag, host clock at
# submit, publish and completion seen (us); device wall clock at admission,
# first claim, distance start, row-sweep start and end (ticks); whether the
# tiles were group-walked; the request's track records.
TRACE_FIELDS = ("submit", "publish", "seen", "admit", "claim", "dist", "sweep", "done", "group_walked", "records")
WARMUP = 10  # requests of each case that are not counted


def parse(text):
    """The cases ("case particles frames first_tag") and the pcs-trace rows
    (tag -> the TRACE_FIELDS values) of a child's stderr."""
    cases, rows = [], {}
    for line in text.splitlines():
        f = line.split()
        if f[:1] == ["case"]:
            cases.append((int(f[1]), int(f[2]), int(f[3])))
        elif f[:1] == ["pcs-trace"]:
            if len(f) != 2 + len(TRACE_FIELDS):
                raise ValueError(f"pcs-trace line with {len(f) - 2} fields, not {len(TRACE_FIELDS)}: {line!r}")
            rows[int(f[1])] = [float(x) for x in f[2:]]
    return cases, rows


def summarize(rows, first, requests):
    """Median records and median stages (us) of the case whose tags start at
    `first`, past its warm-up: submit->publish, admit->claim, walk, distance,
    its record passes, its sweep, ring + return, total; and how many of the
    requests were group-walked, of how many."""
    tags = range(first + WARMUP, first + WARMUP + requests)
    r = np.array([rows[t] for t in tags])
    sub, pub, seen, ta, tc, td, ts, tn, gw, nrec = r.T
    dev = lambda a, b: (b - a) * TICK_US
    stages = np.stack([pub - sub, dev(ta, tc), dev(tc, td), dev(td, tn), dev(td, ts), dev(ts, tn),
                       (seen - pub) - dev(ta, tn), seen - sub], axis=1)
    return int(np.median(nrec)), np.median(stages, axis=0), int(gw.sum()), len(tags)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--particles", type=int, nargs="+", default=[64, 6400])
    ap.add_argument("--frames", type=int, nargs="+", default=[250, 1000])
    ap.add_argument("--requests", type=int, default=50)
    ap.add_argument("--gap-ms", type=float, default=1.0)
    ap.add_argument("--one-wave-tiles", action="store_true")
    ap.add_argument("--child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args(argv)
    if args.child:
        return child(args)
    if not TRACE_LIB.exists():
        raise SystemExit(f"{TRACE_LIB} not found: make -C {ROOT / 'csrc'} trace")
    env = dict(os.environ, PCS_LIBRARY=str(TRACE_LIB))
    cmd = [sys.executable, "-u", "-m", "part_conc_spt_abc_amd.latency_anatomy", "--child",
           "--particles", *map(str, args.particles), "--frames", *map(str, args.frames),
           "--requests", str(args.requests), "--gap-ms", str(args.gap_ms)]
    if args.one_wave_tiles:
        cmd.append("--one-wave-tiles")
    p = subprocess.run(cmd, env=env, cwd=str(ROOT.parent), capture_output=True, text=True, timeout=600)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr[-4000:])
        raise SystemExit(p.returncode)
    cases, rows = parse(p.stderr)
    print(f"{'particles':>9} {'frames':>6} {'records':>7} | {'submit->publish':>15} {'admit->claim':>12} "
          f"{'walk':>8} {'distance':>8} {'(records':>8} {'sweep)':>7} {'ring+return':>11} {'total':>8} | "
          f"group-walked  (median us)")
    for n, f, first in cases:
        rec, med, walked, count = summarize(rows, first, args.requests)
        print(f"{n:9d} {f:6d} {rec:7d} | {med[0]:15.1f} {med[1]:12.1f} {med[2]:8.1f} "
              f"{med[3]:8.1f} {med[4]:8.1f} {med[5]:7.1f} {med[6]:11.1f} {med[7]:8.1f} | "
              f"{walked}/{count}")

if __name__ == "__main__":
    main()
