"""Regenerate the golden fixtures in this directory.

1. ``threefry_kat.json``: Threefry2x64-20 known answers produced by the
   reference's own vendored Random123 header, compiled in place by
   ``make -C oracle ref`` (needs the reference checkout; output binary in
   ``oracle/_ref``).  Inputs are the published all-zero / all-one vectors plus
   the (counter, key) words Branson's RNG forms for the seeds and stream
   numbers its decks use (src/RNG.h:318-330: ctr = {k, seed << 32},
   key = {stream, 0}).
2. ``reference_runs.json``: per-step diagnostics printed by the reference
   executable (src/imc_state.h:264-284) for the two decks in ``decks/``,
   parsed from the run logs captured in this container (``--logs``).

Run from the repository root:  python tests/golden/make_golden.py --logs DIR
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
KAT_BIN = os.path.join(ROOT, "oracle", "_ref", "threefry_kat")
M64 = (1 << 64) - 1


def kat_inputs():
    ins = [(0, 0, 0, 0), (M64, M64, M64, M64),
           (0x243F6A8885A308D3, 0x13198A2E03707344, 0xA4093822299F31D0,
            0x082EFA98EC4E6C89)]
    for seed in (777, 1234, 14706, 0, 0xFFFFFFFF):
        for stream in (0, 1, 2, 99999, 10**13, 10**13 + 123456, 5 * 10**13 + 7,
                       (1 << 63) + 5):
            for ctr in (0, 1, 2, 6, 7, 1000, M64):
                ins.append((ctr, (seed << 32) & M64, stream, 0))
    # carries into the high counter word
    ins.append((M64, 777 << 32, 3, 0))
    ins.append((0, (777 << 32) + 1, 3, 0))
    return ins


def make_kat():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    text = "\n".join("%x %x %x %x" % t for t in kat_inputs()) + "\n"
    out = subprocess.run([KAT_BIN], input=text, capture_output=True, text=True,
                         check=True).stdout
    rows = []
    for line in out.strip().splitlines():
        w = [int(x, 16) for x in line.split()]
        rows.append({"ctr": [hex(w[0]), hex(w[1])], "key": [hex(w[2]), hex(w[3])],
                     "out": [hex(w[4]), hex(w[5])]})
    with open(os.path.join(HERE, "threefry_kat.json"), "w") as f:
        f.write('{"generator": %s,\n "vectors": [\n' % json.dumps(
            "oracle/_ref/threefry_kat (src/random123/threefry.h)"))
        f.write(",\n".join("  " + json.dumps(v) for v in rows))
        f.write("\n ]\n}\n")
    print(f"threefry_kat.json: {len(rows)} vectors")


_STEP = re.compile(r"^Step: (\d+)\s+Start Time: (\S+)\s+End Time: (\S+)\s+dt: (\S+)")
_FIELDS = [
    (re.compile(r"^Total Photons transported: (\d+)"), ["transported"]),
    (re.compile(r"^Emission E: (\S+), Source E: (\S+), Absorption E: (\S+), Exit E: (\S+)"),
     ["emission_E", "source_E", "absorbed_E", "exit_E"]),
    (re.compile(r"^Pre census E: (\S+) Post census E: (\S+) Post census Size: (\d+)"),
     ["pre_census_E", "post_census_E", "census_size"]),
    (re.compile(r"^Pre mat E: (\S+) Post mat E: (\S+)"), ["pre_mat_E", "post_mat_E"]),
    (re.compile(r"^Radiation conservation: (\S+)"), ["rad_conservation"]),
    (re.compile(r"^Material conservation: (\S+)"), ["mat_conservation"]),
]


def parse_log(path):
    steps = []
    cur = None
    for line in open(path):
        line = line.strip()
        m = _STEP.match(line)
        if m:
            cur = {"step": int(m.group(1)), "time": float(m.group(2)),
                   "dt": float(m.group(4))}
            steps.append(cur)
            continue
        if cur is None:
            continue
        for rx, names in _FIELDS:
            m = rx.match(line)
            if m:
                for n, v in zip(names, m.groups()):
                    cur[n] = int(v) if n in ("transported", "census_size") else float(v)
    return steps


def make_runs(logdir):
    runs = {
        "marshak_slab_100": {"deck": "decks/marshak_slab_100.xml", "omp_threads": 1,
                             "log": "cfg1_1thr.out"},
        "cube_40": {"deck": "decks/cube_40.xml", "omp_threads": 8, "log": "cfg2.out"},
    }
    out = {"note": "values as printed by the reference executable (6 significant "
                   "digits, src/imc_state.h:264-284)", "runs": {}}
    for name, r in runs.items():
        out["runs"][name] = {"deck": r["deck"], "omp_threads": r["omp_threads"],
                             "steps": parse_log(os.path.join(logdir, r["log"]))}
    with open(os.path.join(HERE, "reference_runs.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("reference_runs.json:", {k: len(v["steps"]) for k, v in out["runs"].items()})


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--logs", default=None, help="directory with the reference run logs")
    a = ap.parse_args()
    make_kat()
    if a.logs:
        make_runs(a.logs)
