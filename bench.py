"""Benchmark: replicated IMC transport phase on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-chosen]

Workload: the plain gray 3D cube throughput configuration -- 40x40x40 cells of
0.05 cm, vacuum faces, T = 1 keV, emission only, sigma_a = 5/cm,
sigma_s = 1/cm, dt = 0.001 sh, 1e7 photons per step, seed 1234
(branson_amd/decks/cube_40_throughput.xml, 5 steps by default).

A step is one replicated time step.  Its host part (step physics, upload of
cells and per-cell source energies, tally copy-out, temperature update) runs
untimed; the timed part is the device phase with every input resident in HBM:
source sampling + photon transport + census compaction (bh_driver_transport),
bracketed by device synchronisation.  value = histories transported (new +
census photons) / summed device-phase seconds of the plain FP64 form (FP64
tracking that replays the reference's RNG streams and arithmetic; persistent
waves -- the one-work-item-per-history form gives bit-identical events and is
reported beside it).

The JSON line also carries the chosen form (FP32 tracking, FP64 energies and
tallies) with its per-cell tally error against the FP64 form, segments/s, the
roofline of the transport kernel (algorithmic bytes per launch over its
HIP-event time against 8 TB/s HBM; measured HBM traffic per launch from the
committed rocprofv3 counter summary) and the oracle's OpenMP CPU step
(source + transport + census) timed on this host over the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DECK = os.path.join(ROOT, "branson_amd", "decks", "cube_40_throughput.xml")
METRIC = "histories/s & segments/s on 1 MI355X, plain FP64 vs chosen at fixed tally error"
HBM_PEAK_GBPS = 8000.0
DT = 0.001  # the deck's dt_start; t_mult = 1
N_PHOTONS = 10_000_000  # the deck's photons per step
# measured HBM bytes per transport launch (rocprofv3 FETCH_SIZE/WRITE_SIZE,
# calibrated against the copy kernel) for this workload and command
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "transport_persistent_fp64_cube40_traffic.md")

# Algorithmic bytes of one transport launch (per history / segment / cell
# crossing / tally flush): photon state in + out (FP64 SoA, 97 B each way),
# per-segment opacity + Fleck + three face coordinates (48 B), 8 B bc +
# neighbour per crossing, abs + track read-modify-write per flush (32 B).
# The FP32-tracking form keeps the same FP64 storage in HBM (it converts on
# load and store), so both forms move the same algorithmic bytes.
BYTES = {"fp64": dict(hist=194, seg=48, cross=8, flush=32),
         "fp32": dict(hist=194, seg=48, cross=8, flush=32)}


def _algorithmic_bytes(form, hist, segs, flushes):
    b = BYTES[form]
    return (b["hist"] * hist + b["seg"] * segs + b["cross"] * max(flushes - hist, 0)
            + b["flush"] * flushes)


def _run_form(capi, steps, warmup, precision, algorithm, keep_step1=False):
    """Warm up on one driver, then time `steps` steps on a fresh one.  Photon
    buffers are reserved up front (device memory is set-up, not step work):
    the census grows to ~3.7x the photons per step over five steps."""
    reserve = N_PHOTONS * (1 + min(steps, 8))
    if warmup:
        with capi.Driver(DECK, precision=precision, algorithm=algorithm,
                         t_stop=DT * warmup) as w:
            w.run()
    out = dict(hist=0, segs=0, flushes=0, dev_s=0.0, step_s=0.0, k_ms=[], launches=[],
               step1=None)
    with capi.Driver(DECK, precision=precision, algorithm=algorithm, t_stop=DT * steps,
                     reserve_photons=reserve) as d:
        eng = capi.lib()
        h = eng.bh_driver_engine(d._d)
        for k in range(steps):
            tp = time.perf_counter()
            d.prepare()
            eng.bh_sync(h)
            t0 = time.perf_counter()
            d.transport()
            eng.bh_sync(h)
            t1 = time.perf_counter()
            r = d.finish()
            out["step_s"] += time.perf_counter() - tp
            if k == 0 and keep_step1:
                out["step1"] = d.tallies()[0]
            out["hist"] += r["transported"]
            out["segs"] += r["n_segments"]
            out["flushes"] += r["n_flushes"]
            out["dev_s"] += t1 - t0
            # transport_ms: HIP events around the transport kernel on the
            # engine's stream
            out["k_ms"].append(r["transport_ms"])
            out["launches"].append((r["transported"], r["n_segments"], r["n_flushes"]))
    return out


def _roofline(form, run):
    """Average over the timed launches: algorithmic bytes per launch over the
    mean HIP-event duration of the transport kernel."""
    n = len(run["k_ms"])
    bytes_mean = sum(_algorithmic_bytes(form, *x) for x in run["launches"]) / n
    ms_mean = sum(run["k_ms"]) / n
    return bytes_mean, ms_mean, bytes_mean / (ms_mean * 1e-3) / 1e9


def _traffic_from_profile():
    """HBM bytes per launch read from the committed counter summary (the
    line 'traffic_bytes_per_launch_mean: <bytes>'); None when absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            m = re.search(r"traffic_bytes_per_launch_mean:\s*([0-9.eE+]+)", f.read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def _cpu_baseline(steps):
    """The oracle's OpenMP step (schedule(guided), per-thread tallies): the
    same source + transport + census phases over the same workload, on this
    host's cores."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    deck = oracle.Deck.from_xml(DECK)
    deck.t_stop = DT * steps
    sim = oracle.Sim(deck)
    hist = segs = 0
    secs = t_secs = 0.0
    while not sim.finished():
        r = sim.step(threads)
        hist += r["transported"]
        segs += r["segments"]
        secs += r["transport_seconds"] + r["source_seconds"]
        t_secs += r["transport_seconds"]
    return {"value": hist / secs, "unit": "histories/s", "cores": threads, "kind": "port",
            "segments_per_s": segs / secs, "transport_only_histories_per_s": hist / t_secs,
            "sample": f"all {steps} steps of the same 40^3 cube deck at 1e7 photons/step "
                      f"({hist} histories, {secs:.2f} s of source + transport + census), "
                      f"oracle C restatement, OpenMP guided, {threads} threads"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-chosen", action="store_true",
                    help="skip the chosen form and the history-form comparison")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: every run "
              f"here has one GPU (replicated mode would give each rank its own share of "
              f"the photons and all-reduce the tallies)", file=sys.stderr)
        sys.exit(2)
    if a.steps < 1:
        sys.exit("--steps must be >= 1")

    from branson_amd import capi
    import numpy as np

    # streaming-copy probe: the HBM rate this device reaches with a plain
    # copy kernel (k_copy_probe, 1 GiB each way), also the calibration kernel
    # for the FETCH_SIZE / WRITE_SIZE counters of the profiles
    with capi.Engine(0) as e:
        copy_gbps = e.copy_probe(1 << 30, 20)

    plain = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_PERSISTENT,
                      keep_step1=True)
    value = plain["hist"] / plain["dev_s"]
    bytes_launch, kms, achieved = _roofline("fp64", plain)

    history = chosen = None
    if not a.no_chosen:
        hi = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_HISTORY)
        history = {"form": "plain fp64, one work-item per history (bit-identical events)",
                   "histories_per_s": hi["hist"] / hi["dev_s"],
                   "segments_per_s": hi["segs"] / hi["dev_s"],
                   "ms_per_step": 1e3 * hi["dev_s"] / a.steps,
                   "transport_kernel_ms_mean": _roofline("fp64", hi)[1]}
        ch = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP32_TRACK,
                       capi.ALGO_PERSISTENT, keep_step1=True)
        p1, c1 = plain["step1"], ch["step1"]
        m = p1 > 0
        err = float(np.max(np.abs(c1[m] - p1[m]) / p1[m])) if m.any() else 0.0
        cb, cms, cach = _roofline("fp32", ch)
        chosen = {
            "form": "fp32 tracking + fp64 energies and tallies, persistent waves",
            "histories_per_s": ch["hist"] / ch["dev_s"],
            "segments_per_s": ch["segs"] / ch["dev_s"],
            "ms_per_step": 1e3 * ch["dev_s"] / a.steps,
            "step1_max_rel_abs_E_vs_fp64": err,
            "transport_kernel_ms_mean": cms,
            "transport_kernel_GBps_algorithmic": cach,
        }

    cpu = None if a.no_cpu else _cpu_baseline(a.steps)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * plain["dev_s"] / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cube_40_throughput: plain gray 3D cube, 40x40x40 cells of "
                               "0.05 cm, vacuum faces, T=1 keV, emission only, sigma_a=5, "
                               "sigma_s=1, dt=0.001 sh, 1e7 photons/step, seed 1234",
                   "form": "plain fp64 (reference RNG streams and arithmetic), "
                           "persistent waves",
                   "histories": plain["hist"], "segments": plain["segs"]},
        "segments_per_s": plain["segs"] / plain["dev_s"],
        # whole step: host physics, uploads of cells and source energies over
        # PCIe, device phase, tally copy-out and temperature update
        "whole_step_histories_per_s": plain["hist"] / plain["step_s"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": _traffic_from_profile(),
                     "kernel": "k_transport_persistent<GlobalTally<false>, double>",
                     "kernel_ms_mean": kms, "algorithmic_bytes_per_launch": bytes_launch,
                     "launches": len(plain["k_ms"]),
                     "launch_histories_segments_flushes": plain["launches"],
                     "launch_kernel_ms": plain["k_ms"], "copy_probe_GBps": copy_gbps},
        "cpu_baseline": cpu,
        "history_form": history,
        "chosen": chosen,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
