"""Benchmark: replicated IMC transport phase on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: the plain gray 3D cube throughput configuration -- 40x40x40 cells of
0.05 cm, vacuum faces, T = 1 keV, emission only, sigma_a = 5/cm,
sigma_s = 1/cm, dt = 0.001 sh, 1e7 photons per step, seed 1234
(branson_amd/decks/cube_40_throughput.xml, 5 steps by default).

A step is one replicated time step.  Its host part (step physics, upload of
cells and per-cell source energies, tally copy-out, temperature update) runs
untimed; the timed part is the device phase with every input resident in HBM:
source sampling + photon transport + census compaction (bh_driver_transport),
bracketed by device synchronisation.  value = histories transported (new +
census photons) / summed device-phase seconds, for the plain FP64 form.

The JSON line also carries the chosen form (FP32 tracking, persistent waves),
segments/s, the roofline of the transport kernel (algorithmic bytes per launch
over its HIP-event time against 8 TB/s HBM), and the oracle's OpenMP CPU
transport timed on this host on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

DECK = os.path.join(ROOT, "branson_amd", "decks", "cube_40_throughput.xml")
METRIC = "histories/s & segments/s on 1 MI355X, plain FP64 vs chosen at fixed tally error"
HBM_PEAK_GBPS = 8000.0
DT = 0.001  # the deck's dt_start; t_mult = 1

# Algorithmic bytes of one transport launch (per history / segment / crossing
# / tally flush): FP64 state 97 B in + out, per-segment opacity/Fleck/face
# reads, 8 B bc+next per crossing, abs+track read-modify-write per flush.
BYTES = {"fp64": dict(hist=194, seg=48, cross=8, flush=32),
         "fp32": dict(hist=138, seg=36, cross=8, flush=32)}


N_PHOTONS = 10_000_000  # the deck's photons per step


def _run_form(capi, steps, warmup, precision, algorithm, keep_step1=False):
    """Warm up on one driver, then time `steps` steps on a fresh one.  Photon
    buffers are reserved up front (device memory is set-up, not step work):
    the census grows to ~3.7x the photons per step over five steps."""
    reserve = N_PHOTONS * (1 + min(steps, 8))
    if warmup:
        with capi.Driver(DECK, precision=precision, algorithm=algorithm,
                         t_stop=DT * warmup) as w:
            w.run()
    out = dict(hist=0, segs=0, flushes=0, dev_s=0.0, k_ms=[], k_bytes=[], step1=None)
    with capi.Driver(DECK, precision=precision, algorithm=algorithm, t_stop=DT * steps,
                     reserve_photons=reserve) as d:
        eng = capi.lib()
        h = eng.bh_driver_engine(d._d)
        for k in range(steps):
            d.prepare()
            eng.bh_sync(h)
            t0 = time.perf_counter()
            d.transport()
            eng.bh_sync(h)
            t1 = time.perf_counter()
            r = d.finish()
            if k == 0 and keep_step1:
                out["step1"] = d.tallies()[0]
            out["hist"] += r["transported"]
            out["segs"] += r["n_segments"]
            out["flushes"] += r["n_flushes"]
            out["dev_s"] += t1 - t0
            out["k_ms"].append(r["transport_ms"])
            out["k_bytes"].append((r["transported"], r["n_segments"], r["n_flushes"]))
    return out


def _algorithmic_bytes(form, hist, segs, flushes):
    b = BYTES[form]
    return b["hist"] * hist + b["seg"] * segs + b["cross"] * max(flushes - hist, 0) + \
        b["flush"] * flushes


def _cpu_baseline(steps_sample=2):
    """The oracle's OpenMP transport (schedule(guided), per-thread tallies)
    on the first steps of the same workload, on this host's cores."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, 16))
    deck = oracle.Deck.from_xml(DECK)
    deck.t_stop = DT * steps_sample
    sim = oracle.Sim(deck)
    hist = segs = 0
    secs = 0.0
    while not sim.finished():
        r = sim.step(threads)
        hist += r["transported"]
        segs += r["segments"]
        secs += r["transport_seconds"]
    return {"value": hist / secs, "unit": "histories/s", "cores": threads, "kind": "port",
            "segments_per_s": segs / secs,
            "sample": f"first {steps_sample} of the 5 steps of the same 40^3 cube deck at "
                      f"1e7 photons/step ({hist} histories, {secs:.2f} s of transport), "
                      f"oracle C restatement, OpenMP guided, {threads} threads"}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-chosen", action="store_true", help="skip the chosen form")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              f"(multi-GPU replicated mode would shard photons by rank, see DESIGN.md)",
              file=sys.stderr)
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

    plain = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_HISTORY,
                      keep_step1=True)
    value = plain["hist"] / plain["dev_s"]
    # roofline of the dominant kernel: the last (largest) transport launch
    kms = plain["k_ms"][-1]
    h, s, f = plain["k_bytes"][-1]
    bytes_launch = _algorithmic_bytes("fp64", h, s, f)
    achieved = bytes_launch / (kms * 1e-3) / 1e9

    chosen = None
    if not a.no_chosen:
        ch = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP32_TRACK,
                       capi.ALGO_PERSISTENT, keep_step1=True)
        p1, c1 = plain["step1"], ch["step1"]
        m = p1 > 0
        err = float(np.max(np.abs(c1[m] - p1[m]) / p1[m])) if m.any() else 0.0
        ch_kms = ch["k_ms"][-1]
        hh, ss, ff = ch["k_bytes"][-1]
        chosen = {
            "form": "fp32 tracking + fp64 tallies, persistent waves",
            "histories_per_s": ch["hist"] / ch["dev_s"],
            "segments_per_s": ch["segs"] / ch["dev_s"],
            "ms_per_step": 1e3 * ch["dev_s"] / a.steps,
            "step1_max_rel_abs_E_vs_fp64": err,
            "transport_kernel_ms_last": ch_kms,
            "transport_kernel_GBps_algorithmic": _algorithmic_bytes("fp32", hh, ss, ff)
                                                 / (ch_kms * 1e-3) / 1e9,
        }

    cpu = None if a.no_cpu else _cpu_baseline()

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
                   "form": "plain fp64, one work-item per history",
                   "histories": plain["hist"], "segments": plain["segs"]},
        "segments_per_s": plain["segs"] / plain["dev_s"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "k_transport_history<*, double>",
                     "kernel_ms": kms, "algorithmic_bytes": bytes_launch,
                     "copy_probe_GBps": copy_gbps},
        "cpu_baseline": cpu,
        "chosen": chosen,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
