"""Benchmark: replicated IMC transport phase on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-chosen]

Workload: the plain gray 3D cube throughput configuration -- 40x40x40 cells of
0.05 cm, vacuum faces, T = 1 keV, emission only, sigma_a = 5/cm,
sigma_s = 1/cm, dt = 0.001 sh, 1e7 photons per step, seed 1234
(branson_amd/decks/cube_40_throughput.xml, 5 steps by default).

A step is one replicated time step.  Its host part (step physics, upload of
cells and per-cell source energies, tally copy-out, temperature update) runs
untimed; the timed part is the device phase with every input resident in HBM:
source sampling + photon transport (the kernels write the census) + census
order and totals (bh_driver_transport), bracketed by device synchronisation.  value = histories transported (new +
census photons) / summed device-phase seconds of the plain FP64 form (FP64
tracking that replays the reference's RNG streams and arithmetic; persistent
waves -- the one-work-item-per-history form gives bit-identical events and is
reported beside it, and so are the same steps with each device phase launched
as one HIP graph).

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


def _run_form(capi, steps, warmup, precision, algorithm, keep_step1=False, deck=DECK,
              n_photons=N_PHOTONS, n_groups=1, fp32_tol=0.0, step_graph=False):
    """Warm up on one driver, then time `steps` steps on a fresh one.  Photon
    buffers are reserved up front (device memory is set-up, not step work):
    the census grows to ~3.7x the photons per step over five steps.
    step_graph: each device phase is one HIP graph launch (captured on the
    first two graph steps, one graph per parity of the photon stores; the
    captures are inside the timed steps); transport_ms then times the whole
    graph."""
    reserve = n_photons * (1 + min(steps, 8))
    kw = dict(precision=precision, algorithm=algorithm, n_groups=n_groups, fp32_tol=fp32_tol,
              step_graph=step_graph)
    if warmup:
        with capi.Driver(deck, t_stop=DT * warmup, **kw) as w:
            w.run()
    out = dict(hist=0, segs=0, flushes=0, dev_s=0.0, step_s=0.0, k_ms=[], launches=[],
               step1=None, replayed=0, lattice=None, tally=None)
    with capi.Driver(deck, t_stop=DT * steps, reserve_photons=reserve, **kw) as d:
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
            out["replayed"] += r["n_replayed"]
            out["lattice"] = bool(r["lattice"])
            out["tally"] = TALLY_FORMS.get(r["tally"], r["tally"])
            out["segs"] += r["n_segments"]
            out["flushes"] += r["n_flushes"]
            out["dev_s"] += t1 - t0
            # transport_ms: HIP events around the transport kernel on the
            # engine's stream
            out["k_ms"].append(r["transport_ms"])
            out["launches"].append((r["transported"], r["n_segments"], r["n_flushes"]))
    return out


def _run_resident(capi, steps, warmup, deck=DECK, n_photons=N_PHOTONS, n_groups=1):
    """Whole steps with the step physics on the device (driver device_physics):
    wall clock of transport + finish per step, nothing but the step report
    crossing PCIe.  prepare() is untimed: in this mode it uploads the mesh and
    material on the first step (set-up, as the photon store reservation) and
    does nothing after.  Histories per second of whole steps."""
    reserve = n_photons * (1 + min(steps, 8))
    kw = dict(algorithm=capi.ALGO_PERSISTENT, n_groups=n_groups, device_physics=True)
    if warmup:
        with capi.Driver(deck, t_stop=DT * warmup, **kw) as w:
            w.run()
    hist, secs = 0, 0.0
    with capi.Driver(deck, t_stop=DT * steps, reserve_photons=reserve, **kw) as d:
        for _ in range(steps):
            d.prepare()
            t0 = time.perf_counter()
            d.transport()
            r = d.finish()
            secs += time.perf_counter() - t0
            hist += r["transported"]
    return hist / secs


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


def _tally_error(ref, got):
    """Per-cell error of a tally against the FP64 form: max and 99th
    percentile of |got - ref| / ref over cells with ref > 0, and the
    normwise max |got - ref| / max ref."""
    import numpy as np
    m = ref > 0
    if not m.any():
        return {"max_rel": 0.0, "p99_rel": 0.0, "normwise": 0.0}
    rel = np.abs(got[m] - ref[m]) / ref[m]
    return {"max_rel": float(rel.max()), "p99_rel": float(np.percentile(rel, 99)),
            "normwise": float(np.abs(got - ref).max() / ref.max()),
            "zero_cells_differ": int(np.count_nonzero(got[~m]))}


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


HOHLRAUM = os.path.join(ROOT, "branson_amd", "decks", "hohlraum_64.xml")
TALLY_FORMS = {1: "one global atomic per flush (wave ring)", 2: "block-private LDS",
               3: "two global atomics per flush", 4: "per-wave LDS hash table"}


def _common(metric, value, steps, warmup, ms, dtype, config, **extra):
    line = {"metric": metric, "value": value, "unit": "histories/s", "n_gpus": 1,
            "steps": steps, "warmup": warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
            "config": config}
    line.update(extra)
    return line


def _bench_hohlraum(a, capi):
    """30-group 64^3 hohlraum, 5e7 photons/step: plain FP64 against FP32
    tracking (persistent waves) at the stated per-cell tally error."""
    n_ph = 50_000_000
    kw = dict(deck=HOHLRAUM, n_photons=n_ph, n_groups=30, keep_step1=True)
    plain = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_PERSISTENT, **kw)
    ch = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP32_TRACK, capi.ALGO_PERSISTENT,
                   fp32_tol=a.fp32_tol, **kw)
    hist = plain["hist"] / plain["dev_s"]
    chosen = {"form": "fp32 tracking + fp64 energies and tallies, persistent waves",
              "fp32_tol": a.fp32_tol or 1e-6,
              "histories_per_s": ch["hist"] / ch["dev_s"],
              "segments_per_s": ch["segs"] / ch["dev_s"],
              "ms_per_step": 1e3 * ch["dev_s"] / a.steps,
              "speedup_vs_plain_fp64": (ch["hist"] / ch["dev_s"]) / hist,
              "step1_tally_error_vs_fp64": _tally_error(plain["step1"], ch["step1"]),
              "replayed_histories": ch["replayed"]}
    resident = _run_resident(capi, a.steps, a.warmup, deck=HOHLRAUM, n_photons=n_ph,
                             n_groups=30)
    return _common(METRIC, hist, a.steps, a.warmup, 1e3 * plain["dev_s"] / a.steps, "f64",
                   {"workload": "hohlraum_64: 64^3 cells of 0.01 cm, 4-cell wall sigma_a=1000 "
                                "rho=19, fill sigma_a=0.1 sigma_s=0.01 rho=0.001, 0.3 keV "
                                "source on z=0, 30 groups, 5e7 photons/step",
                    "form": "plain fp64, persistent waves",
                    "cell_records": "lattice" if plain["lattice"] else "general",
                    "tally": plain["tally"],
                    "histories": plain["hist"], "segments": plain["segs"]},
                   segments_per_s=plain["segs"] / plain["dev_s"],
                   whole_step_histories_per_s=plain["hist"] / plain["step_s"],
                   resident_whole_step_histories_per_s=resident, chosen=chosen)


def _bench_batched(a, capi):
    """256 independent slab/cube problems in one engine, 3 steps."""
    import numpy as np
    from branson_amd import problems
    from branson_amd.batched import BatchedRun

    steps = min(a.steps, 3)
    res = {}
    for prec in (capi.PRECISION_FP64, capi.PRECISION_FP32_TRACK):
        b = problems.batched_instances()
        with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT, precision=prec,
                         fp32_tol=a.fp32_tol,
                         reserve_photons=int(b.n_photons.sum()) * 3) as e:
            run = BatchedRun(b, e)
            reps = [run.step() for _ in range(steps)]
        res[prec] = reps
    def rate(reps):
        return sum(r["n_transported"] for r in reps) / sum(r["device_s"] for r in reps)
    p64, p32 = res[capi.PRECISION_FP64], res[capi.PRECISION_FP32_TRACK]
    chosen = {"form": "fp32 tracking + fp64 energies and tallies, persistent waves",
              "fp32_tol": a.fp32_tol or 1e-6, "histories_per_s": rate(p32),
              "speedup_vs_plain_fp64": rate(p32) / rate(p64),
              "step1_tally_error_vs_fp64": _tally_error(p64[0]["abs_E"], p32[0]["abs_E"]),
              "replayed_histories": sum(r["n_replayed"] for r in p32)}
    return _common(METRIC, rate(p64), steps, 0,
                   1e3 * sum(r["device_s"] for r in p64) / steps, "f64",
                   {"workload": "batched: 256 slab/cube instances, cells log-uniform in "
                                "[50, 32768], photons log-uniform in [1e4, 1e6], sigma_a "
                                "lognormal (median 10, sigma_log 1), sigma_s U[0,2], T0 "
                                "U[0.01,1] keV, dt 0.01 sh, seeds 0..255, one launch per "
                                "phase for all instances",
                    "form": "plain fp64, persistent waves",
                    "histories": sum(r["n_transported"] for r in p64),
                    "segments": sum(r["n_segments"] for r in p64)},
                   chosen=chosen)


def _bench_adjoint(a, capi):
    """Forward tally and path-replay adjoint on the 32^3 lognormal field,
    1e6 photons, one step."""
    import numpy as np
    from branson_amd import problems
    from branson_amd.adjoint import FixedSource

    prob = problems.adjoint_field_problem()
    with capi.Engine(0, algorithm=capi.ALGO_PERSISTENT) as e:
        src = FixedSource.sample(e, prob, 10**6, seed=2718)
        target = src.absorbed(prob.sigma_a * 1.05, prob.fleck)
        fwd_ms, adj_ms = [], []
        for _ in range(max(a.steps, 1)):
            src._set(prob.sigma_a, prob.fleck)
            ph = src.photons.copy()
            ab, _, st = e.transport_photons(ph)
            fwd_ms.append(st["transport_ms"])
            g = 2.0 * (ab - target)
            src._set(prob.sigma_a, prob.fleck)
            da, df, sa = e.adjoint_photons(src.photons, g)
            adj_ms.append(sa["transport_ms"])
    n = len(src.photons)
    return _common("adjoint histories/s (two passes per history), 32^3 lognormal field",
                   n / (np.mean(adj_ms) * 1e-3), max(a.steps, 1), 0, float(np.mean(adj_ms)),
                   "f64",
                   {"workload": "adjoint: 32^3 cells of 0.05 cm, sigma_a lognormal field "
                                "(median 10, sigma_log 0.7, correlation 4 cells), sigma_s 0.5, "
                                "T 0.5 keV, 1e6 photons, loss sum (abs_E - target)^2",
                    "form": "fp64 path replay", "histories": n},
                   forward_histories_per_s=n / (np.mean(fwd_ms) * 1e-3),
                   adjoint_over_forward_time=float(np.mean(adj_ms) / np.mean(fwd_ms)),
                   grad_norm={"sigma_a": float(np.linalg.norm(da)),
                              "fleck": float(np.linalg.norm(df))})


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-chosen", action="store_true",
                    help="skip the chosen form and the history-form comparison")
    ap.add_argument("--workload", default="cube",
                    choices=["cube", "hohlraum", "batched", "adjoint"],
                    help="cube (default): the throughput line; the others measure the "
                         "30-group hohlraum, the 256-problem batch and the adjoint")
    ap.add_argument("--fp32-tol", type=float, default=0.0,
                    help="FP32 tracking: per-deposit relative error bound (0 = default)")
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

    if a.workload != "cube":
        fn = {"hohlraum": _bench_hohlraum, "batched": _bench_batched,
              "adjoint": _bench_adjoint}[a.workload]
        print(json.dumps(fn(a, capi)), flush=True)
        return

    # streaming-copy probe: the HBM rate this device reaches with a plain
    # copy kernel (k_copy_probe, 1 GiB each way), also the calibration kernel
    # for the FETCH_SIZE / WRITE_SIZE counters of the profiles
    with capi.Engine(0) as e:
        copy_gbps = e.copy_probe(1 << 30, 20)

    plain = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_PERSISTENT,
                      keep_step1=True)
    value = plain["hist"] / plain["dev_s"]
    bytes_launch, kms, achieved = _roofline("fp64", plain)
    # the same steps with each device phase as one HIP graph (the two
    # captures fall inside the timed steps)
    graph = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_PERSISTENT,
                      step_graph=True)

    history = chosen = None
    if not a.no_chosen:
        hi = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_HISTORY)
        history = {"form": "plain fp64, one work-item per history (bit-identical events)",
                   "histories_per_s": hi["hist"] / hi["dev_s"],
                   "segments_per_s": hi["segs"] / hi["dev_s"],
                   "ms_per_step": 1e3 * hi["dev_s"] / a.steps,
                   "transport_kernel_ms_mean": _roofline("fp64", hi)[1]}
        ch = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP32_TRACK,
                       capi.ALGO_PERSISTENT, keep_step1=True, fp32_tol=a.fp32_tol)
        cb, cms, cach = _roofline("fp32", ch)
        chosen = {
            "form": "fp32 tracking + fp64 energies and tallies, persistent waves",
            "histories_per_s": ch["hist"] / ch["dev_s"],
            "segments_per_s": ch["segs"] / ch["dev_s"],
            "ms_per_step": 1e3 * ch["dev_s"] / a.steps,
            "step1_tally_error_vs_fp64": _tally_error(plain["step1"], ch["step1"]),
            "replayed_histories": ch["replayed"],
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
                   "cell_records": "lattice" if plain["lattice"] else "general",
                   "tally": plain["tally"],
                   "histories": plain["hist"], "segments": plain["segs"]},
        "segments_per_s": plain["segs"] / plain["dev_s"],
        # photon counts follow temperatures that carry the tally atomics'
        # round-off, so the graph run may differ by a photon in a rare cell
        "graph_steps": {"histories_per_s": graph["hist"] / graph["dev_s"],
                        "ms_per_step": 1e3 * graph["dev_s"] / a.steps,
                        "histories_segments": [graph["hist"], graph["segs"]]},
        # whole step: host physics, uploads of cells and source energies over
        # PCIe, device phase, tally copy-out and temperature update
        "whole_step_histories_per_s": plain["hist"] / plain["step_s"],
        # whole step with the step physics on the device (no per-step uploads)
        "resident_whole_step_histories_per_s": _run_resident(capi, a.steps, a.warmup),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": _traffic_from_profile(),
                     "kernel": "k_transport_persistent<GlobalTally<false>, double, %s>"
                               % ("true" if plain["lattice"] else "false"),
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
