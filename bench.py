"""Benchmark: replicated IMC transport phase on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                    [--workload headline|cube|batched|adjoint|precision] [--fp32-tol T]

Default workload (headline): the 30-group hohlraum-like configuration -- 64^3
cells of 0.01 cm, a 4-cell wall (sigma_a = 1000/cm, rho = 19) around a
low-density fill (sigma_a = 0.1/cm, sigma_s = 0.01/cm, rho = 0.001), a 0.3
keV source on the z = 0 face, 30 groups, 5e7 photons per step
(branson_amd/decks/hohlraum_64.xml) -- in the plain FP64 form and in the
FP32-tracking form (FP64 energies and tallies) at the per-cell tally error
bound fp32_tol (default 1e-6).  The chosen form is the faster of the two whose
step-1 per-cell absorbed-energy tallies are within 1e-6 (max relative over
cells) of the FP64 form's; value is its histories per second.  The 40^3 cube
throughput configuration (1e7 photons per step) is measured beside it in the
plain FP64 form.

A step is one replicated time step.  Its host part (step physics, upload of
cells and per-cell source energies, tally copy-out, temperature update) runs
untimed; the timed part is the device phase with every input resident in HBM:
source sampling + photon transport (the kernels write the census) + census
order and totals (bh_driver_transport), bracketed by device synchronisation.
Histories = photons transported (new + census).

The JSON line also carries segments/s, the same steps launched as one HIP
graph per step, whole steps with the step physics on the device, the roofline
of the dominant kernel (the transport kernel: algorithmic bytes per launch
over its HIP-event time against 8 TB/s HBM; measured HBM traffic per launch
from the committed rocprofv3 counter summary of the same launches, when that
summary was made with the same --steps and --warmup) and the oracle's OpenMP
CPU step (source + transport + census) timed on this host.
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
              n_photons=N_PHOTONS, n_groups=1, fp32_tol=0.0, step_graph=False,
              census_steps=8):
    """Warm up on one driver, then time `steps` steps on a fresh one.  Photon
    buffers are reserved up front (device memory is set-up, not step work):
    the census grows to ~3.7x the photons per step over five steps.
    step_graph: each device phase is one HIP graph launch (captured on the
    first two graph steps, one graph per parity of the photon stores; the
    captures are inside the timed steps); transport_ms then times the whole
    graph."""
    reserve = n_photons * (1 + min(steps, census_steps))
    kw = dict(precision=precision, algorithm=algorithm, n_groups=n_groups, fp32_tol=fp32_tol,
              step_graph=step_graph)
    if warmup:
        with capi.Driver(deck, t_stop=DT * warmup, **kw) as w:
            w.run()
    out = dict(hist=0, segs=0, flushes=0, dev_s=0.0, step_s=0.0, k_ms=[], launches=[],
               step1=None, replayed=0, lattice=None, tally=None, src_ms=[], cen_ms=[],
               replay_ms=[], allocs=0)
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
            out["src_ms"].append(r["source_ms"])
            out["cen_ms"].append(r["census_ms"])
            out["replay_ms"].append(r["replay_ms"])
            out["allocs"] += r["n_allocs"]
            out["launches"].append((r["transported"], r["n_segments"], r["n_flushes"]))
    return out


def _run_resident(capi, steps, warmup, deck=DECK, n_photons=N_PHOTONS, n_groups=1,
                  census_steps=8):
    """Whole steps with the step physics on the device (driver device_physics):
    wall clock of transport + finish per step, nothing but the step report
    crossing PCIe.  prepare() is untimed: in this mode it uploads the mesh and
    material on the first step (set-up, as the photon store reservation) and
    does nothing after.  Histories per second of whole steps."""
    reserve = n_photons * (1 + min(steps, census_steps))
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


def _traffic_from_profile(path, steps, warmup):
    """Mean HBM bytes per timed transport launch from a committed counter
    summary, when it was made over the same launches: the summary's line
    'bench_args: --steps K --warmup W' must match and its line
    'traffic_bytes_per_launch: [...]' hold one value per timed step.  None
    otherwise."""
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return None
    m = re.search(r"bench_args:\s*--steps\s+(\d+)\s+--warmup\s+(\d+)", text)
    t = re.search(r"traffic_bytes_per_launch:\s*\[([^\]]*)\]", text)
    if not m or not t or (int(m.group(1)), int(m.group(2))) != (steps, warmup):
        return None
    vals = [float(x) for x in t.group(1).split(",") if x.strip()]
    return sum(vals) / len(vals) if len(vals) == steps else None


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


def _host_cores():
    """Physical cores of this host (unique core/socket pairs of lscpu -p),
    or the logical CPU count when lscpu is missing."""
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                             timeout=10).stdout
        pairs = {l.strip() for l in out.splitlines() if l.strip() and not l.startswith("#")}
        if pairs:
            return len(pairs)
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count() or 1


def _cpu_threads():
    """Every physical core, within the CPU share this process is given
    (OMP_NUM_THREADS when set: the GPU boxes allot 16 cores per GPU)."""
    phys = _host_cores()
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(phys, env) if env > 0 else phys), phys


def _cpu_baseline(deck_path, n_groups, n_photons, steps, what):
    """The oracle's OpenMP step (schedule(guided), per-thread tallies; source
    serial as the reference's): source + transport + census of the first
    `steps` steps of the deck, timed on this host."""
    import oracle

    threads, phys = _cpu_threads()
    deck = oracle.Deck.from_xml(deck_path, n_groups=n_groups)
    deck.n_photons = n_photons
    deck.t_stop = deck.t_start + DT * steps
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
            "host_physical_cores": phys,
            "segments_per_s": segs / secs, "transport_only_histories_per_s": hist / t_secs,
            "sample": f"{what}: {hist} histories, {secs:.2f} s of source + transport + census; "
                      f"oracle C restatement, OpenMP guided, {threads} threads "
                      f"(lscpu: {phys} physical cores on this host)"}


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


HOHLRAUM_WORKLOAD = ("hohlraum_64: 64^3 cells of 0.01 cm, 4-cell wall sigma_a=1000 rho=19, "
                     "fill sigma_a=0.1 sigma_s=0.01 rho=0.001, 0.3 keV source on z=0, 30 "
                     "groups, 5e7 photons/step, dt=0.001 sh, seed 4321")
CUBE_WORKLOAD = ("cube_40_throughput: plain gray 3D cube, 40x40x40 cells of 0.05 cm, vacuum "
                 "faces, T=1 keV, emission only, sigma_a=5, sigma_s=1, dt=0.001 sh, 1e7 "
                 "photons/step, seed 1234")
HOHLRAUM_PHOTONS = 50_000_000
TALLY_BAR = 1e-6  # per-cell absorbed-energy error the chosen form must meet
CUBE_TRAFFIC = os.path.join(ROOT, "profiles", "transport_persistent_fp64_cube40_traffic.md")
HOHLRAUM_TRAFFIC = os.path.join(ROOT, "profiles",
                                "transport_persistent_fp64_hohlraum64_traffic.md")


def _mean(x):
    return sum(x) / len(x) if x else 0.0


def _kernel_split(run):
    """Mean device time per step of the source, transport and census phases
    (HIP events on the engine's stream)."""
    return {"source_ms": _mean(run["src_ms"]), "transport_ms": _mean(run["k_ms"]),
            "census_ms": _mean(run["cen_ms"])}


def _bench_headline(a, capi, cube=True):
    """The 30-group hohlraum: plain FP64 against FP32 tracking at fp32_tol;
    the chosen form is the faster one within TALLY_BAR of FP64 per cell.
    Beside it: graph and resident whole steps, the roofline of the transport
    kernel, the cube throughput line and the CPU baseline."""
    with capi.Engine(0) as e:
        copy_gbps = e.copy_probe(1 << 30, 20)
    # the wall absorbs nearly every photon, so the census stays a small
    # fraction of a step's photons: stores for two steps' worth
    kw = dict(deck=HOHLRAUM, n_photons=HOHLRAUM_PHOTONS, n_groups=30, census_steps=1)
    plain = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_PERSISTENT,
                      keep_step1=True, **kw)
    fp32 = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP32_TRACK, capi.ALGO_PERSISTENT,
                     keep_step1=True, fp32_tol=a.fp32_tol, **kw)
    plain_rate = plain["hist"] / plain["dev_s"]
    fp32_rate = fp32["hist"] / fp32["dev_s"]
    err = _tally_error(plain["step1"], fp32["step1"])
    fp32_ok = err["max_rel"] <= TALLY_BAR
    use_fp32 = fp32_ok and fp32_rate > plain_rate
    ch, ch_rate = (fp32, fp32_rate) if use_fp32 else (plain, plain_rate)
    fp32_block = {
        "form": "fp32 tracking + fp64 energies and tallies, persistent waves",
        "fp32_tol": a.fp32_tol or 1e-6,
        "histories_per_s": fp32_rate,
        "segments_per_s": fp32["segs"] / fp32["dev_s"],
        "ms_per_step": 1e3 * fp32["dev_s"] / a.steps,
        "speedup_vs_plain_fp64": fp32_rate / plain_rate,
        "step1_tally_error_vs_fp64": err,
        "meets_tally_bar": fp32_ok,
        "replayed_histories": fp32["replayed"],
        "replay_fraction": fp32["replayed"] / max(fp32["hist"], 1),
        "device_allocations_in_timed_steps": fp32["allocs"],
        "kernel_ms_per_step": dict(_kernel_split(fp32),
                                   fp32_pass_ms=_mean(fp32["k_ms"]) - _mean(fp32["replay_ms"]),
                                   fp64_replay_ms=_mean(fp32["replay_ms"])),
    }
    graph = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_PERSISTENT,
                      step_graph=True, **kw)
    event = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_EVENT, **kw)
    resident = None if a.no_resident else _run_resident(
        capi, a.steps, a.warmup, deck=HOHLRAUM, n_photons=HOHLRAUM_PHOTONS, n_groups=30,
        census_steps=1)
    bytes_launch, kms, achieved = _roofline("fp64", plain)
    line = _common(
        METRIC, ch_rate, a.steps, a.warmup, 1e3 * ch["dev_s"] / a.steps, "f64",
        {"workload": HOHLRAUM_WORKLOAD,
         "form": ("fp32 tracking + fp64 energies and tallies" if use_fp32 else
                  "plain fp64 (reference RNG streams and arithmetic)") + ", persistent waves",
         "chosen_by": f"faster of plain fp64 and fp32 tracking with step-1 max per-cell "
                      f"abs_E error <= {TALLY_BAR:g} against fp64",
         "cell_records": "lattice" if ch["lattice"] else "general", "tally": ch["tally"],
         "histories": ch["hist"], "segments": ch["segs"]},
        segments_per_s=ch["segs"] / ch["dev_s"],
        plain_fp64={"histories_per_s": plain_rate,
                    "segments_per_s": plain["segs"] / plain["dev_s"],
                    "ms_per_step": 1e3 * plain["dev_s"] / a.steps,
                    "kernel_ms_per_step": _kernel_split(plain),
                    "device_allocations_in_timed_steps": plain["allocs"]},
        fp32_tracking=fp32_block,
        graph_steps={"histories_per_s": graph["hist"] / graph["dev_s"],
                     "ms_per_step": 1e3 * graph["dev_s"] / a.steps},
        event_form={"form": "plain fp64, event-split passes over the compacted list of "
                            "live photons (same events)",
                    "histories_per_s": event["hist"] / event["dev_s"],
                    "ms_per_step": 1e3 * event["dev_s"] / a.steps,
                    "kernel_ms_per_step": _kernel_split(event)},
        whole_step_histories_per_s=plain["hist"] / plain["step_s"],
        resident_whole_step_histories_per_s=resident,
        roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBPS,
                  "traffic": _traffic_from_profile(HOHLRAUM_TRAFFIC, a.steps, a.warmup),
                  "kernel": "k_transport_persistent<%s, double, %s> (plain fp64, hohlraum)"
                            % ("HashTally" if plain["tally"] == TALLY_FORMS[4] else "tally",
                               "true" if plain["lattice"] else "false"),
                  "kernel_ms_mean": kms, "algorithmic_bytes_per_launch": bytes_launch,
                  "launches": len(plain["k_ms"]),
                  "launch_histories_segments_flushes": plain["launches"],
                  "launch_kernel_ms": plain["k_ms"], "copy_probe_GBps": copy_gbps})
    if cube:
        c = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_PERSISTENT)
        cb, cms, cach = _roofline("fp64", c)
        ce = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_EVENT)
        line["cube_throughput"] = {
            "workload": CUBE_WORKLOAD, "form": "plain fp64, persistent waves",
            "histories_per_s": c["hist"] / c["dev_s"], "segments_per_s": c["segs"] / c["dev_s"],
            "ms_per_step": 1e3 * c["dev_s"] / a.steps, "tally": c["tally"],
            "histories": c["hist"], "segments": c["segs"],
            "kernel_ms_per_step": _kernel_split(c),
            "transport_kernel_roofline": {"achieved_GBps": cach, "frac": cach / HBM_PEAK_GBPS,
                                          "kernel_ms_mean": cms,
                                          "algorithmic_bytes_per_launch": cb,
                                          "launch_histories_segments_flushes": c["launches"],
                                          "launch_kernel_ms": c["k_ms"]},
            "event_form": {"histories_per_s": ce["hist"] / ce["dev_s"],
                           "ms_per_step": 1e3 * ce["dev_s"] / a.steps,
                           "kernel_ms_per_step": _kernel_split(ce)}}
    line["cpu_baseline"] = None if a.no_cpu else _cpu_baseline(
        HOHLRAUM, 30, HOHLRAUM_PHOTONS, 1,
        "step 1 of the same 30-group hohlraum deck at 5e7 photons/step")
    return line


def _bench_batched(a, capi):
    """256 independent slab/cube problems in one engine, 3 steps, each step
    one bh_step_resident_instances call (physics on the device; steps after
    the first one graph launch each)."""
    from branson_amd import problems
    from branson_amd.batched import BatchedRun

    steps = min(a.steps, 3)
    res, step1 = {}, {}
    for prec, algo in ((capi.PRECISION_FP64, capi.ALGO_PERSISTENT),
                       (capi.PRECISION_FP32_TRACK, capi.ALGO_PERSISTENT),
                       (capi.PRECISION_FP64, capi.ALGO_EVENT)):
        b = problems.batched_instances()
        key = prec if algo == capi.ALGO_PERSISTENT else "event"
        with capi.Engine(0, algorithm=algo, precision=prec,
                         fp32_tol=a.fp32_tol,
                         reserve_photons=int(b.n_photons.sum()) * 3) as e:
            run = BatchedRun(b, e)
            reps = []
            for k in range(steps):
                reps.append(run.step())
                if k == 0:
                    step1[key] = run.tallies()[0]
        res[key] = reps

    def rate(reps):
        return sum(r["n_transported"] for r in reps) / sum(r["device_s"] for r in reps)
    p64, p32 = res[capi.PRECISION_FP64], res[capi.PRECISION_FP32_TRACK]
    chosen = {"form": "fp32 tracking + fp64 energies and tallies, persistent waves",
              "fp32_tol": a.fp32_tol or 1e-6, "histories_per_s": rate(p32),
              "speedup_vs_plain_fp64": rate(p32) / rate(p64),
              "step1_tally_error_vs_fp64": _tally_error(step1[capi.PRECISION_FP64],
                                                        step1[capi.PRECISION_FP32_TRACK]),
              "replayed_histories": sum(r["n_replayed"] for r in p32)}
    return _common(METRIC, rate(p64), steps, 0,
                   1e3 * sum(r["device_s"] for r in p64) / steps, "f64",
                   {"workload": "batched: 256 slab/cube instances, cells log-uniform in "
                                "[50, 32768], photons log-uniform in [1e4, 1e6], sigma_a "
                                "lognormal (median 10, sigma_log 1), sigma_s U[0,2], T0 "
                                "U[0.01,1] keV, dt 0.01 sh, seeds 0..255, one launch per "
                                "phase for all instances",
                    "form": "plain fp64, persistent waves, whole steps (physics on the "
                            "device, steps after the first one graph launch)",
                    "histories": sum(r["n_transported"] for r in p64),
                    "segments": sum(r["n_segments"] for r in p64)},
                   max_instance_balance=max(r["max_instance_balance"] for r in p64),
                   event_form={"histories_per_s": rate(res["event"])},
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


def _bench_precision(a, capi):
    """FP32 tracking against plain FP64 on the 30-group hohlraum at 1e7
    photons per step (2 steps each): per-step transport and FP64-replay time,
    replayed histories and the step-1 per-cell tally error, at fp32_tol 1e-6
    and 1e-2.  The command the VALU-counter profile of the two forms runs."""
    forms = [("plain fp64", capi.PRECISION_FP64, 0.0),
             ("fp32 tracking, fp32_tol 1e-6", capi.PRECISION_FP32_TRACK, 1e-6),
             ("fp32 tracking, fp32_tol 1e-2", capi.PRECISION_FP32_TRACK, 1e-2)]
    out, ref = [], None
    for name, prec, tol in forms:
        with capi.Driver(HOHLRAUM, n_groups=30, n_photons=10_000_000, t_stop=2 * DT,
                         algorithm=capi.ALGO_PERSISTENT, precision=prec, fp32_tol=tol,
                         reserve_photons=20_000_000) as d:
            reps, tal = [], None
            while not d.finished():
                d.prepare()
                d.transport()
                reps.append(d.finish())
                if tal is None:
                    tal = d.tallies()[0]
        if ref is None:
            ref = tal
        out.append({"form": name, "histories": sum(r["transported"] for r in reps),
                    "replayed": sum(r["n_replayed"] for r in reps),
                    "transport_ms_per_step": _mean([r["transport_ms"] for r in reps]),
                    "fp64_replay_ms_per_step": _mean([r["replay_ms"] for r in reps]),
                    "step1_tally_error_vs_fp64": _tally_error(ref, tal)})
    f64 = out[0]
    ms = f64["transport_ms_per_step"]
    return _common("hohlraum histories/s of the transport kernel, plain FP64 beside FP32 "
                   "tracking", f64["histories"] / 2 / (ms * 1e-3), 2, 0, ms, "f64",
                   {"workload": "precision study: hohlraum_64, 30 groups, 1e7 photons/step, "
                                "2 steps per form", "form": "plain fp64"},
                   forms=out)


def _bench_cube(a, capi):
    """The 40^3 cube throughput configuration alone, plain FP64, with the
    one-work-item-per-history form and FP32 tracking beside it."""
    with capi.Engine(0) as e:
        copy_gbps = e.copy_probe(1 << 30, 20)
    plain = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_PERSISTENT,
                      keep_step1=True)
    value = plain["hist"] / plain["dev_s"]
    bytes_launch, kms, achieved = _roofline("fp64", plain)
    hi = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP64, capi.ALGO_HISTORY)
    ch = _run_form(capi, a.steps, a.warmup, capi.PRECISION_FP32_TRACK, capi.ALGO_PERSISTENT,
                   keep_step1=True, fp32_tol=a.fp32_tol)
    return _common(
        METRIC, value, a.steps, a.warmup, 1e3 * plain["dev_s"] / a.steps, "f64",
        {"workload": CUBE_WORKLOAD, "form": "plain fp64, persistent waves",
         "cell_records": "lattice" if plain["lattice"] else "general", "tally": plain["tally"],
         "histories": plain["hist"], "segments": plain["segs"]},
        segments_per_s=plain["segs"] / plain["dev_s"],
        kernel_ms_per_step=_kernel_split(plain),
        history_form={"form": "plain fp64, one work-item per history (bit-identical events)",
                      "histories_per_s": hi["hist"] / hi["dev_s"],
                      "transport_kernel_ms_mean": _roofline("fp64", hi)[1]},
        fp32_tracking={"histories_per_s": ch["hist"] / ch["dev_s"],
                       "step1_tally_error_vs_fp64": _tally_error(plain["step1"], ch["step1"]),
                       "replay_fraction": ch["replayed"] / max(ch["hist"], 1)},
        roofline={"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK_GBPS,
                  "traffic": _traffic_from_profile(CUBE_TRAFFIC, a.steps, a.warmup),
                  "kernel": "k_transport_persistent<GlobalTally<false>, double, 1> "
                            "(plain fp64, cube)",
                  "kernel_ms_mean": kms, "algorithmic_bytes_per_launch": bytes_launch,
                  "launches": len(plain["k_ms"]),
                  "launch_histories_segments_flushes": plain["launches"],
                  "launch_kernel_ms": plain["k_ms"], "copy_probe_GBps": copy_gbps},
        cpu_baseline=None if a.no_cpu else _cpu_baseline(
            DECK, 1, N_PHOTONS, 2, "steps 1-2 of the same 40^3 cube deck at 1e7 photons/step"))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-resident", action="store_true",
                    help="skip the resident whole-step leg (the other legs' launches are "
                         "unchanged; profiling runs use it)")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "cube", "batched", "adjoint", "precision"],
                    help="headline (default): the 30-group hohlraum, plain FP64 against "
                         "FP32 tracking, with the cube throughput beside it; cube: the cube "
                         "alone; batched: the 256-problem batch; adjoint: the adjoint; "
                         "precision: FP32 tracking against FP64 on the hohlraum")
    ap.add_argument("--fp32-tol", type=float, default=0.0,
                    help="FP32 tracking: per-deposit relative error bound (0 = default 1e-6)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: every run "
              f"here has one GPU (replicated mode would give each rank its own share of "
              f"the photons and all-reduce the tallies)", file=sys.stderr)
        sys.exit(2)
    if a.steps < 1:
        sys.exit("--steps must be >= 1")

    from branson_amd import capi

    fn = {"headline": _bench_headline, "cube": _bench_cube, "batched": _bench_batched,
          "adjoint": _bench_adjoint, "precision": _bench_precision}[a.workload]
    print(json.dumps(fn(a, capi)), flush=True)


if __name__ == "__main__":
    main()
