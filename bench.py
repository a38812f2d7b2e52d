#!/usr/bin/env python3
"""FDAS hot-path benchmark on one MI355X.

Metric: DM-trials/sec at N = 2^21 bins, 85 templates x 421 taps, 8 harmonics
(4,096 candidates per harmonic).  A trial is one dedispersed spectrum through
the whole device path: fused FT convolution -> harmonic summing + thresholding
+ compaction -> per-plane selection -> candidate lists in host memory.

The timed form (``value``) is the batched engine on the batched-uniform
workload: 64 trials of N = 2^21 (seeds 1000..1063, 0-2 drifting tones each),
spectra resident in HBM before the timed region, one step = one batch of 64
trials.  Two plans alternate so that one batch's tail overlaps the next
batch's head; inside a plan, groups of trials pipeline FT convolution on one
stream against harmonic summing on another.  The engine's autotuner fixes
the tile, rows per item and rows per summing workgroup; the shape is recorded
in ``config``.

Also reported in the same JSON line:
  * plain_gpu_form (primary plain form): one trial per launch on the same
    resident spectra at the same tile: device-to-device input copy, FT
    convolution, summing, selection, list read-back, stream synchronised per
    trial; p50 / p99 latency;
  * plain_gpu_form_h2d: the same with the 16.8 MB host spectrum uploaded per
    trial;
  * single_trial_form: the drop-in class FDAS path, two trials in flight (A/B);
  * graph_form: the same batches replayed as HIP graphs;
  * uneven_form: 48 trials of N in {2^19..2^22}, banks of 85 / 169 templates,
    8 or 16 harmonics, in one plan;
  * served_form: open-loop Poisson request streams through the serving loop,
    p50 / p99 latency and throughput beside the plain one-request-per-launch
    form on the same arrivals, and the sustained capacity of both;
  * power_form (only with --forms ...,power): board power and shader clock
    while the FT convolution alone, then the summing alone, runs;
  * roofline of the dominant kernel (HIP-event spans on its own stream over
    the timed region) and the CPU baseline (the oracle's C restatement timed
    on this host: all threads and one thread, at N = 2^21 and N = 32,768).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch (rocprofv3 FETCH_SIZE x2.000 + WRITE_SIZE x1.000, calibrated
# on a copy kernel) for the batched workload at N=2^21, T=85, tile 1024, 8 trials
# per launch; profiles/batched_n2p21_t85_h8_tile1024_traffic.txt.
MEASURED_TRAFFIC = {("ftconv1w_kernel", 8): 6869.7e6, ("hsum_kernel<8,true>", 8): 9302.7e6}
# VALU wave-instructions per launch (rocprofv3 SQ_INSTS_VALU) at the same shape;
# profiles/batched_n2p21_t85_h8_tile1024_instructions.txt.  Issue rate: 256 CUs x
# 4 SIMDs, one wave64 VALU instruction per 4 cycles per SIMD at 2.4 GHz.
MEASURED_VALU = {("ftconv1w_kernel", 8): 1050558056, ("hsum_kernel<8,true>", 8): 510989349}
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4
# Shader clock the kernel runs at (amd-smi median while it runs alone, bench.py --forms batched,power);
# profiles/power_clock_n2p21_t85_tile1024.txt.  The FT convolution holds the board at its power limit.
MEASURED_CLOCK_GHZ = {"ftconv1w_kernel": 2.112, "hsum_kernel<8,true>": 2.385}
FP32_PEAK_TFLOPS = 157.3
N_METRIC = 1 << 21
T_METRIC = 85
H_METRIC = 8
CAP = 4096


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ftc_bytes(n_bins, n_templates, tile, trials=1):
    """Algorithmic HBM bytes of one FT-convolution launch over `trials` trials:
    spectrum read + FOP write per trial, template spectra once."""
    return trials * (8 * n_bins + 4 * n_templates * n_bins) + 8 * n_templates * tile


def hsum_bytes(n_bins, n_templates, n_cands, trials=1):
    """Algorithmic HBM bytes of one harmonic-summing launch: FOP read once + candidates out, per trial."""
    return trials * (4 * n_templates * n_bins) + 8 * n_cands


def path_bytes(n_bins, n_templates):
    """Algorithmic bytes of one trial through the whole path: spectrum in, FOP written and read back
    (the template spectra are resident and candidates are a few hundred KB)."""
    return 8 * n_bins + 8 * n_templates * n_bins


def ftc_kernel_name(tile):
    return "ftconv1w_kernel" if tile == 1024 else f"ftconv_kernel<{tile}>"


def ftc_flops(shape, tile):
    """FP32 flops of one trial's FT convolution from the transforms the kernel runs: 5 S log2 S per
    transform (forward per work item, inverse per tile and template row), 6 S per spectrum product
    and 3 per stored |x|^2."""
    per = 5.0 * tile * np.log2(tile)
    return ((shape["forward_transforms"] + shape["inverse_transforms"]) * per
            + shape["inverse_transforms"] * 6.0 * tile + shape["stored_outputs"] * 3.0)


def device_thresholds(x, n_harmonics, sigma=6.0):
    """T_k = (k + sigma sqrt(k)) * mean|X|^2 (the noise_thresholds rule), mean taken on the device in FP64."""
    import torch
    p = float((x.real.double() ** 2 + x.imag.double() ** 2).mean().item())
    k = np.arange(1, n_harmonics + 1, dtype=np.float64)
    return ((k + sigma * np.sqrt(k)) * p).astype(np.float32)


def valu_issue(kernel, trials_per_launch, span_ms):
    """VALU issue time of one launch at 2.4 GHz and at the clock the kernel was measured running at,
    each as a fraction of the measured span."""
    n = MEASURED_VALU[(kernel, trials_per_launch)]
    at_peak = n / VALU_ISSUE_PER_S
    out = {"instructions": n, "issue_ms": round(at_peak * 1e3, 4), "frac_of_span": round(at_peak / (span_ms * 1e-3), 3)}
    ghz = MEASURED_CLOCK_GHZ.get(kernel)
    if ghz:
        at_clk = at_peak * 2.4 / ghz
        out.update({"clock_ghz": ghz, "issue_ms_at_clock": round(at_clk * 1e3, 4),
                    "frac_of_span_at_clock": round(at_clk / (span_ms * 1e-3), 3)})
    return out


def pct(a, q):
    return round(float(np.percentile(np.asarray(a), q)), 3) if len(a) else None


def _cpu_trial_times(x, bank, thr, cap, threads, budget_s, min_reps, max_reps):
    from oracle import cpu
    times = []
    t_start = time.perf_counter()
    while len(times) < min_reps or (time.perf_counter() - t_start < budget_s and len(times) < max_reps):
        t0 = time.perf_counter()
        fop = cpu.fop(x, bank, threads=threads)
        cpu.hsum(fop, thr, len(thr), cap, threads=threads)
        times.append(time.perf_counter() - t0)
        del fop
    return float(np.median(times)), len(times)


def cpu_baseline(x, bank, thr, cap, budget_s=12.0):
    """The oracle's C restatement of ft_conv.py + hsum.py (FP32, OpenMP, own radix-4 FFT, AVX2/FMA
    build where the host has them) on whole trials: N = 2^21 on all host threads (the baseline
    value) and on one thread, and the emulator-size N = 32,768 search on all threads."""
    from oracle import cpu
    from fdas_amd import synth
    n_threads = cpu.threads()
    cpu.fop(x[:4096], bank)  # warm-up (page-in, thread pool)
    med, reps = _cpu_trial_times(x, bank, thr, cap, n_threads, budget_s, 3, 50)
    med1, reps1 = _cpu_trial_times(x, bank, thr, cap, 1, 0.0, 1, 1)
    w1 = synth.emulator_workload()
    x1 = w1.spectrum()
    thr1 = synth.noise_thresholds(x1, 8)
    _cpu_trial_times(x1, bank, thr1, 64, n_threads, 0.0, 1, 1)
    med_e, reps_e = _cpu_trial_times(x1, bank, thr1, 64, n_threads, 1.0, 5, 200)
    return {"value": round(1.0 / med, 4), "unit": "DM-trials/s", "cores": n_threads, "kind": "port",
            "sample": f"{reps} full trials (N=2^21, 85 templates x 421 taps, 8 harmonics, 4096/plane) through "
                      f"oracle/fdas_cpu.c ({cpu.isa()} build: FP32 overlap-save radix-4 FFT convolution + "
                      f"harmonic summing + selection), median {med:.3f} s per trial on {n_threads} OpenMP threads",
            "ms_per_trial": round(med * 1000.0, 2), "isa": cpu.isa(),
            "one_thread": {"value": round(1.0 / med1, 4), "ms_per_trial": round(med1 * 1000.0, 1),
                           "sample": f"{reps1} full N=2^21 trial on 1 thread"},
            "emulator_size": {"value": round(1.0 / med_e, 2), "ms_per_trial": round(med_e * 1000.0, 3),
                              "cores": n_threads,
                              "sample": f"median of {reps_e} N=32,768 trials (85 templates, 8 harmonics, 64/plane)"}}


# --------------------------------------------------------------------------- single-trial drop-in path

def single_trial_forms(args, x, bank, thr, forms, tile, device_trials):
    """The drop-in class FDAS path at the batched form's tile.  device_trials: [(spectrum tensor on the
    device, thresholds)] of the batched workload, the inputs of the device-resident plain form."""
    from fdas_amd.host import FDAS
    f = FDAS(log=lambda lvl, msg: (log(msg) if lvl >= 1 else None))
    assert f.initialise_accelerator(N_METRIC, tile=tile, n_templates=T_METRIC, max_taps=421,
                                    n_harmonics=H_METRIC, cands_per_harmonic=CAP), f.last_error
    for ab in (FDAS.A, FDAS.B):
        assert f.upload_template_taps(bank, ab), f.last_error
        assert f.perform_input_tiling(x, ab), f.last_error
    n_list = f.get_candidate_list_sz()
    loc = {ab: np.zeros(n_list, np.uint32) for ab in (FDAS.A, FDAS.B)}
    pwr = {ab: np.zeros(n_list, np.float32) for ab in (FDAS.A, FDAS.B)}
    part = FDAS.AllAccelerations

    def run(n_steps, depth, record=None):
        in_flight = []
        for i in range(n_steps):
            ab = i % 2
            if len(in_flight) == depth:
                prev = in_flight.pop(0)
                assert f.wait(prev), f.last_error
                if record is not None:
                    record.append(f.stage_times(prev))
            assert f.launch_resident(thr, loc[ab], pwr[ab], part, ab), f.last_error
            in_flight.append(ab)
        for prev in in_flight:
            assert f.wait(prev), f.last_error
            if record is not None:
                record.append(f.stage_times(prev))

    out = {}
    if "single" in forms:
        run(4, 2)
        rec = []
        t0 = time.perf_counter()
        run(30, 2, rec)
        ms = (time.perf_counter() - t0) * 1000.0 / 30
        st = np.array(rec)
        out["single_trial_form"] = {
            "value": round(1000.0 / ms, 2), "unit": "DM-trials/s", "ms_per_trial": round(ms, 4), "tile": tile,
            "ftconv_ms": round(float(st[:, 1].mean()), 4), "hsum_ms": round(float(st[:, 2].mean()), 4),
            "select_ms": round(float(st[:, 3].mean()), 4),
            "form": "class FDAS drop-in path, launch_resident on buffer sets A/B, two trials in flight"}
        log(f"single-trial: {out['single_trial_form']}")
    if "plain" in forms:
        import torch

        def plain_stats(lat, total_s, form):
            lat = np.array(lat)
            return {"value": round(float(len(lat) / total_s), 2), "unit": "DM-trials/s",
                    "ms_per_trial": round(float(lat.mean()), 4), "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                    "trials": int(len(lat)), "tile": tile, "form": form}

        # device-resident: the batched workload's spectra, one trial per launch, synchronised per trial
        n_dev = len(device_trials)
        for i in range(4):
            xd, th = device_trials[i % n_dev]
            assert f.set_input_device(xd.data_ptr(), FDAS.A), f.last_error
            assert f.launch_resident(th, loc[FDAS.A], pwr[FDAS.A], part, FDAS.A), f.last_error
            assert f.wait(FDAS.A), f.last_error
        torch.cuda.synchronize()
        lat = []
        t_all = time.perf_counter()
        for i in range(args.plain_trials):
            xd, th = device_trials[i % n_dev]
            t = time.perf_counter()
            assert f.set_input_device(xd.data_ptr(), FDAS.A), f.last_error
            assert f.launch_resident(th, loc[FDAS.A], pwr[FDAS.A], part, FDAS.A), f.last_error
            assert f.wait(FDAS.A), f.last_error
            lat.append((time.perf_counter() - t) * 1000.0)
        out["plain_gpu_form"] = plain_stats(
            lat, time.perf_counter() - t_all,
            "device-resident spectrum: device-to-device input copy + FT conv + HSUM + selection + list read-back "
            "+ sync, one trial per launch, the batched workload's trials")
        log(f"plain (device-resident): {out['plain_gpu_form']}")
        # host upload per trial (the reference's launch(): host spectrum in, lists out)
        for _ in range(2):
            assert f.launch(x, thr, loc[FDAS.A], pwr[FDAS.A], part, FDAS.A), f.last_error
            assert f.wait(FDAS.A), f.last_error
        lat = []
        t_all = time.perf_counter()
        for _ in range(args.plain_trials):
            t = time.perf_counter()
            assert f.launch(x, thr, loc[FDAS.A], pwr[FDAS.A], part, FDAS.A), f.last_error
            assert f.wait(FDAS.A), f.last_error
            lat.append((time.perf_counter() - t) * 1000.0)
        out["plain_gpu_form_h2d"] = plain_stats(
            lat, time.perf_counter() - t_all,
            "host spectrum upload (16.8 MB, pageable) + FT conv + HSUM + selection + read-back + sync, "
            "one trial per launch")
        log(f"plain (host upload): {out['plain_gpu_form_h2d']}")
    f.close()
    return out


def parse_smi_metric(txt):
    """(board power W, mean shader clock MHz, min, max) from the text of `amd-smi metric -p -c`, or
    None when either field is missing."""
    import re
    pw = re.search(r"SOCKET_POWER:\s*(\d+)", txt)
    clk = [int(c) for c in re.findall(r"GFX_\d+:\s*\n\s*CLK:\s*(\d+)", txt)]
    if pw is None or not clk:
        return None
    return float(pw.group(1)), float(np.mean(clk)), min(clk), max(clk)


def _smi_sample(smi):
    """One `amd-smi metric -g 0 -p -c` call, parsed; None if it fails."""
    import subprocess
    try:
        txt = subprocess.run([smi, "metric", "-g", "0", "-p", "-c"], capture_output=True, text=True,
                             timeout=5).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    return parse_smi_metric(txt)


def power_form(x, bank, tile, seconds, batched_plan=None):
    """Board power and shader clock while the FT convolution alone, then the harmonic summing alone,
    runs back to back on one N = 2^21 trial through the class FDAS path, and then, given a plan of the
    batched form, while that plan runs batch after batch (`amd-smi metric -p -c` sampled about every
    0.3 s; medians over each phase after its first 2 s).  The summing phase uses thresholds that no
    bin passes, so that selection has nothing to sort."""
    import shutil

    import torch
    from fdas_amd.host import FDAS
    smi = shutil.which("amd-smi")
    if smi is None:
        return {"error": "amd-smi not found"}
    f = FDAS(log=lambda lvl, msg: (log(msg) if lvl >= 1 else None))
    assert f.initialise_accelerator(N_METRIC, tile=tile, n_templates=T_METRIC, max_taps=421,
                                    n_harmonics=H_METRIC, cands_per_harmonic=CAP), f.last_error
    assert f.upload_template_taps(bank, FDAS.A), f.last_error
    assert f.perform_input_tiling(x, FDAS.A), f.last_error
    none_pass = np.full(H_METRIC, np.finfo(np.float32).max, np.float32)
    samples, stop = [], threading.Event()

    def sampler():
        while not stop.is_set():
            t = time.time()
            s = _smi_sample(smi)
            if s is not None:
                samples.append((t,) + s)
            stop.wait(0.25)

    th = threading.Thread(target=sampler, daemon=True)
    th.start()
    time.sleep(2.0)
    phases = {"idle": (time.time() - 1.5, time.time(), 0)}
    def batch():
        batched_plan.launch()
        batched_plan.wait(results=False)
        return True

    steps = [("ftconv", lambda: f.perform_ft_convolution(FDAS.AllAccelerations, FDAS.A), 100),
             ("hsum", lambda: f.perform_harmonic_summing(none_pass, FDAS.AllAccelerations, FDAS.A), 100)]
    if batched_plan is not None:
        steps.append(("batched", batch, 4))
    for name, fn, per_sync in steps:
        assert fn(), f.last_error
        torch.cuda.synchronize()
        t0, k = time.time(), 0
        while time.time() - t0 < seconds:
            for _ in range(per_sync):
                assert fn(), f.last_error
            torch.cuda.synchronize()
            k += per_sync
        phases[name] = (t0, time.time(), k)
        log(f"power: {name} phase done, {k} launches")
        time.sleep(1.5)
    stop.set()
    th.join()
    f.close()
    out = {"sampler": "amd-smi metric -g 0 -p -c", "n_bins": N_METRIC, "tile": tile}
    for name, (t0, t1, k) in phases.items():
        lo = t0 if name == "idle" else t0 + 2.0
        sel = [s for s in samples if lo <= s[0] <= t1]
        rec = {"samples": len(sel)}
        if k:
            rec["launches"] = k
            rec["ms_per_launch"] = round((t1 - t0) * 1e3 / k, 4)
            if name == "batched":
                rec["trials_per_s"] = round(k * batched_plan.n / (t1 - t0), 1)
        if sel:
            rec["power_w"] = float(np.median([s[1] for s in sel]))
            rec["gfx_clk_mhz"] = float(np.median([s[2] for s in sel]))
            rec["gfx_clk_mhz_range"] = [min(s[3] for s in sel), max(s[4] for s in sel)]
        out[name] = rec
    return out


# --------------------------------------------------------------------------- batched engine

def make_engine(args, **kw):
    from fdas_amd.engine import Engine
    return Engine(tile=args.tile, rows_per_item=args.rows_per_item, hsum_rows=args.hsum_rows,
                  cands_per_harmonic=CAP, group_trials=args.group_trials,
                  ftc_workgroups_per_cu=args.ftc_wg_per_cu, schedule=args.schedule,
                  log=lambda lvl, msg: (log(msg) if (lvl >= 1 or args.verbose) else None), **kw)


def run_batches(plans, n, on_done=None):
    """n batches, alternating over the plans, at most len(plans) in flight."""
    in_flight = []
    for i in range(n):
        p = plans[i % len(plans)]
        if len(in_flight) == len(plans):
            q = in_flight.pop(0)
            q.wait(results=False)
            if on_done:
                on_done(q)
        p.launch()
        in_flight.append(p)
    for q in in_flight:
        q.wait(results=False)
        if on_done:
            on_done(q)


def batch_workload(args):
    """The batched-uniform trials on the device: [(spectrum tensor, thresholds)], seeds 1000.."""
    import torch
    from fdas_amd import synth
    log(f"generating {args.batch} trials of N=2^21 on the device")
    out = []
    for i in range(args.batch):
        x = synth.device_spectrum(synth.random_trial(1000 + i, N_METRIC))
        out.append((x, device_thresholds(x, H_METRIC)))
    torch.cuda.synchronize()
    return out


def batched_form(args, eng, bank85, workload):
    import torch
    from fdas_amd.engine import Trial
    trials = [Trial(x, bank85, thr, 2, tag=1000 + i) for i, (x, thr) in enumerate(workload)]
    tile = eng.prepare(N_METRIC, bank85, 2)
    shape = eng.shape_info(N_METRIC, bank85, 2)
    log(f"batched: shape {shape}")
    out = {"tile": tile, "shape": shape}
    def timed(graph):
        """(ms per batch, mean FT-conv span, mean HSUM span per group) over args.steps batches."""
        spans = {"ftc": [], "hsum": [], "fused": []}

        def collect(p):
            info = p.info()
            n_unfused = info["n_groups"] - info["n_fused"]
            if info["n_fused"]:
                spans["fused"].append(info["fused_ms"] / info["n_fused"])
            if n_unfused:
                spans["ftc"].append(info["ftc_ms"] / n_unfused)
                spans["hsum"].append(info["hsum_ms"] / n_unfused)

        plans = [eng.plan(trials, graph=graph) for _ in range(2)]
        out["groups_per_batch"] = plans[0].info()["n_groups"]
        out["state_gb_per_plan"] = round(plans[0].info()["region_bytes"] / 1e9, 2)
        run_batches(plans, args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run_batches(plans, args.steps, collect)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / args.steps
        # Results of one more batch, for the record: every trial completed within the spill capacity.
        plans[0].launch()
        res = plans[0].wait(raise_overflow=False)
        out["statuses_ok"] = all(r.status == 0 for r in res)
        out["candidates_per_trial_mean"] = float(np.mean([int(np.sum(r.n_kept)) for r in res]))
        # Batch latency: one batch in flight, from launch to every list in host memory.
        lat = []
        for _ in range(max(5, args.steps)):
            t0 = time.perf_counter()
            plans[0].launch()
            plans[0].wait(results=False)
            lat.append((time.perf_counter() - t0) * 1000.0)
        out["lagged"] = bool(plans[0].info()["lagged"])
        for p in plans:
            p.close()
        mean = lambda v: float(np.mean(v)) if v else 0.0
        return ms, {k: mean(v) for k, v in spans.items()}, lat

    log(f"batched: timing {args.steps} batches of {args.batch} (eager launches)")
    ms, sp, lat = timed(False)
    out["eager"] = {"value": round(args.batch * 1000.0 / ms, 2), "ms_per_step": round(ms, 4),
                    "ftc_span_ms": round(sp["ftc"], 4), "hsum_span_ms": round(sp["hsum"], 4),
                    "fused_span_ms": round(sp["fused"], 4),
                    "batch_latency_p50_ms": pct(lat, 50), "batch_latency_p99_ms": pct(lat, 99)}
    log(f"batched eager: {out['eager']}")
    if not args.no_graph:
        # The same launches captured once per plan and replayed; kernel spans are
        # not recorded inside the graph, so the headline stays the eager form.
        log(f"batched: timing {args.steps} batches of {args.batch} (HIP graph replays)")
        gms, _, glat = timed(True)
        out["graph"] = {"value": round(args.batch * 1000.0 / gms, 2), "ms_per_step": round(gms, 4),
                        "batch_latency_p50_ms": pct(glat, 50), "batch_latency_p99_ms": pct(glat, 99)}
        log(f"batched graph: {out['graph']}")
    form = "eager"
    sel = out[form]
    out["form"] = form
    out["ms_per_step"] = sel["ms_per_step"]
    out["value"] = args.batch * 1000.0 / sel["ms_per_step"]
    out["ftc_span_ms"] = sel["ftc_span_ms"]
    out["hsum_span_ms"] = sel["hsum_span_ms"]
    out["fused_span_ms"] = sel["fused_span_ms"]
    out["batch_latency_p50_ms"] = sel["batch_latency_p50_ms"]
    out["batch_latency_p99_ms"] = sel["batch_latency_p99_ms"]
    return out


def uneven_form(args, eng, banks):
    import torch
    from fdas_amd import synth
    from fdas_amd.engine import Trial
    trials, bins = [], 0
    sizes = [1 << 19, 1 << 20, 1 << 21, 1 << 22]
    for i in range(48):
        r = np.random.default_rng(2000 + i)
        n = int(r.choice(sizes))
        T = int(r.choice([85, 169]))
        H = 16 if i % 4 == 0 else 8
        w = synth.random_trial(2000 + i, n, z_max=84.0 if T == 85 else 168.0)
        x = synth.device_spectrum(w)
        trials.append(Trial(x, banks[T], device_thresholds(x, H), 2, tag=2000 + i))
        bins += n * T
    torch.cuda.synchronize()
    for t in trials:
        eng.prepare(int(t.spectrum.numel()), t.bank, 2)
    tiles = {}
    plan = eng.plan(trials, graph=not args.no_graph)
    run_batches([plan], 1)
    n = max(2, args.steps // 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_batches([plan], n)
    ms = (time.perf_counter() - t0) * 1000.0 / n
    res = plan.run()
    for t, r in zip(trials, res):
        key = f"N={int(t.spectrum.numel())},T={85 if t.bank == banks[85] else 169}"
        tiles[key] = r.tile
    plan.close()
    out = {"value": round(48 * 1000.0 / ms, 2), "unit": "DM-trials/s", "ms_per_batch": round(ms, 3),
           "fop_gsamples_per_s": round(bins / (ms * 1e-3) / 1e9, 1), "tiles": tiles,
           "statuses_ok": all(r.status == 0 for r in res), "graph": not args.no_graph,
           "workload": "48 trials, N in {2^19,2^20,2^21,2^22}, banks 85 (z in [-84,84]) / 169 (z in [-168,168]) "
                       "x 421 taps, H = 16 for every 4th trial else 8"}
    log(f"uneven: {out}")
    return out


# --------------------------------------------------------------------------- serving

def served_form(args, eng, banks, bank_taps):
    import torch
    from fdas_amd import synth
    from fdas_amd.engine import Trial
    from fdas_amd.host import FDAS
    sizes, weights = [1 << 20, 1 << 21, 1 << 22], [0.3, 0.5, 0.2]
    pool = {}
    for n in sizes:
        for j in range(3):
            for T in (85, 169):
                w = synth.random_trial(3000 + 10 * j + (n.bit_length()), n, z_max=84.0 if T == 85 else 168.0)
                x = synth.device_spectrum(w)
                pool.setdefault((n, T), []).append((x, device_thresholds(x, H_METRIC)))
    torch.cuda.synchronize()
    tiles = {}
    for (n, T) in pool:
        tiles[(n, T)] = eng.prepare(n, banks[T], 2)

    # plain one-request-per-launch server: one class FDAS context per shape, at the engine's tile
    plain_ctx = {}
    for (n, T) in pool:
        f = FDAS(log=lambda lvl, msg: (log(msg) if lvl >= 1 else None))
        assert f.initialise_accelerator(n, n_templates=T, max_taps=421, tile=tiles[(n, T)], n_harmonics=H_METRIC,
                                        cands_per_harmonic=CAP), f.last_error
        assert f.upload_template_taps(bank_taps[T], FDAS.A), f.last_error
        plain_ctx[(n, T)] = f

    def plain_one(key, x, thr):
        f = plain_ctx[key]
        assert f.set_input_device(x.data_ptr(), FDAS.A), f.last_error
        assert f.perform_ft_convolution(FDAS.AllAccelerations, FDAS.A), f.last_error
        assert f.perform_harmonic_summing(thr, FDAS.AllAccelerations, FDAS.A), f.last_error
        f.retrieve_candidates_wide(FDAS.A)

    # warm every path once
    for key, lst in pool.items():
        plain_one(key, *lst[0])

    def picks_for(n_req, seed):
        rng = np.random.default_rng(seed)
        picks = []
        for _ in range(n_req):
            n = int(rng.choice(sizes, p=weights))
            T = int(rng.choice([85, 169]))
            x, thr = pool[(n, T)][int(rng.integers(0, 3))]
            picks.append(((n, T), x, thr))
        return picks

    def serve(picks, arrivals):
        """Served form on the given arrival times (s from the start): (throughput, latencies, statuses ok)."""
        n_req = len(picks)
        results = []

        def poller():
            while len(results) < n_req:
                results.extend(eng.poll(max_results=64, timeout_ms=200))

        th = threading.Thread(target=poller, daemon=True)
        th.start()
        t_start = time.perf_counter()
        for i, (key, x, thr) in enumerate(picks):
            delay = t_start + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            eng.submit(Trial(x, banks[key[1]], thr, 2, tag=i))
        th.join(timeout=300)
        t_end = time.perf_counter()
        assert len(results) == n_req, f"served {len(results)} of {n_req}"
        return n_req / (t_end - t_start), [r.latency_ms for r in results], all(r.status == 0 for r in results)

    def plain(picks, arrivals):
        """Plain form on the same arrivals: FIFO, one request per launch."""
        lat = []
        t_start = time.perf_counter()
        for i, (key, x, thr) in enumerate(picks):
            arr = t_start + arrivals[i]
            delay = arr - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            plain_one(key, x, thr)
            lat.append((time.perf_counter() - arr) * 1000.0)
        return len(picks) / (time.perf_counter() - t_start), lat

    def at_rate(rate, seed):
        n_req = max(40, int(rate * args.serve_seconds))
        rng = np.random.default_rng(seed)
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
        picks = picks_for(n_req, seed + 1)
        thr_s, lat_s, ok = serve(picks, arrivals)
        thr_p, lat_p = plain(picks, arrivals)
        served = {"requests": n_req, "throughput": round(thr_s, 2), "p50_ms": pct(lat_s, 50),
                  "p99_ms": pct(lat_s, 99), "statuses_ok": ok}
        plain_r = {"throughput": round(thr_p, 2), "p50_ms": pct(lat_p, 50), "p99_ms": pct(lat_p, 99)}
        return {"offered": round(rate, 1), "served": served, "plain": plain_r}

    out = {"rates": {}, "workload": "Poisson arrivals; N in {2^20,2^21,2^22} (0.3/0.5/0.2), banks 85/169 x 421 taps, "
                                    "H = 8, 4096 candidates/plane, spectra resident in HBM; plain form at the "
                                    "engine's tile per shape",
           "hbm_budget_gb": 200, "tiles": {f"N={n},T={T}": t for (n, T), t in tiles.items()}}
    for rate in [float(r) for r in args.serve_rates.split(",") if r]:
        out["rates"][str(int(rate))] = r = at_rate(rate, int(rate) + 17)
        log(f"serve {rate:.0f} req/s: {r}")
    # Capacity: every request offered at once, served and plain, on the same requests.
    picks = picks_for(args.capacity_requests, 7)
    zero = np.zeros(len(picks))
    cap_s, lat_s, ok = serve(picks, zero)
    cap_p, _ = plain(picks, zero)
    out["capacity"] = {"requests": len(picks), "served_req_per_s": round(cap_s, 1),
                       "plain_req_per_s": round(cap_p, 1), "served_vs_plain": round(cap_s / cap_p, 2),
                       "statuses_ok": ok}
    log(f"serve capacity: {out['capacity']}")
    # Latency at 50 %, 80 % and 100 % of the served capacity (plain on the same arrivals).
    out["load"] = {}
    for frac in (0.5, 0.8, 1.0):
        out["load"][f"{int(frac * 100)}%"] = r = at_rate(frac * cap_s, 100 + int(frac * 100))
        log(f"serve at {int(frac * 100)} % of capacity: {r}")
    st = eng.stats()
    out["engine_stats"] = {k: int(v) for k, v in st.items()}
    for f in plain_ctx.values():
        f.close()
    return out


# --------------------------------------------------------------------------- main

def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed batches of the batched form")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64, help="trials per batch")
    ap.add_argument("--forms", default="single,plain,batched,uneven,serve",
                    help="comma list of single, plain, batched, uneven, serve, power (power: not by default)")
    ap.add_argument("--power-seconds", type=float, default=8.0, help="length of each phase of the power form")
    ap.add_argument("--tile", type=int, default=0, help="engine tile size (0: autotuned per shape)")
    ap.add_argument("--rows-per-item", type=int, default=0)
    ap.add_argument("--hsum-rows", type=int, default=0)
    ap.add_argument("--group-trials", type=int, default=8)
    ap.add_argument("--ftc-wg-per-cu", type=int, default=0,
                    help="FT-convolution workgroups resident per CU (0: one per work item)")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--schedule", default="lagged", choices=["lagged", "fused", "streams"],
                    help="plan schedule of each group's FT convolution against the previous group's summing")
    ap.add_argument("--plain-trials", type=int, default=20)
    ap.add_argument("--serve-rates", default="50,100,200,400")
    ap.add_argument("--capacity-requests", type=int, default=400)
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only; the FDAS hot path would shard DM trials across GPUs with no "
              "collective (each GPU runs its own trials; candidate lists merge on the host), so a multi-GPU "
              "figure would be N independent copies of this run",
              file=sys.stderr)
        sys.exit(2)
    forms = set(args.forms.split(","))

    import fdas_amd
    from fdas_amd import synth

    w = synth.full_workload()
    log(f"workload: {w.name}")
    x = w.spectrum()
    bank85 = fdas_amd.make_drift_bank(synth.drift_grid(85), 421)
    bank169 = fdas_amd.make_drift_bank(synth.drift_grid(169), 421)
    thr = synth.noise_thresholds(x, H_METRIC)

    batched = uneven = served = None
    eng = make_engine(args)
    banks = {85: eng.add_bank(bank85), 169: eng.add_bank(bank169)}
    tile = eng.prepare(N_METRIC, banks[85], 2)
    shape = eng.shape_info(N_METRIC, banks[85], 2)
    workload_trials = batch_workload(args) if forms & {"batched", "plain"} else []
    if "batched" in forms:
        batched = batched_form(args, eng, banks[85], workload_trials)
    out_forms = single_trial_forms(args, x, bank85, thr, forms, tile, workload_trials) \
        if forms & {"single", "plain"} else {}
    if "uneven" in forms:
        uneven = uneven_form(args, eng, banks)
    if "serve" in forms:
        served = served_form(args, eng, banks, {85: bank85, 169: bank169})
    power = None
    if "power" in forms:
        plan = None
        if workload_trials:
            from fdas_amd.engine import Trial
            plan = eng.plan([Trial(xd, banks[85], th, 2, tag=i) for i, (xd, th) in enumerate(workload_trials)])
        power = power_form(x, bank85, tile, args.power_seconds, plan)
        if plan is not None:
            plan.close()
        log(f"power: {power}")
    eng.close()
    del workload_trials

    if batched is not None:
        value, ms_per_step = batched["value"], batched["ms_per_step"]
        steps, trials_per_launch = args.steps, args.group_trials
        n_cands = trials_per_launch * batched["candidates_per_trial_mean"]
        if batched["fused_span_ms"] > 0:
            # one launch: FT convolution of one group + harmonic summing of the group before
            kernels = {"fused1k_kernel<8,true>": {
                "ms": batched["fused_span_ms"],
                "bytes": ftc_bytes(N_METRIC, T_METRIC, tile, trials_per_launch)
                + hsum_bytes(N_METRIC, T_METRIC, n_cands, trials_per_launch)}}
        else:
            kernels = {
                ftc_kernel_name(tile): {"ms": batched["ftc_span_ms"],
                                        "bytes": ftc_bytes(N_METRIC, T_METRIC, tile, trials_per_launch)},
                "hsum_kernel<8,true>": {"ms": batched["hsum_span_ms"],
                                        "bytes": hsum_bytes(N_METRIC, T_METRIC, n_cands, trials_per_launch)},
            }
        workload = (f"batched uniform: {args.batch} DM trials per step, N=2^21 bins, 85 templates x 421 taps, "
                    f"8 harmonics, 4096 candidates/harmonic, spectra resident, groups of {args.group_trials}, "
                    f"{batched['form']} launches")
        in_flight = f"2 batches, groups of {args.group_trials} trials"
    else:
        st = out_forms.get("single_trial_form")
        if st is None:
            raise SystemExit("bench.py: --forms needs batched or single for the headline value")
        value, ms_per_step, steps, trials_per_launch = st["value"], st["ms_per_trial"], 30, 1
        kernels = {ftc_kernel_name(tile): {"ms": st["ftconv_ms"], "bytes": ftc_bytes(N_METRIC, T_METRIC, tile)},
                   "hsum_kernel<8,true>": {"ms": st["hsum_ms"], "bytes": hsum_bytes(N_METRIC, T_METRIC, 8 * CAP)}}
        workload = "single-trial FDAS, N=2^21 bins, 85 templates x 421 taps, 8 harmonics, spectra resident"
        in_flight = "2 trials (buffer sets A/B)"

    dominant = max(kernels, key=lambda k: kernels[k]["ms"])
    kd = kernels[dominant]
    achieved = kd["bytes"] / (kd["ms"] * 1e-3) / 1e9
    ftc_name = [k for k in kernels if k.startswith("ftconv") or k.startswith("fused")][0]
    traffic = MEASURED_TRAFFIC.get((dominant, trials_per_launch)) if batched is not None else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": int(traffic) if traffic else None,
                "traffic_source": ("profiles/batched_n2p21_t85_h8_tile1024_traffic.txt (rocprofv3 --pmc, bytes per "
                                   "launch)") if traffic else None,
                "kernel": dominant, "kernel_ms": round(kd["ms"], 4), "algorithmic_bytes": int(kd["bytes"]),
                "trials_per_launch": trials_per_launch,
                "kernels": {k: {"ms": round(v["ms"], 4), "algorithmic_bytes": int(v["bytes"]),
                                "gbs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)} for k, v in kernels.items()},
                "ftconv_fft_tflops": round(trials_per_launch * ftc_flops(shape, tile)
                                           / (kernels[ftc_name]["ms"] * 1e-3) / 1e12, 2),
                "ftconv_flops_per_trial": int(ftc_flops(shape, tile)),
                "path_achieved_gbs": round(value * path_bytes(N_METRIC, T_METRIC) / 1e9, 1),
                "path_frac": round(value * path_bytes(N_METRIC, T_METRIC) / 1e9 / HBM_PEAK_GBS, 4),
                "fp32_vector_peak_tflops": FP32_PEAK_TFLOPS,
                "valu_issue": ({k: valu_issue(k, trials_per_launch, v["ms"])
                                for k, v in kernels.items() if (k, trials_per_launch) in MEASURED_VALU}
                               if batched is not None else None),
                "valu_source": ("profiles/batched_n2p21_t85_h8_tile1024_instructions.txt (rocprofv3 SQ_INSTS_VALU "
                                "per launch); clocks: profiles/power_clock_n2p21_t85_tile1024.txt (amd-smi)")
                if batched is not None else None}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (bounded sample)")
        cpu = cpu_baseline(x, bank85, thr, CAP)
        log(f"cpu: {cpu}")

    plain = out_forms.get("plain_gpu_form")
    plain_h2d = out_forms.get("plain_gpu_form_h2d")
    out = {
        "metric": "DM-trials/sec (2^21 bins, 85 tmpl, 8 harm)",
        "value": round(value, 2),
        "unit": "DM-trials/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 N(0,1) series + 0-2 drifting 8-harmonic tones per trial; drift banks "
                "z=2(j-c) of 85/169 templates)",
        "config": {"workload": workload, "n_bins": N_METRIC, "n_templates": T_METRIC, "tile": tile,
                   "rows_per_item": shape["rows_per_item"], "hsum_rows": shape["hsum_rows"],
                   "row_groups": shape["n_row_groups"], "n_harmonics": H_METRIC, "cands_per_harmonic": CAP,
                   "batch": args.batch if batched else 1, "in_flight": in_flight},
        "vs_plain": round(value / plain["value"], 2) if plain else None,
        "vs_plain_h2d": round(value / plain_h2d["value"], 2) if plain_h2d else None,
        "batched_form": batched,
        "single_trial_form": out_forms.get("single_trial_form"),
        "plain_gpu_form": plain,
        "plain_gpu_form_h2d": plain_h2d,
        "uneven_form": uneven,
        "served_form": served,
        "power_form": power,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
