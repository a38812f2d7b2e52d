#!/usr/bin/env python3
"""FDAS hot-path benchmark on one MI355X.

Metric: DM-trials/sec at N = 2^21 bins, 85 templates, 8 harmonics (4,096
candidates per harmonic), the single-trial configuration of the reference's
search.  One step = one DM trial through the whole device path: fused FT
convolution -> harmonic summing + thresholding + compaction -> per-plane
selection -> candidate read-back.  Inputs (spectrum, template spectra) are
resident in HBM before the timed region; two buffer sets A/B keep two trials
in flight, as the reference's pipelined driver does (main_gtest.cpp:388-437).

Also reported, in the same JSON line:
  * plain GPU form: one trial per launch, host spectrum uploaded, candidates
    read back and the stream synchronised per trial (p50 / p99 latency);
  * roofline of the dominant kernel (HIP-event durations on its own stream);
  * CPU baseline: the oracle's restatement of the reference path timed on a
    bounded sample on this host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ftc_bytes(n_bins, n_templates, tile):
    """Algorithmic HBM bytes of one FT-convolution launch: spectrum read + FOP write + template spectra."""
    return 8 * n_bins + 4 * n_templates * n_bins + 8 * n_templates * tile


def hsum_bytes(n_bins, n_templates, n_cands):
    """Algorithmic HBM bytes of one harmonic-summing launch: FOP read once + candidates out."""
    return 4 * n_templates * n_bins + 8 * n_cands


def fft_flops(n_bins, n_templates, tile, overlap):
    """Nominal 5 S log2 S per transform, forward per tile + inverse per (tile, template)."""
    payload = tile - overlap
    n_tiles = -(-n_bins // payload)
    per = 5.0 * tile * np.log2(tile)
    return n_tiles * per + n_tiles * n_templates * (per + 6 * tile + 3 * payload)


def cpu_baseline(x, bank, thr, cap, budget_s=10.0):
    """The oracle's C restatement of ft_conv.py + hsum.py (FP32, OpenMP, own FFT) on full trials.

    Runs whole N = 2^21 trials back to back until ~budget_s of CPU time has passed (at least 2).
    """
    from oracle import cpu
    n_threads = cpu.threads()
    cpu.fop(x[:4096], bank)  # warm-up (page-in, thread pool)
    times = []
    t_start = time.perf_counter()
    while len(times) < 2 or (time.perf_counter() - t_start < budget_s and len(times) < 50):
        t0 = time.perf_counter()
        fop = cpu.fop(x, bank)
        cpu.hsum(fop, thr, len(thr), cap)
        times.append(time.perf_counter() - t0)
        del fop
    med = float(np.median(times))
    return {"value": round(1.0 / med, 4), "unit": "DM-trials/s", "cores": n_threads, "kind": "port",
            "sample": f"{len(times)} full trials (N=2^21, 85 templates x 421 taps, 8 harmonics, 4096/plane) through "
                      f"oracle/fdas_cpu.c (FP32 overlap-save FFT convolution + harmonic summing + selection), "
                      f"median {med:.3f} s per trial on {n_threads} OpenMP threads",
            "ms_per_trial": round(med * 1000.0, 2)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="timed DM trials")
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--in-flight", type=int, default=2, choices=(1, 2),
                    help="trials in flight in the timed region (buffer sets A/B)")
    ap.add_argument("--isolated-steps", type=int, default=10,
                    help="extra one-at-a-time trials for uncontended kernel durations (0: skip)")
    ap.add_argument("--plain-trials", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile", type=int, default=2048)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU; the FDAS hot path shards DM trials across GPUs with no "
              "collective, so a multi-GPU figure is N independent copies of this run (see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import fdas_amd
    from fdas_amd import synth
    from fdas_amd.host import FDAS

    w = synth.full_workload()
    log(f"workload: {w.name}")
    x = w.spectrum()
    bank = fdas_amd.make_drift_bank(synth.drift_grid(w.n_templates), w.max_taps)
    thr = synth.noise_thresholds(x, w.n_harmonics)
    cap = w.cands_per_harmonic

    f = FDAS(log=lambda lvl, msg: (log(msg) if lvl >= 1 else None))
    assert f.initialise_accelerator(w.n_bins, tile=args.tile, n_templates=w.n_templates, max_taps=w.max_taps,
                                    n_harmonics=w.n_harmonics, cands_per_harmonic=cap), f.last_error
    for ab in (FDAS.A, FDAS.B):
        assert f.upload_template_taps(bank, ab), f.last_error
        assert f.perform_input_tiling(x, ab), f.last_error
    n_list = f.get_candidate_list_sz()
    loc = {ab: np.zeros(n_list, np.uint32) for ab in (FDAS.A, FDAS.B)}
    pwr = {ab: np.zeros(n_list, np.float32) for ab in (FDAS.A, FDAS.B)}
    part = FDAS.AllAccelerations

    def run(n_steps, depth, record=None):
        """n_steps trials, `depth` in flight on sets A/B; per-trial stage times appended to record."""
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

    isolated = None
    if args.isolated_steps > 0:
        run(2, 1)
        rec = []
        t0 = time.perf_counter()
        run(args.isolated_steps, 1, rec)
        ms = (time.perf_counter() - t0) * 1000.0 / args.isolated_steps
        st = np.array(rec)
        isolated = {"ms_per_trial": round(ms, 4), "trials_per_s": round(1000.0 / ms, 2),
                    "ftconv_ms": round(float(st[:, 1].mean()), 4), "hsum_ms": round(float(st[:, 2].mean()), 4),
                    "select_ms": round(float(st[:, 3].mean()), 4)}
        log(f"isolated (1 in flight): {isolated}")

    log(f"warmup {args.warmup} trials, {args.in_flight} in flight")
    run(args.warmup, args.in_flight)
    rec = []
    log(f"timing {args.steps} trials")
    t0 = time.perf_counter()
    run(args.steps, args.in_flight, rec)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = 1000.0 / ms_per_step
    st = np.array(rec)
    ftc_ms, hsum_ms, sel_ms = (float(st[:, i].mean()) for i in (1, 2, 3))
    log(f"timed: {ms_per_step:.3f} ms/trial ({value:.1f} trials/s); kernel means ftconv {ftc_ms:.3f} ms, "
        f"hsum {hsum_ms:.3f} ms, select {sel_ms:.3f} ms")

    # Plain GPU form: one trial per launch, host upload, read-back and synchronisation each time.
    plain = None
    if args.plain_trials > 0:
        lat = []
        for i in range(args.plain_trials + 2):
            t = time.perf_counter()
            assert f.launch(x, thr, loc[FDAS.A], pwr[FDAS.A], part, FDAS.A), f.last_error
            assert f.wait(FDAS.A), f.last_error
            if i >= 2:
                lat.append((time.perf_counter() - t) * 1000.0)
        lat = np.array(lat)
        plain = {"value": round(float(1000.0 / lat.mean()), 2), "unit": "DM-trials/s",
                 "p50_ms": round(float(np.percentile(lat, 50)), 3),
                 "p99_ms": round(float(np.percentile(lat, 99)), 3),
                 "form": "host spectrum upload + FT conv + HSUM + selection + read-back + sync, one trial per launch"}
        log(f"plain: {plain}")

    _, _, _, _, det = f.retrieve_candidates_wide(FDAS.A)
    n_cands = int(np.minimum(det, cap).sum())
    kernels = {
        "ftconv_kernel<2048>": {"ms": ftc_ms, "bytes": ftc_bytes(w.n_bins, w.n_templates, args.tile)},
        "hsum_kernel<8,true>": {"ms": hsum_ms, "bytes": hsum_bytes(w.n_bins, w.n_templates, n_cands)},
    }
    dominant = max(kernels, key=lambda k: kernels[k]["ms"])
    kd = kernels[dominant]
    achieved = kd["bytes"] / (kd["ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dominant, "kernel_ms": round(kd["ms"], 4), "algorithmic_bytes": int(kd["bytes"]),
                "kernels": {k: {"ms": round(v["ms"], 4), "algorithmic_bytes": int(v["bytes"]),
                                "gbs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)} for k, v in kernels.items()},
                "ftconv_fft_tflops": round(fft_flops(w.n_bins, w.n_templates, args.tile, 420) / (ftc_ms * 1e-3) / 1e12,
                                           2),
                "fp32_vector_peak_tflops": FP32_PEAK_TFLOPS,
                "select_ms": round(sel_ms, 4), "detections_per_plane": [int(v) for v in det]}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (bounded sample)")
        cpu = cpu_baseline(x, bank, thr, cap)
        log(f"cpu: {cpu}")

    f.close()
    out = {
        "metric": "DM-trials/sec (2^21 bins, 85 tmpl, 8 harm)",
        "value": round(value, 2),
        "unit": "DM-trials/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded N(0,1) series + drifting 8-harmonic tone; 85-template drift bank z=2(j-42))",
        "config": {"workload": "single-trial FDAS, N=2^21 bins, 85 templates x 421 taps, FFT 2048 overlap-save, "
                               "8 harmonics, 4096 candidates/harmonic, spectra resident",
                   "n_bins": w.n_bins, "n_templates": w.n_templates, "tile": args.tile,
                   "n_harmonics": w.n_harmonics, "cands_per_harmonic": cap, "in_flight": args.in_flight},
        "isolated_form": isolated,
        "plain_gpu_form": plain,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
