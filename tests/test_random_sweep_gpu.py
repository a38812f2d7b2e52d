"""Seeded random sweeps of shapes through both device paths: single trials against the FP64 oracle,
and batched plans at large N against the single-trial path.

Single trials: each case draws N (1 .. 70,000), a bank (9, 33, 85 or 169 templates), 1 .. 16 harmonics, a FOP part,
a tile (1024 / 2048 / 4096), a cap (1 .. 256) and a threshold scale (including thresholds every bin
passes), then checks:
  * the single-trial path step by step (tiling, FT convolution, harmonic summing, wide candidates):
    FOP within 1e-5 of the plane maximum of the FP64 oracle, lists as in test_gpu_parity;
  * the engine at the same tile: every plane's list and detection count bit-equal to the single-trial
    path;
  * when the detections exceed the spill capacity, both paths report FDAS_EOVERFLOW.
Batched plans: see test_random_plans_at_large_n.
"""

import numpy as np
import pytest

import fdas_amd
from fdas_amd import synth
from fdas_amd.engine import Engine, EngineError, Trial
from fdas_amd.host import FDAS
from oracle import fdas_oracle as O

from test_engine_gpu import planes_of
from test_gpu_parity import FOP_TOL, assert_lists_match, rel_err, wide_by_plane

pytestmark = pytest.mark.gpu

PART_NAMES = {FDAS.NegativeAccelerations: "neg", FDAS.PositiveAccelerations: "pos", FDAS.AllAccelerations: "all"}
N_CASES = 80


@pytest.fixture(scope="module")
def sweep_banks():
    return {9: fdas_amd.make_drift_bank(synth.drift_grid(9), 421),
            33: fdas_amd.make_drift_bank(synth.drift_grid(33), 97),
            85: fdas_amd.make_drift_bank(synth.drift_grid(85), 421),
            169: fdas_amd.make_drift_bank(synth.drift_grid(169), 421)}


def _draw(rng):
    lo, hi = [(1, 200), (200, 5000), (5000, 40000), (40000, 70000)][int(rng.integers(0, 4))]
    n = int(rng.integers(lo, hi))
    case = {"n": n, "T": int(rng.choice([9, 33, 85, 169])), "H": int(rng.integers(1, 17)),
            "part": int(rng.choice(list(PART_NAMES))), "tile": int(rng.choice([1024, 2048, 4096])),
            "cap": int(rng.choice([1, 7, 32, 256])), "scale": float(rng.choice([0.3, 0.8, 1.0, 1.5, -1.0])),
            "seed": int(rng.integers(0, 1 << 30))}
    tone = synth.Tone(max(1.0, n * float(rng.uniform(0.05, 0.9))), float(rng.uniform(-80, 80)),
                      int(rng.integers(1, 9)), float(rng.uniform(0.0, 0.2)))
    return case, tone


def test_random_shapes_both_paths(sweep_banks):
    rng = np.random.default_rng(20261203)
    engines = {}
    n_compared = n_overflow = 0
    try:
        for i in range(N_CASES):
            c, tone = _draw(rng)
            n, T, H, part, tile, cap = c["n"], c["T"], c["H"], c["part"], c["tile"], c["cap"]
            print(f"case {i}: {c}", flush=True)
            bank = sweep_banks[T]
            x = synth.Workload("sweep", n, c["seed"], [tone]).spectrum()
            if c["scale"] > 0:
                thr = (synth.noise_thresholds(x, H) * c["scale"]).astype(np.float32)
            else:  # every bin is a detection
                thr = -np.arange(1, H + 1, dtype=np.float32)
            if (tile, cap) not in engines:
                eng = Engine(tile=tile, rows_per_item=16, hsum_rows=11, cands_per_harmonic=cap,
                             log=lambda lvl, msg: None)
                engines[(tile, cap)] = (eng, {t: eng.add_bank(b) for t, b in sweep_banks.items()})
            eng, ids = engines[(tile, cap)]

            f = FDAS(log=lambda lvl, msg: None)
            try:
                assert f.initialise_accelerator(n, n_templates=T, max_taps=bank.shape[1], tile=tile,
                                                n_harmonics=H, cands_per_harmonic=cap), f.last_error
                assert f.upload_template_taps(bank, FDAS.A), f.last_error
                assert f.perform_input_tiling(x, FDAS.A), f.last_error
                assert f.perform_ft_convolution(part, FDAS.A), f.last_error
                if not f.perform_harmonic_summing(thr, part, FDAS.A):
                    assert f.last_status == 7, f"case {i}: status {f.last_status}: {f.last_error}"
                    with pytest.raises(EngineError) as ex:
                        eng.run([Trial(x, ids[T], thr, part)])
                    assert ex.value.status == 7, f"case {i}"
                    n_overflow += 1
                    continue
                fop = np.zeros(f.get_fop_sz(), np.float32)
                assert f.retrieve_FOP(fop, FDAS.A), f.last_error
                planes, det = wide_by_plane(f, FDAS.A, H)
            finally:
                f.close()

            rows = list(O.part_rows(T, PART_NAMES[part]))
            ref_fop, _, sel = O.run(x, bank, thr, H, cap, PART_NAMES[part])
            assert rel_err(fop.reshape(T, n)[rows], ref_fop[rows]) <= FOP_TOL, f"case {i}"
            assert_lists_match(planes, det, sel, thr, cap, T // 2)

            (r,) = eng.run([Trial(x, ids[T], thr, part)])
            assert r.status == 0, f"case {i}"
            assert np.array_equal(np.asarray(det[:H]), np.asarray(r.n_detected[:H])), f"case {i}"
            for k, ((gt, gb, gp), (et, eb, ep)) in enumerate(zip(planes, planes_of(r, T)), start=1):
                assert np.array_equal(gt, et) and np.array_equal(gb, eb) and np.array_equal(gp, ep), \
                    f"case {i}: plane {k}"
            n_compared += 1
    finally:
        for eng, _ in engines.values():
            eng.close()
    # the draw mixes both outcomes
    assert n_compared >= N_CASES // 2 and n_overflow >= 1


def _single_wide(x, bank, thr, part, tile, cap):
    """The single-trial path step by step with wide candidates; None when its spill overflows."""
    f = FDAS(log=lambda lvl, msg: None)
    try:
        assert f.initialise_accelerator(x.size, n_templates=bank.shape[0], max_taps=bank.shape[1], tile=tile,
                                        n_harmonics=len(thr), cands_per_harmonic=cap), f.last_error
        assert f.upload_template_taps(bank, FDAS.A), f.last_error
        assert f.perform_input_tiling(x, FDAS.A), f.last_error
        assert f.perform_ft_convolution(part, FDAS.A), f.last_error
        if not f.perform_harmonic_summing(thr, part, FDAS.A):
            assert f.last_status == 7, f.last_error
            return None
        return wide_by_plane(f, FDAS.A, len(thr))
    finally:
        f.close()


def test_random_plans_at_large_n(sweep_banks):
    """Seeded batched plans of 3-6 trials mixing N in {2^17 .. 2^22, ragged}, banks 85/169, H in
    {1, 4, 8, 16} and every part, eager and graph-replayed: each trial bit-equal to the single-trial
    path at the plan's tile, or overflowing the spill capacity where that path does."""
    rng = np.random.default_rng(77)
    cap = 256
    n_compared = 0
    with Engine(tile=1024, rows_per_item=16, hsum_rows=11, cands_per_harmonic=cap, log=lambda lvl, msg: None) as eng:
        ids = {T: eng.add_bank(sweep_banks[T]) for T in (85, 169)}
        for pi in range(3):
            trials, meta = [], []
            for i in range(int(rng.integers(3, 7))):
                n = int(rng.choice([1 << 17, 1 << 19, 1 << 20, 1 << 21, 1 << 22, int(rng.integers(1 << 17, 1 << 21))]))
                T, H, part = int(rng.choice([85, 169])), int(rng.choice([1, 4, 8, 16])), int(rng.integers(0, 3))
                x = synth.random_trial(int(rng.integers(0, 1 << 30)), n).spectrum()
                thr = (synth.noise_thresholds(x, H) * float(rng.choice([0.9, 1.0, 1.3]))).astype(np.float32)
                trials.append(Trial(x, ids[T], thr, part, tag=i))
                meta.append((x, T, thr, part))
            print(f"plan {pi}: " + ", ".join(f"N={m[0].size} T={m[1]} H={m[2].size} part={m[3]}" for m in meta),
                  flush=True)
            singles = [_single_wide(x, sweep_banks[T], thr, part, 1024, cap) for x, T, thr, part in meta]
            for graph in (False, True):
                if any(s is None for s in singles):
                    with pytest.raises(EngineError) as ex:
                        eng.run(trials, graph=graph)
                    assert ex.value.status == 7
                    continue
                for r, (x, T, thr, part), (planes, det) in zip(eng.run(trials, graph=graph), meta, singles):
                    H = thr.size
                    assert r.status == 0 and r.tile == 1024
                    assert np.array_equal(np.asarray(det[:H]), np.asarray(r.n_detected[:H]))
                    for (gt, gb, gp), (et, eb, ep) in zip(planes, planes_of(r, T)):
                        assert np.array_equal(gt, et) and np.array_equal(gb, eb) and np.array_equal(gp, ep)
                    n_compared += 1
    assert n_compared >= 10  # the seeded plans are not all overflows
