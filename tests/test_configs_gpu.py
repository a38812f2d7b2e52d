"""The batched configurations at their own workload sizes, through the engine with
the autotuner choosing tile, rows per item and rows per summing workgroup.

* batched uniform: 12 trials of N = 2^21, 85 templates, 8 harmonics, 4096 candidates
  per plane (groups of 8 and 4), eager and replayed as a HIP graph;
* batched uneven: one trial of every (N in {2^19, 2^20, 2^21, 2^22}) x (bank of 85 /
  169 templates) shape in one plan, 16 harmonics on some of them.

Every trial's lists equal, bit for bit, those of the single-trial path (class FDAS)
run at the tile the engine chose; trials at N = 2^21 / 85, 2^20 / 85 and 2^19 / 169
also meet the FP64 oracle's bars (FOP error <= 1e-5 of the plane maximum, lists
identical except within 1e-4 of the cut).  The 2^22 shapes are held to bit-equality
with the single-trial path, whose kernels the oracle checks at every other size.
"""

import numpy as np
import pytest

from fdas_amd import synth
from fdas_amd.engine import Engine, Trial
from fdas_amd.host import FDAS
from oracle import fdas_oracle as O

from test_engine_gpu import assert_same, planes_of, quiet
from test_gpu_parity import assert_lists_match, bank_for, make, wide_by_plane

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

CAP = 4096


def single_trial_lists(f, x, thr):
    """Lists of one trial through a class FDAS context: the blocking single-step calls
    (any T and H; the reference's 32-bit lists only hold T <= 127, H <= 8) and wide retrieval."""
    assert f.perform_input_tiling(x, FDAS.A), f.last_error
    assert f.perform_ft_convolution(FDAS.AllAccelerations, FDAS.A), f.last_error
    assert f.perform_harmonic_summing(thr, FDAS.AllAccelerations, FDAS.A), f.last_error
    return wide_by_plane(f, FDAS.A, len(thr))


def assert_equals_single_trial(r, planes, det, n_templates):
    assert np.array_equal(r.n_detected, det)
    for (gt, gb, gp), (et, eb, ep) in zip(planes, planes_of(r, n_templates)):
        assert np.array_equal(gt, et) and np.array_equal(gb, eb) and np.array_equal(gp, ep)


def assert_oracle(x, bank, thr, r, n_templates):
    ref = O.fop_fft(x, bank)
    sel, _ = O.detect_streaming(ref, len(thr), thr)
    del ref
    sel = O.select(sel, CAP)
    assert_lists_match(planes_of(r, n_templates), r.n_detected, sel, thr, CAP, n_templates // 2)


def test_batched_uniform_full_size():
    n = 1 << 21
    bank = bank_for(85)
    xs, thrs = [], []
    for i in range(12):
        x = synth.random_trial(1000 + i, n).spectrum()
        xs.append(x)
        thrs.append(synth.noise_thresholds(x, 8))
    print(f"batched uniform: 12 trials of N=2^21 generated", flush=True)
    with Engine(cands_per_harmonic=CAP, log=quiet) as eng:
        b = eng.add_bank(bank)
        tile = eng.prepare(n, b, 2)
        assert tile in (1024, 2048, 4096)
        trials = [Trial(x, b, t, 2, tag=1000 + i) for i, (x, t) in enumerate(zip(xs, thrs))]
        plan = eng.plan(trials)
        assert plan.info()["n_groups"] == 2
        results = plan.run()
        plan.close()
        graph = eng.plan(trials, graph=True)
        replay = graph.run()
        graph.close()
    for r, g in zip(results, replay):
        assert r.status == 0 and r.tile == tile
        assert_same(r, g)
    print(f"batched uniform: tile {tile}; comparing with the single-trial path", flush=True)
    f = make(n, tile=tile, cands_per_harmonic=CAP)
    assert f.upload_template_taps(bank, FDAS.A), f.last_error
    for x, thr, r in zip(xs, thrs, results):
        planes, det = single_trial_lists(f, x, thr)
        assert_equals_single_trial(r, planes, det, 85)
    f.close()
    # the FP64 oracle on the first trial that carries a tone, and on one without
    with_tone = [i for i in range(12) if synth.random_trial(1000 + i, n).tones]
    without = [i for i in range(12) if not synth.random_trial(1000 + i, n).tones]
    for i in (with_tone[:1] + without[:1]):
        print(f"batched uniform: FP64 oracle on trial {1000 + i}", flush=True)
        assert_oracle(xs[i], bank, thrs[i], results[i], 85)


def test_batched_uneven_every_shape():
    banks = {85: bank_for(85), 169: bank_for(169)}
    cases = []  # (n_bins, templates, harmonics)
    for j, n in enumerate([1 << 19, 1 << 20, 1 << 21, 1 << 22]):
        for T in (85, 169):
            cases.append((n, T, 16 if (j + T) % 3 == 0 else 8))
    assert {h for _, _, h in cases} == {8, 16}
    xs, thrs = [], []
    for i, (n, T, H) in enumerate(cases):
        x = synth.random_trial(2000 + i, n, z_max=84.0 if T == 85 else 168.0).spectrum()
        xs.append(x)
        thrs.append(synth.noise_thresholds(x, H))
    print(f"batched uneven: {len(cases)} trials generated", flush=True)
    with Engine(cands_per_harmonic=CAP, log=quiet) as eng:
        ids = {T: eng.add_bank(banks[T]) for T in banks}
        for n, T, _ in cases:
            eng.prepare(n, ids[T], 2)
        trials = [Trial(x, ids[T], t, 2, tag=2000 + i) for i, ((n, T, H), x, t) in enumerate(zip(cases, xs, thrs))]
        plan = eng.plan(trials, graph=True)
        results = plan.run()
        again = plan.run()
        plan.close()
    for (n, T, H), x, thr, r, g in zip(cases, xs, thrs, results, again):
        assert r.status == 0 and r.tile in (1024, 2048, 4096)
        assert_same(r, g)
        print(f"batched uneven: N={n} T={T} H={H} tile {r.tile}: single-trial path", flush=True)
        f = make(n, n_templates=T, tile=r.tile, n_harmonics=H, cands_per_harmonic=CAP)
        assert f.upload_template_taps(banks[T], FDAS.A), f.last_error
        planes, det = single_trial_lists(f, x, thr)
        f.close()
        assert_equals_single_trial(r, planes, det, T)
    for i, (n, T, H) in enumerate(cases):
        if (n, T) in ((1 << 19, 169), (1 << 20, 85)):
            print(f"batched uneven: FP64 oracle on N={n} T={T} H={H}", flush=True)
            assert_oracle(xs[i], banks[T], thrs[i], results[i], T)
