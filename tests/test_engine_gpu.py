"""GPU tests of the batched / served engine (through the C ABI).

* a trial's lists from a plan equal those of the single-trial path at the same
  tile size, bit for bit (same kernels, same row groups, same FP32 order);
* uneven batches (N, banks of 85 and 169 templates, 8 and 16 harmonics, all
  three parts) against the FP64 oracle, with the bars of test_gpu_parity;
* a plan replayed as a HIP graph gives the eager result on every replay;
* served requests give the same result alone and in a mix, and admission
  keeps device state within the HBM budget.
"""

import random
import threading
import time

import numpy as np
import pytest

import fdas_amd
from fdas_amd import synth
from fdas_amd.engine import Engine, EngineError, Trial
from fdas_amd.host import FDAS
from oracle import fdas_oracle as O

from test_gpu_parity import assert_lists_match, bank_for, make, spectrum, wide_by_plane

pytestmark = pytest.mark.gpu

PART_NAMES = {0: "neg", 1: "pos", 2: "all"}


@pytest.fixture(scope="module")
def banks():
    return {85: bank_for(85), 169: bank_for(169)}


def quiet(level, msg):
    pass


def planes_of(res, n_templates):
    c = n_templates // 2
    out = []
    for ky, pw in zip(res.keys, res.powers):
        out.append(((ky >> 22).astype(np.int64) - c, (ky & ((1 << 22) - 1)).astype(np.int64), pw))
    return out


def single_trial(x, bank, thr, part, tile, cap):
    f = make(x.size, n_templates=bank.shape[0], max_taps=bank.shape[1], tile=tile, n_harmonics=len(thr),
             cands_per_harmonic=cap)
    assert f.upload_template_taps(bank, FDAS.A), f.last_error
    loc = np.zeros(f.get_candidate_list_sz(), np.uint32)
    pwr = np.zeros(f.get_candidate_list_sz(), np.float32)
    assert f.launch(x, thr, loc, pwr, part, FDAS.A), f.last_error
    assert f.wait(FDAS.A), f.last_error
    planes, det = wide_by_plane(f, FDAS.A, len(thr))
    f.close()
    return planes, det


def assert_same(res_a, res_b):
    assert res_a.status == res_b.status
    assert res_a.tile == res_b.tile
    assert np.array_equal(res_a.n_detected, res_b.n_detected)
    for a, b in zip(res_a.keys, res_b.keys):
        assert np.array_equal(a, b)
    for a, b in zip(res_a.powers, res_b.powers):
        assert np.array_equal(a, b)


def test_plan_equals_single_trial_path(banks):
    cap = 256
    with Engine(tile=2048, cands_per_harmonic=cap, group_trials=2, groups_in_flight=2, log=quiet) as eng:
        b = eng.add_bank(banks[85])
        cases = [(1 << 14, 11, 2), (12007, 12, 0), (1 << 14, 13, 1)]
        trials, xs = [], []
        for n, seed, part in cases:
            x = spectrum(n, seed)
            xs.append(x)
            trials.append(Trial(x, b, synth.noise_thresholds(x, 8), part, tag=seed))
        results = eng.run(trials)
        for (n, seed, part), x, t, r in zip(cases, xs, trials, results):
            assert r.tag == seed and r.status == 0 and r.tile == 2048
            planes, det = single_trial(x, banks[85], t.thresholds, part, 2048, cap)
            assert np.array_equal(r.n_detected, det)
            for (gt, gb, gp), (et, eb, ep) in zip(planes, planes_of(r, 85)):
                assert np.array_equal(gt, et) and np.array_equal(gb, eb) and np.array_equal(gp, ep)


def test_uneven_batch_against_oracle(banks):
    cap = 128
    cases = [  # (n_bins, templates, harmonics, part)
        (1 << 13, 85, 8, 2), (10007, 169, 16, 2), (1 << 14, 85, 16, 0), (6007, 169, 8, 1),
        (3 * 4096, 85, 8, 1), (5003, 169, 16, 0), (1 << 13, 85, 8, 2),
    ]
    with Engine(cands_per_harmonic=cap, group_trials=3, groups_in_flight=2, log=quiet) as eng:
        ids = {t: eng.add_bank(banks[t]) for t in banks}
        trials, xs = [], []
        for i, (n, T, H, part) in enumerate(cases):
            x = spectrum(n, 100 + i, f0_frac=0.2 + 0.05 * i, z=-30.0 + 10 * i)
            xs.append(x)
            trials.append(Trial(x, ids[T], synth.noise_thresholds(x, H), part, tag=i))
        plan = eng.plan(trials)
        results = plan.run()
        info = plan.info()
        assert info["n_groups"] == 3 and info["n_trials"] == len(cases)
        assert info["ftc_ms"] > 0 and info["hsum_ms"] > 0
        plan.close()
        for (n, T, H, part), x, t, r in zip(cases, xs, trials, results):
            assert r.status == 0 and r.tile in (1024, 2048, 4096)
            _, _, sel = O.run(x, banks[T], t.thresholds, H, cap, PART_NAMES[part])
            assert_lists_match(planes_of(r, T), r.n_detected, sel, t.thresholds, cap, T // 2)


def test_graph_replays_match_eager(banks):
    torch = pytest.importorskip("torch")
    cap = 64
    with Engine(tile=2048, cands_per_harmonic=cap, group_trials=2, groups_in_flight=2, log=quiet) as eng:
        b = eng.add_bank(banks[85])
        trials = []
        for i, n in enumerate([1 << 13, 9001, 1 << 13, 7777, 1 << 12]):
            x = spectrum(n, 300 + i)
            trials.append(Trial(x, b, synth.noise_thresholds(x, 8), 2, tag=i))
        ref = eng.run(trials)
        plan = eng.plan(trials, graph=True)
        for rep in range(3):
            out = plan.run()
            assert plan.info()["graph"] == 1
            for a, r in zip(out, ref):
                assert_same(a, r)
            # eager device work between replays
            y = torch.randn(1 << 20, device="cuda")
            assert torch.isfinite(y.sum()).item()
            torch.cuda.synchronize()
        plan.close()


@pytest.mark.parametrize("graph", [False, True])
def test_plan_request_replacement(banks, graph):
    cap = 64
    with Engine(tile=2048, cands_per_harmonic=cap, log=quiet) as eng:
        b = eng.add_bank(banks[85])
        x0, x1, x2 = spectrum(1 << 13, 1), spectrum(1 << 13, 2), spectrum(1 << 13, 3)
        t0 = Trial(x0, b, synth.noise_thresholds(x0, 8), 2, tag=0)
        t1 = Trial(x1, b, synth.noise_thresholds(x1, 8), 2, tag=1)
        plan = eng.plan([t0, t1], graph=graph)
        first = plan.run()
        thr2 = synth.noise_thresholds(x2, 8) * np.float32(0.9)
        plan.set_trial(1, spectrum=x2, thresholds=thr2)
        second = plan.run()
        assert_same(second[0], first[0])
        alone = eng.run([Trial(x2, b, thr2, 2, tag=1)])[0]
        assert_same(second[1], alone)
        plan.close()


def test_device_spectrum_equals_host(banks):
    torch = pytest.importorskip("torch")
    with Engine(tile=2048, cands_per_harmonic=64, log=quiet) as eng:
        b = eng.add_bank(banks[85])
        x = spectrum(12345, 9)
        thr = synth.noise_thresholds(x, 8)
        host = eng.run([Trial(x, b, thr, 2)])[0]
        d = torch.from_numpy(x).to("cuda")
        dev = eng.run([Trial(d, b, thr, 2)])[0]
        assert_same(dev, host)


def test_served_results_same_alone_and_in_mix(banks):
    cap = 128
    with Engine(cands_per_harmonic=cap, group_trials=4, groups_in_flight=2, log=quiet) as eng:
        ids = {t: eng.add_bank(banks[t]) for t in banks}
        # two long requests first: while they occupy both groups in flight, the
        # burst behind them queues and is admitted in groups
        heavy = []
        for i in range(2):
            x = spectrum(1 << 22, 900 + i, f0_frac=0.2, amp=0.002)  # weak tone: detections stay within the spill
            heavy.append(Trial(x, ids[169], synth.noise_thresholds(x, 8), 2, tag=100 + i))
        specs = [(1 << 13, 85, 8), (1 << 14, 169, 8), (9001, 85, 16), (1 << 13, 169, 16), (5000, 85, 8)]
        burst = []
        for i in range(14):
            n, T, H = specs[i % len(specs)]
            x = spectrum(n, 500 + i % 7, f0_frac=0.1 + 0.03 * i)
            burst.append(Trial(x, ids[T], synth.noise_thresholds(x, H), i % 3, tag=i))
        mix = heavy + burst
        for t in mix:  # fix every shape's tile before the comparison
            eng.prepare(t.spectrum.size, t.bank, t.part)
        alone = {}
        for t in mix:
            eng.submit(t)
            (r,) = eng.drain()
            alone[t.tag] = r
        for t in mix:
            eng.submit(t)
        got = {r.tag: r for r in eng.drain()}
        assert sorted(got) == sorted(t.tag for t in mix)
        for t in mix:
            assert got[t.tag].status == 0
            assert_same(got[t.tag], alone[t.tag])
            assert got[t.tag].latency_ms > 0
        st = eng.stats()
        assert st["completed"] == st["submitted"] == 2 * len(mix)
        assert st["max_group"] >= 2  # requests queued behind running groups were batched
        assert st["in_flight"] == 0 and st["queued"] == 0


def test_served_admission_within_budget(banks):
    cap, spill = 64, 4096
    n = 1 << 13
    # state of one trial: FOP 85 x n floats + lists; the budget holds one and a half
    per_trial = 85 * n * 4 + 8 * n + 8 * (2 * spill + 2 * (cap + 4096) + 2 * cap + 2048 + 3) * 4 + 16 * 256
    budget = int(1.5 * per_trial)
    with Engine(tile=2048, rows_per_item=16, hsum_rows=11, cands_per_harmonic=cap, spill_per_harmonic=spill,
                hbm_budget_bytes=budget, log=quiet) as eng:
        b = eng.add_bank(banks[85])
        trials = []
        for i in range(5):
            x = spectrum(n, 700 + i)
            trials.append(Trial(x, b, synth.noise_thresholds(x, 8), 2, tag=i))
        for t in trials:
            eng.submit(t)
        big = spectrum(4 * n, 777)
        eng.submit(Trial(big, b, synth.noise_thresholds(big, 8), 2, tag=99))
        got = {r.tag: r for r in eng.drain()}
        assert got[99].status == 2  # FDAS_ENOMEM: beyond the budget on its own
        for t in trials:
            assert got[t.tag].status == 0
            _, _, sel = O.run(t.spectrum, banks[85], t.thresholds, 8, cap, "all")
            assert_lists_match(planes_of(got[t.tag], 85), got[t.tag].n_detected, sel, t.thresholds, cap, 42)
        st = eng.stats()
        assert st["arena_peak_bytes"] <= budget
        assert st["arena_reserved_bytes"] <= budget


def test_request_validation(banks):
    with Engine(tile=2048, cands_per_harmonic=64, log=quiet) as eng:
        b = eng.add_bank(banks[85])
        x = spectrum(4096, 1)
        with pytest.raises(EngineError):
            eng.run([Trial(x, b + 5, synth.noise_thresholds(x, 8))])
        with pytest.raises(EngineError):
            eng.run([Trial(x, b, synth.noise_thresholds(x, 8), part=7)])
        with pytest.raises(ValueError):
            eng.run([Trial(x, b, np.full(17, 1.0, np.float32))])
        with pytest.raises(EngineError):
            eng.add_bank(np.zeros((84, 421), np.complex64))
        # 1500-tap templates need a 4096-point tile; a 2048 engine refuses them
        with pytest.raises(EngineError):
            eng.add_bank(fdas_amd.make_drift_bank(synth.drift_grid(5, 300.0), 1500))


@pytest.mark.parametrize("n,T,H,part", [(4099, 9, 1, 2), (1001, 85, 3, 0), (20011, 169, 16, 1), (64, 85, 8, 2)])
def test_edge_shapes_against_oracle(banks, n, T, H, part):
    """N not a multiple of 4 (scalar loads), one harmonic, tiny N, wide banks, each part."""
    bank = banks.get(T)
    if bank is None:
        bank = bank_for(T)
    cap = 32
    x = spectrum(n, 77 + n, f0_frac=0.37)
    thr = synth.noise_thresholds(x, H)
    with Engine(cands_per_harmonic=cap, log=quiet) as eng:
        b = eng.add_bank(bank)
        (r,) = eng.run([Trial(x, b, thr, part)])
        assert r.status == 0
        _, _, sel = O.run(x, bank, thr, H, cap, PART_NAMES[part])
        assert_lists_match(planes_of(r, T), r.n_detected, sel, thr, cap, T // 2)
        # served: the same result
        eng.submit(Trial(x, b, thr, part, tag=5))
        (s,) = eng.drain()
        assert_same(s, r)


def test_plan_schedules(banks):
    """Plans order group g's FT convolution against group g-1's harmonic summing as consecutive
    kernels on one stream (lagged, the default), as one fused kernel whose workgroups
    interleave both kinds of work (1024-point tiles, one summing shape per group), or on two
    streams side by side.  The fused kernel's four-wave transform rounds its products in
    another order than the one-wave kernel, so every schedule is held to the oracle's bars;
    each is bit-identical between eager runs and graph replays, and the lagged and two-stream
    schedules' lists equal the single-trial path's."""
    cap = 128
    cases = [  # (n_bins, templates, harmonics): groups of 2 -> fused pairs, a mixed-H group, ragged N
        (1 << 14, 85, 8), (12007, 85, 8), (1 << 14, 169, 16), (9000, 85, 16),
        (1 << 13, 85, 8), (1 << 14, 169, 8), (6007, 169, 16), (1 << 13, 85, 16), (5000, 85, 8),
    ]
    spec = []
    for i, (n, T, H) in enumerate(cases):
        x = spectrum(n, 700 + i, f0_frac=0.15 + 0.07 * i, z=-40.0 + 9 * i)
        spec.append((x, T, synth.noise_thresholds(x, H)))
    out = {}
    for schedule in ("fused", "lagged", "streams"):
        fused = schedule == "fused"
        with Engine(tile=1024, cands_per_harmonic=cap, group_trials=2, groups_in_flight=2, schedule=schedule,
                    log=quiet) as eng:
            ids = {T: eng.add_bank(banks[T]) for T in banks}
            trials = [Trial(x, ids[T], thr, 2, tag=i) for i, (x, T, thr) in enumerate(spec)]
            plan = eng.plan(trials)
            res = plan.run()
            info = plan.info()
            assert info["lagged"] == (0 if schedule == "streams" else 1)
            assert (info["n_fused"] >= 2) if fused else (info["n_fused"] == 0)
            assert (info["fused_ms"] > 0) == fused
            if not fused:  # eager runs time each kernel on its own, also when both share one stream
                assert info["ftc_ms"] > 0 and info["hsum_ms"] > 0
            plan.close()
            graph = eng.plan(trials, graph=True)
            for _ in range(2):
                for a, b in zip(graph.run(), res):
                    assert_same(a, b)
            graph.close()
            out[schedule] = res
    for a, b in zip(out["lagged"], out["streams"]):
        assert_same(a, b)
    for i, (x, T, thr) in enumerate(spec):
        _, _, sel = O.run(x, banks[T], thr, len(thr), cap, "all")
        for schedule in ("fused", "lagged"):
            r = out[schedule][i]
            assert r.status == 0 and r.tile == 1024
            assert_lists_match(planes_of(r, T), r.n_detected, sel, thr, cap, T // 2)
    # the default schedule equals the single-trial path bit for bit
    for i in (0, 4):  # 85 templates, 8 harmonics: the reference's 32-bit lists hold them
        x, T, thr = spec[i]
        planes, det = single_trial(x, banks[T], thr, 2, 1024, cap)
        r = out["lagged"][i]
        assert np.array_equal(r.n_detected, det)
        for (gt, gb, gp), (et, eb, ep) in zip(planes, planes_of(r, T)):
            assert np.array_equal(gt, et) and np.array_equal(gb, eb) and np.array_equal(gp, ep)


def test_engine_close_releases_open_plans(banks):
    """Closing an engine with plans still open (one finished, one in flight, one graph-captured)
    destroys them with it: their handles become inert and later calls fail cleanly.  A plan that
    an exception leaves open inside `with Engine(...)` takes this path."""
    x = spectrum(1 << 16, 31)
    thr = synth.noise_thresholds(x, 8)
    eng = Engine(tile=1024, cands_per_harmonic=32, log=quiet)
    b = eng.add_bank(banks[85])
    done = eng.plan([Trial(x, b, thr, 2)])
    ref = done.run()[0]
    running = eng.plan([Trial(x, b, thr, 2), Trial(x, b, thr, 1)])
    running.launch()
    captured = eng.plan([Trial(x, b, thr, 0)], graph=True)
    assert captured.run()[0].status == 0
    eng.close()
    for p in (done, running, captured):
        assert not p._h
        p.close()  # a no-op now
    with pytest.raises(EngineError):
        done.launch()
    # a new engine in the same process works and gives the same lists
    with Engine(tile=1024, cands_per_harmonic=32, log=quiet) as eng2:
        (r,) = eng2.run([Trial(x, eng2.add_bank(banks[85]), thr, 2)])
        assert_same(r, ref)


def test_served_requests_can_be_canceled_while_queued(banks):
    """Requests withdrawn while still queued complete at once with FDAS_ECANCELED; the others
    complete normally with the lists they give alone."""
    cap = 64
    with Engine(tile=2048, cands_per_harmonic=cap, group_trials=2, groups_in_flight=2, log=quiet) as eng:
        b = eng.add_bank(banks[85])
        n = 1 << 18
        eng.prepare(n, b, 2)
        x = spectrum(n, 811)
        thr = synth.noise_thresholds(x, 8)
        alone = eng.run([Trial(x, b, thr, 2, tag=0)])[0]
        outs = [eng.submit(Trial(x, b, thr, 2, tag=100 + i)) for i in range(24)]
        withdrawn = eng.cancel(123) + eng.cancel(122)  # the last two submitted
        assert eng.cancel(99999) == 0
        done = {r.tag: r for r in eng.drain()}
        assert len(done) == 24
        canceled = [t for t, r in done.items() if r.status == 8]
        assert len(canceled) == withdrawn and set(canceled) <= {122, 123}
        for t, r in done.items():
            if r.status != 8:
                assert r.status == 0
                assert np.array_equal(r.n_detected, alone.n_detected)
                for a, c in zip(r.keys, alone.keys):
                    assert np.array_equal(a, c)
        assert eng.stats()["completed"] == eng.stats()["submitted"]


def test_served_concurrent_submitters_and_cancels(banks):
    """Two threads submit mixed shapes (N 2^18..2^20 and ragged, banks 85/169, H 8/16) while a
    third withdraws recent tags, under an HBM budget that makes admission wait.  Every request
    completes exactly once, as FDAS_ECANCELED when a cancel withdrew it (the counts agree), else
    with the lists its spectrum gives alone, bit for bit."""
    cap = 128
    shapes = [(1 << 18, 85, 8), (1 << 19, 169, 8), (1 << 20, 85, 16), (1 << 18, 169, 16), (200003, 85, 8)]
    with Engine(cands_per_harmonic=cap, group_trials=4, groups_in_flight=3, hbm_budget_bytes=6 * 10**9,
                log=quiet) as eng:
        ids = {T: eng.add_bank(banks[T]) for T in banks}
        specs = []
        for i, (n, T, H) in enumerate(shapes):
            x = spectrum(n, 5000 + i)
            eng.prepare(n, ids[T], 2)
            specs.append((x, T, synth.noise_thresholds(x, H)))
        alone = [eng.run([Trial(x, ids[T], thr, 2, tag=0)])[0] for (x, T, thr) in specs]

        per_thread = 150
        sent, lock, stop, withdrawn = {}, threading.Lock(), threading.Event(), [0]

        def submitter(seed, base):
            rnd = random.Random(seed)
            for j in range(per_thread):
                s = rnd.randrange(len(specs))
                x, T, thr = specs[s]
                with lock:
                    sent[base + j] = s
                eng.submit(Trial(x, ids[T], thr, 2, tag=base + j))
                if rnd.random() < 0.3:
                    time.sleep(rnd.random() * 0.002)

        def canceller():
            rnd = random.Random(7)
            while not stop.is_set():
                with lock:
                    tags = list(sent)[-20:]
                if tags:
                    withdrawn[0] += eng.cancel(rnd.choice(tags))
                time.sleep(0.001)

        threads = [threading.Thread(target=submitter, args=(1, 100000)),
                   threading.Thread(target=submitter, args=(2, 200000)),
                   threading.Thread(target=canceller)]
        for t in threads:
            t.start()
        done, t0 = [], time.time()
        while len(done) < 2 * per_thread and time.time() - t0 < 60:
            done.extend(eng.poll(max_results=64, timeout_ms=1000))
        stop.set()
        for t in threads:
            t.join()
        assert len(done) == 2 * per_thread
        assert len({r.tag for r in done}) == 2 * per_thread
        canceled = [r for r in done if r.status == 8]
        assert len(canceled) == withdrawn[0]
        for r in done:
            if r.status != 8:
                assert_same(r, alone[sent[r.tag]])
        st = eng.stats()
        assert st["completed"] == st["submitted"] and st["arena_in_use_bytes"] == 0


def test_tuning_cache_shared_between_engines(banks, tmp_path):
    """An engine appends the shape it tunes to its cache file (malformed and invalid lines are
    skipped); a second engine reading the file takes that shape without timing anything, and gives
    the same lists for the same request, bit for bit."""
    path = tmp_path / "tune.txt"
    path.write_text("not a tuning line\nsomekey tile=999 rows_per_item=1 hsum_rows=1\n")
    n = 1 << 16
    x = spectrum(n, 901)
    thr = synth.noise_thresholds(x, 8)
    logs_a, logs_b = [], []
    with Engine(cands_per_harmonic=64, log=lambda level, msg: logs_a.append(msg)) as a:
        a.tune_cache(path)
        ba = a.add_bank(banks[85])
        a.prepare(n, ba, 2)
        shape_a = a.shape_info(n, ba, 2)
        ra = a.run([Trial(x, ba, thr, 2)])[0]
    assert any("autotune" in m for m in logs_a)
    lines = path.read_text().splitlines()
    assert len(lines) == 3 and f"tile={shape_a['tile']} rows_per_item={shape_a['rows_per_item']}" in lines[-1]
    with Engine(cands_per_harmonic=64, log=lambda level, msg: logs_b.append(msg)) as b:
        b.tune_cache(path)
        bb = b.add_bank(banks[85])
        b.prepare(n, bb, 2)
        assert b.shape_info(n, bb, 2) == shape_a
        rb = b.run([Trial(x, bb, thr, 2)])[0]
    assert any("(tuning cache)" in m for m in logs_b) and not any("autotune" in m for m in logs_b)
    assert len(path.read_text().splitlines()) == 3  # nothing new was timed
    assert_same(ra, rb)
