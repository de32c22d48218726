"""The batch schedule's autotuner (gpuzip_amd/autotune.py) without a GPU: its coordinate search
over a made-up cost, its rejection of candidates whose results are wrong, and its cache."""
import pytest

from gpuzip_amd import autotune


def test_search_finds_the_best_of_a_separable_cost():
    best_of = dict(autotune.DEFAULT, workers=1, chunk_kib=8, probe_spacing_kbit=256, probe_lead_bits=8192)
    calls = []

    def measure(cfg):
        calls.append(cfg)
        return 100.0 - sum(5.0 * abs(autotune.SPACE[k].index(cfg[k]) - autotune.SPACE[k].index(best_of[k]))
                           for k in autotune.SPACE)

    best, score, trials = autotune.coordinate_search(measure)
    assert best == best_of and score == 100.0
    assert trials[0][0] == autotune.DEFAULT
    # each candidate measured once in the search, then the start and the winner again
    assert len(calls) == len({tuple(sorted(c.items())) for c in calls}) + 3 == len(trials)
    assert [c for c, _ in trials[-3:]] == [autotune.DEFAULT, best_of, autotune.DEFAULT]


def test_search_never_takes_a_rejected_or_marginal_candidate():
    def measure(cfg):
        if cfg["chunk_kib"] == 16:
            return None  # wrong results: rejected however fast
        if cfg["workers"] == 1:
            return 101.0  # within min_gain of the start: not a change
        if cfg["probe_lead_bits"] == 8192:
            return 110.0
        return 100.0

    best, score, trials = autotune.coordinate_search(measure, min_gain=0.02)
    assert best == dict(autotune.DEFAULT, probe_lead_bits=8192) and score == 110.0
    assert any(s is None for _, s in trials)


def test_a_noisy_first_trial_does_not_decide():
    """the start's first score is low by chance (a device warming up): the search moves away,
    the confirmation measures both again and keeps the start; a real gain survives it"""
    seen = {"start": 0}

    def measure(cfg):
        if cfg == autotune.DEFAULT:
            seen["start"] += 1
            return 80.0 if seen["start"] == 1 else 100.0
        return 93.0 if cfg["chunk_kib"] == 8 else 90.0

    best, score, trials = autotune.coordinate_search(measure)
    assert best == autotune.DEFAULT and score == 100.0
    assert trials[-3][0] == trials[-1][0] == autotune.DEFAULT and trials[-2][0] != autotune.DEFAULT

    def measure2(cfg):
        return 110.0 if cfg["chunk_kib"] == 8 else 100.0

    best, score, _ = autotune.coordinate_search(measure2)
    assert best == dict(autotune.DEFAULT, chunk_kib=8) and score == 110.0


def test_rejected_start_is_an_error():
    with pytest.raises(RuntimeError):
        autotune.coordinate_search(lambda cfg: None)


def test_cache_round_trip(tmp_path, monkeypatch):
    path = str(tmp_path / "sub" / "tune.json")
    assert autotune.load(path, "k") is None
    autotune.save(path, "k", {"params": dict(autotune.DEFAULT), "GBs": 1.0})
    autotune.save(path, "k2", {"params": {}, "GBs": 2.0})
    assert autotune.load(path, "k")["GBs"] == 1.0 and autotune.load(path, "k2")["GBs"] == 2.0
    # a cached choice is returned without measuring anything
    monkeypatch.setattr(autotune, "measure_batch", lambda *a, **k: pytest.fail("measured"))
    got = autotune.tune_batch([], [], cache=path, key="k")
    assert got["cached"] and got["params"] == autotune.DEFAULT


def test_tuned_fields_are_the_config_tail():
    from gpuzip_amd import gpuinflate as gi
    assert set(autotune.TUNED_FIELDS) <= set(gi.TUNING_RANGES)
    for k in autotune.TUNED_FIELDS:
        lo, hi = gi.TUNING_RANGES[k]
        # 0 keeps a field's default (for persistent_per_cu: a workgroup per queue item)
        assert all(v == 0 or lo <= v <= hi for v in autotune.SPACE[k])
    assert set(autotune.SPACE["lanes_per_group"]) <= set(gi.LANES_PER_GROUP)
