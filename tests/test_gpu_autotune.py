"""The autotuner on the device: it times real launches of every knob value,
ends no slower than the defaults as measured, and reuses its cache."""
import json
import os

import pytest

from branson_amd import autotune, capi

pytestmark = pytest.mark.gpu

DECK = os.path.join(os.path.dirname(__file__), "golden", "decks", "cube_40.xml")


def test_tune_deck_and_cache(tmp_path, monkeypatch):
    cache = str(tmp_path / "tuned.json")
    best = autotune.tune_deck(DECK, sample_photons=300_000, sample_steps=1, cache=cache)
    for k, v in autotune.DEFAULTS.items():
        assert best[k] in autotune.CANDIDATES[k]
    entry = next(iter(json.load(open(cache)).values()))
    assert entry["transport_ms"] <= entry["default_ms"]
    assert entry["evaluations"] >= sum(len(v) for v in autotune.CANDIDATES.values()) - 4

    def no_driver(*a, **k):
        raise AssertionError("cached result expected")

    monkeypatch.setattr(capi, "Driver", no_driver)
    assert autotune.tune_deck(DECK, sample_photons=300_000, sample_steps=1, cache=cache) == best
    # the tuned options run
    monkeypatch.undo()
    with capi.Driver(DECK, t_stop=0.001, **best) as d:
        r = d.run()
    assert r[0]["census_size"] == 862677
