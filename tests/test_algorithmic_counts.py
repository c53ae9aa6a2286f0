"""The committed per-photon step counts that price the benchmark's algorithmic FLOP
(tests/golden/algorithmic_counts.json) are what the oracle counts: seed 1, photons [0, 1e5)."""
import json
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

import golden.make_algorithmic_counts as G  # noqa: E402


def _committed():
    with open(os.path.join(HERE, "golden", "algorithmic_counts.json")) as f:
        return json.load(f)


def test_header():
    c = _committed()
    assert c["seed"] == G.SEED and c["photons"] == G.PHOTONS
    assert set(c["configs"]) == set(G.CONFIGS)


@pytest.mark.parametrize("name", list(G.CONFIGS))
def test_counts_match_oracle(name):
    got = G.count(name)
    assert got == _committed()["configs"][name]
    # every step is a boundary hit or an interaction (the step's uniform is below 1)
    assert got["steps"] == got["boundary"] + got["interaction"]
