#!/usr/bin/env python3
"""Per-photon step counts of each benchmark configuration, counted by the CPU oracle at
seed 1 over photons [0, 1e5): the n_int and n_bnd of the algorithmic FLOP per photon,
62 n_int + 44 n_bnd, that bench.py prices its transport launches with.  Committed as
algorithmic_counts.json beside this script; tests/test_algorithmic_counts.py recomputes them.

Usage (from the repository root): python tests/golden/make_algorithmic_counts.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import fixtures as F  # noqa: E402
import oracle_ffi as O  # noqa: E402

SEED = 1
PHOTONS = 100_000
CONFIGS = {
    "cfg0": lambda n: F.config("cfg0", n),
    "cfg1": lambda n: F.config("cfg1", n),
    "cfg2": lambda n: F.config("cfg2", n),
    "head_model": lambda n: F.example("head_model", n),
    "hw8a": lambda n: F.example("HW8_A", n),
}
OUT = os.path.join(HERE, "algorithmic_counts.json")


def count(name, photons=PHOTONS, seed=SEED):
    """Totals [steps, boundary, interaction] of the oracle's first pass over [0, photons)."""
    _, c = O.first_pass(O.derive(CONFIGS[name](photons)), seed, 0, photons)
    return {"steps": int(c[0]), "boundary": int(c[1]), "interaction": int(c[2])}


def main():
    out = {"seed": SEED, "photons": PHOTONS, "configs": {name: count(name) for name in CONFIGS}}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(json.dumps(out, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
