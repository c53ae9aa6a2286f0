"""Batched workloads: many tissue parameter sets with their own photon budgets.

The reference runs the lines of input.txt one after another, each a full DoOneSimulation
(main.cu:130-135).  Here a whole batch goes to `mcml_run_batch`, one persistent launch
whose ticket queue spans every set.

`cfg3_sets` draws the 512-set, three-layer batch from a seed: uniform mua 0.05-0.5 /cm,
mus 50-300 /cm, g 0.7-0.95, n 1.33-1.45 per layer, L1/L2 thickness 0.05-1.0 cm, budgets
log-uniform 1e5-1e7.  Its probes are ring detectors that accept every escaping direction
(annuli of half-width 0.1 cm centred 0.4, 0.8 and 1.5 cm from a pencil beam, tissue under
air).  These give enough detected weight at these budgets for per-SDS reflectance to
be a usable training label.  The reference's own NA 0.12, r = 0.02 cm fibres detect about
one photon in 1e8.
"""
import numpy as np

from . import abi

RANGES = {"mua": (0.05, 0.5), "mus": (50.0, 300.0), "g": (0.7, 0.95), "n": (1.33, 1.45),
          "h": (0.05, 1.0)}
SDS_POS = (0.4, 0.8, 1.5)
SDS_HALF_WIDTH = 0.1


def ring_probes(positions=SDS_POS, half_width=SDS_HALF_WIDTH):
    return {"source": {"NA": 0.0, "radius": 0.0, "angle": 0.0}, "num_SDS": len(positions),
            "detectors": [{"pos": p, "NA": 1.0, "radius": half_width, "angle": 0.0} for p in positions]}


def base_config(photons):
    return {"number_simulation": 1, "number_photons": int(photons), "number_layers": 3,
            "detector_reflectance": 0.0, "upper_n": 1.0, "lower_n": 1.0,
            "source_probe_oblique": 0, "detector_probe_oblique": 0, "probes": ring_probes()}


def draw_parameters(n_sets, seed):
    """Optical parameters as an (n_sets, 14) array: mua[3], mus[3], g[3], n[3], h[2]."""
    rng = np.random.default_rng(seed)
    mua = rng.uniform(*RANGES["mua"], (n_sets, 3))
    mus = rng.uniform(*RANGES["mus"], (n_sets, 3))
    g = rng.uniform(*RANGES["g"], (n_sets, 3))
    n = rng.uniform(*RANGES["n"], (n_sets, 3))
    h = rng.uniform(*RANGES["h"], (n_sets, 2))
    return np.concatenate([mua, mus, g, n, h], axis=1)


def sim_from_params(p, photons):
    """abi.Sim for one 14-vector of parameters (mua[3], mus[3], g[3], n[3], h[2])."""
    mua, mus, g, n, h = p[0:3], p[3:6], p[6:9], p[9:12], p[12:14]
    layers = [(float(h[0]), mua[0], mus[0], n[0], g[0]), (float(h[1]), mua[1], mus[1], n[1], g[1]),
              (0.0, mua[2], mus[2], n[2], g[2])]
    return abi.make_sim(base_config(photons), layers)


def cfg3_sets(n_sets=512, seed=1, budget_lo=1e5, budget_hi=1e7):
    """The cfg3 batch: (list of abi.Sim, parameter array, budgets)."""
    params = draw_parameters(n_sets, seed)
    rng = np.random.default_rng(seed + 1_000_003)
    budgets = np.round(10 ** rng.uniform(np.log10(budget_lo), np.log10(budget_hi), n_sets)).astype(np.int64)
    sims = [sim_from_params(params[j], budgets[j]) for j in range(n_sets)]
    return sims, params, budgets
