"""Simulation inputs used by the tests.

The reference ships inputs only (example/HW8_A, example/head_model and the README example);
they are copied under tests/golden/ as data.  The configurations built here add
detection-heavy probes so that per-photon parity has many detected photons to compare:
with the reference's NA 0.12, r = 0.02 cm fibres only about 1e-8 of the photons are
detected."""
import json
import os

from mcml_gpu_amd import abi

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CONFIGS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "configs")


def parse_input_line(vals, L):
    layers, i = [], 0
    for l in range(L):
        h = 0.0
        if l < L - 1:
            h = vals[i]
            i += 1
        layers.append((h, vals[i], vals[i + 1], vals[i + 2], vals[i + 3]))
        i += 4
    return layers


def from_files(json_path, txt_path, photons=None, **overrides):
    cfg = json.load(open(json_path))
    cfg.update(overrides)
    if photons is not None:
        cfg["number_photons"] = photons
    vals = [float(x) for x in open(txt_path).read().split()]
    return abi.make_sim(cfg, parse_input_line(vals, cfg["number_layers"]))


def example(name, photons):
    return from_files(os.path.join(GOLDEN, name, "sim_setup.json"),
                      os.path.join(GOLDEN, name, "GPUMC_input.txt"), photons)


def config(name, photons=None, **kw):
    return from_files(os.path.join(CONFIGS, name + "_sim_set.json"),
                      os.path.join(CONFIGS, name + "_input.txt"), photons, **kw)


def probes(src=(0.0, 0.0, 0.0), dets=()):
    return {"source": {"NA": src[0], "radius": src[1], "angle": src[2]}, "num_SDS": len(dets),
            "detectors": [{"pos": p, "NA": na, "radius": r, "angle": a} for (p, na, r, a) in dets]}


def base_cfg(photons, L, upper_n, lower_n, src_obl, det_obl, prb, refl=0.0):
    return {"number_simulation": 1, "number_photons": photons, "number_layers": L,
            "detector_reflectance": refl, "upper_n": upper_n, "lower_n": lower_n,
            "source_probe_oblique": src_obl, "detector_probe_oblique": det_obl, "probes": prb}


def hemisphere(photons, mua=10.0, mus=90.0, g=0.0, n=1.0):
    """Semi-infinite medium, matched or mismatched boundary, one detector that accepts
    every escaping photon (fiber mode, pos 0, radius 1e3 cm, NA = upper_n)."""
    cfg = base_cfg(photons, 1, 1.0, 1.0, 1, 1, probes((0, 0, 0), [(0.0, 1.0, 1000.0, 0.0)]))
    return abi.make_sim(cfg, [(0.0, mua, mus, n, g)])


def vdh_slab(photons, n_slab=1.0):
    """van de Hulst's slab: d = 0.02 cm, mua 10, mus 90, g 0.75, index n_slab in air,
    realised as a 2-layer model whose second layer is a clear index-1 medium (mus tiny,
    mua 0) leading to the lower ambient medium."""
    cfg = base_cfg(photons, 2, 1.0, 1.0, 1, 1, probes((0, 0, 0), [(0.0, 1.0, 1000.0, 0.0)]))
    return abi.make_sim(cfg, [(0.02, 10.0, 90.0, n_slab, 0.75), (0.0, 0.0, 1e-6, 1.0, 0.0)])


def fiber_busy(photons):
    """Three-layer fibre-mode case with wide, tilted fibres, an angled source, and a
    partially reflecting detector face, so every branch of detection is exercised."""
    prb = probes((0.37, 0.075, 8.0), [(0.25, 0.5, 0.15, 10.0), (0.5, 0.6, 0.2, 0.0),
                                      (0.9, 0.6, 0.3, 25.0)])
    cfg = base_cfg(photons, 3, 1.457, 1.457, 1, 1, prb, refl=0.35)
    return abi.make_sim(cfg, [(0.1, 0.3, 60.0, 1.4, 0.8), (0.3, 0.05, 120.0, 1.37, 0.9),
                              (0.0, 0.2, 80.0, 1.45, 0.85)])


def ring_busy(photons):
    """Ring mode with an NA + radius source and wide annuli (launch sampling paths)."""
    prb = probes((0.37, 0.05, 0.0), [(0.1, 0.4, 0.05, 0.0), (0.3, 0.4, 0.1, 0.0),
                                     (0.7, 0.5, 0.2, 0.0)])
    cfg = base_cfg(photons, 2, 1.457, 1.3, 0, 0, prb, refl=0.1)
    return abi.make_sim(cfg, [(0.15, 0.4, 70.0, 1.4, 0.85), (0.0, 0.1, 110.0, 1.38, 0.9)])


def max_stack(photons, ring=False):
    """The largest problem the formats allow: 10 layers (header.h:38) with alternating
    indices, and 30 detectors (header.h:44) spread from 0.05 to 1.5 cm, fibre or ring mode."""
    dets = [(0.05 + 0.05 * i, 0.6, 0.04, 5.0 * (i % 3) if not ring else 0.0) for i in range(30)]
    cfg = base_cfg(photons, 10, 1.45, 1.33, 0 if ring else 1, 0 if ring else 1, probes((0.3, 0.05, 0.0), dets),
                   refl=0.2)
    layers = []
    for i in range(10):
        h = 0.0 if i == 9 else 0.04 + 0.02 * (i % 4)
        layers.append((h, 0.05 + 0.03 * i, 40.0 + 9.0 * i, 1.33 + 0.015 * (i % 5), 0.7 + 0.025 * (i % 8)))
    return abi.make_sim(cfg, layers)
