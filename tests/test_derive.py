"""Derived constants (read_mua_mus, io.cpp:179-237): library vs oracle, and closed forms (CPU)."""
import math

import numpy as np
import pytest

import fixtures as F
import mcml_gpu_amd as M
import oracle_ffi as O

CASES = {
    "cfg0": lambda: F.config("cfg0", 10),
    "cfg1": lambda: F.config("cfg1", 10),
    "cfg2": lambda: F.config("cfg2", 10),
    "HW8_A": lambda: F.example("HW8_A", 10),
    "head_model": lambda: F.example("head_model", 10),
    "fiber_busy": lambda: F.fiber_busy(10),
    "ring_busy": lambda: F.ring_busy(10),
    "vdh": lambda: F.vdh_slab(10),
}


@pytest.mark.parametrize("name", list(CASES))
def test_library_and_oracle_derive_identically(name):
    sim = CASES[name]()
    assert bytes(M.derive(sim)) == bytes(O.derive(sim))


def test_cfg1_layer_table():
    d = M.derive(F.config("cfg1", 10))
    h = np.float32([0.3, 0.5, 0.2, 0.3])
    z = np.float32(0)
    for l in range(1, 5):
        assert d.layer[l].z_min == z
        z = np.float32(z + h[l - 1])
        assert d.layer[l].z_max == z
    assert d.layer[5].z_min == z and d.layer[5].z_max == 10.0  # last layer forced to 10 cm
    mua = np.float32([0.2, 0.1, 0.3, 0.4, 0.25])
    mus = np.float32([150, 200, 80, 250, 120])
    for l in range(1, 6):
        assert d.layer[l].mutr == np.float32(np.float32(1) / np.float32(mua[l - 1] + mus[l - 1]))
    assert d.layer[0].n == np.float32(1.457) and d.layer[6].n == np.float32(1.457)


def test_start_weight_closed_form():
    for n0, n1 in [(1.457, 1.4), (1.0, 1.4), (1.0, 1.0), (1.33, 1.45)]:
        cfg = F.base_cfg(1, 1, n0, n0, 1, 1, F.probes((0, 0, 0), [(1.0, 0.1, 0.1, 0.0)]))
        d = M.derive(F.abi.make_sim(cfg, [(0.0, 1.0, 10.0, n1, 0.0)]))
        a, b = float(np.float32(n0)), float(np.float32(n1))
        r = ((a - b) / (a + b)) ** 2
        assert d.start_weight == np.float32(int(4294967295.0 * (1 - r)))


def test_deep_stack_keeps_sum_of_thicknesses():
    # when the upper layers already exceed 10 cm the last layer has zero thickness
    cfg = F.base_cfg(1, 3, 1.4, 1.4, 1, 1, F.probes((0, 0, 0), [(1.0, 0.1, 0.1, 0.0)]))
    d = M.derive(F.abi.make_sim(cfg, [(6.0, 1, 10, 1.4, 0), (5.0, 1, 10, 1.4, 0), (99.0, 1, 10, 1.4, 0)]))
    assert d.layer[3].z_min == 11.0 and d.layer[3].z_max == 11.0


def test_probe_angles_and_critical_angles():
    sim = F.fiber_busy(10)
    d = M.derive(sim)
    pi_f = np.float32(3.141592654)
    for i, deg in enumerate([10.0, 0.0, 25.0], start=1):
        a = np.float32(np.float32(np.float32(deg) * pi_f) / np.float32(180))
        assert d.det_cos_a[i] == np.float32(math.cos(float(a)))
        assert d.det_sin_a[i] == np.float32(math.sin(float(a)))
    for i, na in enumerate([0.5, 0.6, 0.6], start=1):
        crit = np.float32(math.asin(float(np.float32(np.float32(na) / np.float32(1.457)))))
        assert abs(d.det_crit[i] - crit) <= np.spacing(crit)
        assert d.det_cos_crit[i] == np.float32(math.cos(float(d.det_crit[i])))


def test_grid_defaults_and_overrides():
    d = M.derive(F.config("cfg1", 10))
    assert (d.grid_nr, d.grid_nz) == (1000, 500)
    assert d.grid_dr == np.float32(0.01) and d.grid_dz == np.float32(0.01)
    d2 = M.derive(F.config("cfg2", 10))
    assert (d2.grid_nr, d2.grid_nz) == (200, 400)
    assert d2.grid_dr == np.float32(0.02) and d2.grid_dz == np.float32(0.005)
