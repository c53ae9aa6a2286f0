"""The deterministic math of mcml_detmath.h against float64 references (CPU).

These pin the functions both the oracle and the HIP kernels use: a wrong coefficient,
sign or quadrant fails here, independently of any transport test."""
import numpy as np

import oracle_ffi as O


def ulp_err(y32, ref64):
    y = y32.astype(np.float64)
    sp = np.spacing(np.abs(ref64).astype(np.float32)).astype(np.float64)
    return np.abs(y - ref64) / sp


def test_log_all_generator_outputs():
    # every value the generator can produce: 1 - k * 2^-23, k = 2^23 - 1 .. 0 (ascending)
    k = np.arange(2 ** 23 - 1, -1, -1, dtype=np.float64)
    u = (1.0 - k * 2.0 ** -23).astype(np.float32)
    assert np.array_equal(u.astype(np.float64), 1.0 - k * 2.0 ** -23)  # exact in binary32
    y = np.zeros_like(u)
    O.lib().oracle_log_unit_n(u, y, len(u))
    ref = np.log(u.astype(np.float64))
    nz = ref != 0
    assert ulp_err(y[nz], ref[nz]).max() < 1.5
    assert y[-1] == 0.0  # log(1) = 0 exactly: u = 1 gives a zero-length step
    y2 = np.zeros_like(u)
    O.lib().oracle_log_n(u, y2, len(u))
    assert np.array_equal(y.view(np.uint32), y2.view(np.uint32))


def test_log_general_and_subnormal():
    rng = np.random.default_rng(3)
    x = np.concatenate([
        rng.uniform(1e-3, 1e3, 200000), 10.0 ** rng.uniform(-37, 38, 200000),
        np.array([1.0, 2.0, 0.5, np.sqrt(2.0), 1.4142135, 1.4142137]),
        np.array([1e-40, 1e-45, 3e-39, 1.1754942e-38])]).astype(np.float32)
    y = np.zeros_like(x)
    O.lib().oracle_log_n(x, y, len(x))
    ref = np.log(x.astype(np.float64))
    nz = ref != 0
    assert ulp_err(y[nz], ref[nz]).max() < 1.5


def test_sincos_turns():
    rng = np.random.default_rng(4)
    t = np.concatenate([rng.uniform(0, 1, 1000000), np.arange(0, 1.0001, 1 / 64)]).astype(np.float32)
    s = np.zeros_like(t)
    c = np.zeros_like(t)
    O.lib().oracle_sincos_turns_n(t, s, c, len(t))
    a = 2 * np.pi * t.astype(np.float64)
    assert np.abs(s - np.sin(a)).max() < 1.2e-7
    assert np.abs(c - np.cos(a)).max() < 1.2e-7
    # quarter turns are exact
    q = np.array([0, 0.25, 0.5, 0.75, 1.0], np.float32)
    s = np.zeros_like(q)
    c = np.zeros_like(q)
    O.lib().oracle_sincos_turns_n(q, s, c, len(q))
    assert list(np.abs(s)) == [0, 1, 0, 1, 0] and list(np.abs(c)) == [1, 0, 1, 0, 1]
    assert list(np.sign(s)[[1, 3]]) == [1, -1] and list(np.sign(c)[[0, 2, 4]]) == [1, -1, 1]


def test_acos():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-1, 1, 1000000),
                        np.array([-1, -0.5, -0.4999999, 0, 0.5, 0.5000001, 1], np.float64)]).astype(np.float32)
    y = np.zeros_like(x)
    O.lib().oracle_acos_n(x, y, len(x))
    ref = np.arccos(x.astype(np.float64))
    assert np.abs(y - ref).max() < 3e-7
    assert y[-1] == 0.0
    bad = np.array([1.0000001, -1.0000001, np.inf, np.nan], np.float32)
    yb = np.zeros_like(bad)
    O.lib().oracle_acos_n(bad, yb, len(bad))
    assert np.isnan(yb).all()


def test_rcp_within_half_ulp():
    """mcml_rcp (seed + three Newton steps) against float64 1/x over the ranges the spin uses
    (1 - dz^2 down to 2^-23 times the HG denominator) and across binades."""
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(1e-8, 2.5, 2000000), 10.0 ** rng.uniform(-30, 30, 500000),
                        2.0 ** np.arange(-60, 60), np.nextafter(2.0 ** np.arange(-60, 60), 0)]).astype(np.float32)
    y = np.zeros_like(x)
    O.lib().oracle_rcp_n(x, y, len(x))
    assert ulp_err(y, 1.0 / x.astype(np.float64)).max() < 0.52


def test_rsqrt_within_one_ulp_and_square_root():
    """mcml_rsqrt (seed + three Newton steps) against float64 1/sqrt(x); x * rsqrt(x) as the
    square root the spin uses, within 2 ulp; x = 0 gives 0 for that square root."""
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.uniform(0, 1, 2000000), 10.0 ** rng.uniform(-30, 30, 500000),
                        2.0 ** np.arange(-120, 120), np.nextafter(2.0 ** np.arange(-120, 120), 0)]).astype(np.float32)
    y = np.zeros_like(x)
    O.lib().oracle_rsqrt_n(x, y, len(x))
    ref = 1.0 / np.sqrt(x.astype(np.float64))
    assert ulp_err(y, ref).max() < 1.0
    s = (x * y).astype(np.float32)
    sref = np.sqrt(x.astype(np.float64))
    assert ulp_err(s, sref).max() < 2.0
    z = np.zeros(1, np.float32)
    yz = np.zeros(1, np.float32)
    O.lib().oracle_rsqrt_n(z, yz, 1)
    assert np.isfinite(yz[0]) and z[0] * yz[0] == 0.0
