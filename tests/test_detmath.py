"""The deterministic math of mcml_detmath.h against float64 references (CPU).

These pin the functions both the oracle and the HIP kernels use: a wrong coefficient,
sign or quadrant fails here, independently of any transport test."""
import numpy as np

import oracle_ffi as O


def ulp_err(y32, ref64):
    y = y32.astype(np.float64)
    sp = np.spacing(np.abs(ref64).astype(np.float32)).astype(np.float64)
    return np.abs(y - ref64) / sp


def _tables():
    import ctypes as C
    sc, lg = C.POINTER(C.c_float)(), C.POINTER(C.c_float)()
    nsc, nlg = C.c_int(0), C.c_int(0)
    O.lib().oracle_tables(C.byref(sc), C.byref(nsc), C.byref(lg), C.byref(nlg))
    return (np.ctypeslib.as_array(sc, (2 * nsc.value,)).copy(),
            np.ctypeslib.as_array(lg, (2 * nlg.value,)).copy())


def test_tables_are_their_formulas():
    """Every table entry is the float64 formula rounded to nearest binary32: cos/sin of
    2 pi j / 256, and c_j = 1 / (1 + j / 128) with L_j = -ln(c_j)."""
    sc, lg = _tables()
    j = np.arange(256)
    a = 2 * np.pi * j / 256
    assert np.array_equal(sc[0::2], np.cos(a).astype(np.float32))
    assert np.array_equal(sc[1::2], np.sin(a).astype(np.float32))
    j = np.arange(128)
    c = (1.0 / (1.0 + j / 128.0)).astype(np.float32)
    assert np.array_equal(lg[0::2], c)
    assert np.array_equal(lg[1::2], (-np.log(c.astype(np.float64))).astype(np.float32))


def neglog(x):
    y = np.zeros(len(x), np.float32)
    O.lib().oracle_neglog_u32_n(np.ascontiguousarray(x, np.uint32), y, len(x))
    return y


def test_neglog_every_step_value():
    """-ln(u) for all 2^23 values u = (2k + 1) 2^-24 the step can draw (k = the draw's low
    23 bits): within 7.5e-8 + 9e-8 |ln u| of float64 (6.3e-8 absolute below 1, 8.7e-8
    relative above), positive everywhere (no zero-length step), smallest at k = 2^23 - 1
    (-ln(1 - 2^-24) = 5.96e-8) and largest at k = 0 (-ln(2^-24) = 16.64 mean free paths)."""
    k = np.arange(2 ** 23, dtype=np.uint64)
    x = k.astype(np.uint32)
    y = neglog(x)
    ref = -np.log((2 * k + 1).astype(np.float64) * 2.0 ** -24)
    err = np.abs(y.astype(np.float64) - ref)
    assert (err <= 7.5e-8 + 9e-8 * ref).all(), err.max()
    assert (y > 0).all()
    assert y.argmin() == len(y) - 1 and abs(y[-1] - 2.0 ** -24) < 1e-14
    assert abs(y[0] - 24 * np.log(2)) < 2e-6
    # the high 9 bits of the draw are not used
    hi = neglog(x[::4097] | np.uint32(0xA5800000))
    assert np.array_equal(hi.view(np.uint32), y[::4097].view(np.uint32))


def test_sincos_u32():
    """sin/cos of 2 pi x / 2^32 against float64, on random draws and at every cell edge."""
    rng = np.random.default_rng(6)
    j = np.arange(256, dtype=np.uint64)
    x = np.concatenate([rng.integers(0, 2 ** 32, 2000000, dtype=np.uint64), j << 24, (j << 24) | 0xFFFFFF,
                        (j << 24) | 0x800000]).astype(np.uint32)
    s = np.zeros(len(x), np.float32)
    c = np.zeros(len(x), np.float32)
    O.lib().oracle_sincos_u32_n(x, s, c, len(x))
    a = 2 * np.pi * x.astype(np.float64) / 2 ** 32
    assert np.abs(s - np.sin(a)).max() < 1.2e-7
    assert np.abs(c - np.cos(a)).max() < 1.2e-7
    n2 = s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2
    assert np.abs(n2 - 1).max() < 3e-7
    # cell starts are the table entries themselves
    k = len(x) - 3 * 256
    assert np.array_equal(s[k:k + 256], np.sin(2 * np.pi * np.arange(256) / 256).astype(np.float32))


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
