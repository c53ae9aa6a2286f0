"""Multiply-with-carry streams: recurrence, multipliers, seeding and uniformity (CPU)."""
import ctypes as C

import numpy as np
import sympy

import oracle_ffi as O


def test_recurrence_matches_bigint_restatement():
    L = O.lib()
    for seed, st, k in [(1, 0, 0), (1, 0, 1), (99, 3, 123456789), (2 ** 63 + 5, 511, 2 ** 39)]:
        s = C.c_uint64()
        a = C.c_uint32()
        L.oracle_stream_state(seed, st, k, C.byref(s), C.byref(a))
        state, mult = s.value, a.value
        n = 2000
        got = np.zeros(n, np.float32)
        L.oracle_uniforms(seed, st, k, got, n)
        exp = []
        for _ in range(n):
            t = mult * (state & 0xFFFFFFFF) + (state >> 32)  # x' = a x + c, carry = high word
            state = t
            x = t & 0xFFFFFFFF
            exp.append(1.0 - (x & 0x7FFFFF) / 2.0 ** 23)  # 2 - 1.k of the low 23 bits, exact
        assert np.array_equal(got, np.array(exp, np.float32))


def test_multiplier_table_safe_primes():
    n = C.c_uint32()
    ptr = O.lib().oracle_multiplier_table(C.byref(n))
    tab = np.ctypeslib.as_array(ptr, (n.value,)).copy()
    assert n.value == 4096
    assert (tab > 2 ** 31).all() and len(np.unique(tab)) == len(tab)
    assert np.all(np.diff(tab.astype(np.int64)) < 0)
    for a in [int(x) for x in tab[::97]] + [int(tab[-1])]:
        p = (a << 32) - 1
        assert sympy.isprime(p) and sympy.isprime((p - 1) // 2)


def test_stream_state_in_range_and_distinct():
    L = O.lib()
    seen = set()
    for k in range(2000):
        s = C.c_uint64()
        a = C.c_uint32()
        L.oracle_stream_state(7, 0, k, C.byref(s), C.byref(a))
        c, x = s.value >> 32, s.value & 0xFFFFFFFF
        assert c < a.value and (c, x) != (0, 0)
        seen.add(s.value)
    assert len(seen) == 2000
    u1 = np.zeros(64, np.float32)
    u2 = np.zeros(64, np.float32)
    L.oracle_uniforms(7, 0, 5, u1, 64)
    L.oracle_uniforms(7, 1, 5, u2, 64)
    assert not np.array_equal(u1, u2)


def test_uniform_moments_and_lag_correlation():
    n = 2_000_000
    u = np.zeros(n, np.float32)
    O.lib().oracle_uniforms(11, 0, 42, u, n)
    assert u.min() > 0 and u.max() <= 1
    u = u.astype(np.float64)
    se = np.sqrt(1 / 12 / n)
    assert abs(u.mean() - 0.5) < 5 * se
    assert abs(u.var() - 1 / 12) < 5 * np.sqrt(1 / 180 / n)
    r = np.corrcoef(u[:-1], u[1:])[0, 1]
    assert abs(r) < 5 / np.sqrt(n)
    h, _ = np.histogram(u, bins=256, range=(0, 1))
    chi2 = ((h - n / 256) ** 2 / (n / 256)).sum()
    assert chi2 < 255 + 6 * np.sqrt(2 * 255)


def test_streams_uncorrelated_across_photons():
    m, n = 400, 2000
    rows = np.zeros((m, n), np.float32)
    for k in range(m):
        O.lib().oracle_uniforms(3, 0, k, rows[k], n)
    means = rows.mean(axis=1)
    assert abs(means.mean() - 0.5) < 5 * np.sqrt(1 / 12 / (m * n))
    c = np.corrcoef(rows[:-1].ravel(), rows[1:].ravel())[0, 1]
    assert abs(c) < 5 / np.sqrt(m * n)
