"""White Monte Carlo rescaling: exactness of w * exp(-sum dmua_l L_l) (CPU, oracle) and the
HIP kernel against a float64 restatement and a direct GPU run (GPU)."""
import numpy as np
import pytest

import oracle_ffi as O
from mcml_gpu_amd import batch, make_opts


def white_numpy(rows, offsets, mua_old, mua_new, power=1):
    """Reference in float64: sum over each group's rows of (w * exp(-sum (mua'-mua) L))^power."""
    L = len(mua_old)
    out = np.zeros(len(offsets) - 1)
    d = np.asarray(mua_new, np.float64) - np.asarray(mua_old, np.float64)
    for g in range(len(offsets) - 1):
        r = rows[offsets[g]:offsets[g + 1]].astype(np.float64)
        out[g] = ((r[:, 0] * np.exp(-(r[:, 1:L + 1] @ d))) ** power).sum()
    return out


def oracle_run(sim, seed, n):
    """Oracle first pass + replay of every detection: per-photon weights per SDS and rows."""
    d = O.derive(sim)
    rec, _ = O.first_pass(d, seed, 0, n)
    S, L = sim.num_sds, sim.number_layers
    rows = {s: [] for s in range(1, S + 1)}
    for key, sds in zip(rec["key"], rec["sds"]):
        hit, r = O.replay(d, seed, int(key), int(sds))
        assert hit == 1
        rows[int(sds)].append(r)
    return {s: np.array(v, np.float32).reshape(-1, L + 2) for s, v in rows.items()}


def per_photon_estimate(rows_by_sds, n, dmua=None):
    """Mean and standard error of the detected weight per launched photon, per SDS."""
    est = []
    for s in sorted(rows_by_sds):
        r = rows_by_sds[s].astype(np.float64)
        w = r[:, 0] if dmua is None else r[:, 0] * np.exp(-(r[:, 1:1 + len(dmua)] @ dmua))
        full = np.zeros(n)
        full[:len(w)] = w
        est.append((full.mean(), full.std() / np.sqrt(n)))
    return est


@pytest.mark.slow
def test_white_rescaling_is_exact_in_expectation():
    """Rescaling photons simulated at mua0 to mua1 agrees with photons simulated at mua1
    (independent seeds), within 4 combined standard errors per SDS."""
    p = batch.draw_parameters(1, seed=11)[0].copy()
    p[0:3] = [0.08, 0.15, 0.10]
    mua1 = np.array([0.35, 0.05, 0.40])
    n = 60000
    sim0 = batch.sim_from_params(p, n)
    p1 = p.copy()
    p1[0:3] = mua1
    sim1 = batch.sim_from_params(p1, n)
    rows0 = oracle_run(sim0, 101, n)
    rows1 = oracle_run(sim1, 202, n)
    resc = per_photon_estimate(rows0, n, mua1 - np.array(p[0:3]))
    direct = per_photon_estimate(rows1, n)
    for (m0, s0), (m1, s1) in zip(resc, direct):
        assert m0 > 0 and m1 > 0
        assert abs(m0 - m1) < 4 * np.hypot(s0, s1), (m0, m1, s0, s1)


def test_identity_rescaling_returns_the_weights():
    p = batch.draw_parameters(1, seed=12)[0]
    n = 3000
    rows = oracle_run(batch.sim_from_params(p, n), 5, n)
    for s, r in rows.items():
        offs = np.array([0, len(r)])
        got = white_numpy(r, offs, p[0:3], p[0:3])
        assert got[0] == pytest.approx(r[:, 0].astype(np.float64).sum(), rel=1e-12)


@pytest.mark.gpu
def test_white_kernel_matches_float64_restatement(ctx):
    sims, params, _ = batch.cfg3_sets(6, seed=5, budget_lo=2e4, budget_hi=5e4)
    res = ctx.run_batch(sims, make_opts(replay=True, pathlength=True), seed=9)
    rng = np.random.default_rng(0)
    m = 7
    mua_new = rng.uniform(0.05, 0.5, (len(sims), m, 3)).astype(np.float32)
    mua_new[:, 0, :] = params[:, 0:3]  # identity vector first
    out, sq = ctx.white_rescale(sims, res, mua_new, with_squares=True)
    assert np.array_equal(out, ctx.white_rescale(sims, res, mua_new))
    S = res.num_sds
    for j in range(len(sims)):
        offs = res.pl_offset[j * S:(j + 1) * S + 1] - res.pl_offset[j * S]
        rows = res.pathlength[res.pl_offset[j * S]:res.pl_offset[(j + 1) * S]]
        for i in range(m):
            ref = white_numpy(rows, offs, params[j, 0:3].astype(np.float32), mua_new[j, i])
            assert np.allclose(out[j, i], ref, rtol=2e-6, atol=0)
            ref_sq = white_numpy(rows, offs, params[j, 0:3].astype(np.float32), mua_new[j, i], power=2)
            assert np.allclose(sq[j, i], ref_sq, rtol=4e-6, atol=0)
        assert np.allclose(out[j, 0], res.reflectance[j], rtol=2e-6)


@pytest.mark.gpu
def test_white_rescaling_matches_direct_gpu_run(ctx):
    p = batch.draw_parameters(1, seed=21)[0].copy()
    p[0:3] = [0.06, 0.12, 0.09]
    mua1 = np.array([0.30, 0.40, 0.20], np.float32)
    n = 4_000_000
    s0 = batch.sim_from_params(p, n)
    p1 = p.copy()
    p1[0:3] = mua1
    s1 = batch.sim_from_params(p1, n)
    r0 = ctx.run(s0, make_opts(replay=True, pathlength=True), seed=31)
    r1 = ctx.run(s1, make_opts(replay=True, pathlength=True), seed=32)
    resc = ctx.white_rescale([s0], r0, mua1.reshape(1, 1, 3))[0, 0]
    for s in range(1, s0.num_sds + 1):
        a = r0.pathlength_records(0, s).astype(np.float64)
        b = r1.pathlength_records(0, s).astype(np.float64)
        wa = np.zeros(n)
        wa[:len(a)] = a[:, 0] * np.exp(-(a[:, 1:4] @ (mua1.astype(np.float64) - np.float32(p[0:3]))))
        wb = np.zeros(n)
        wb[:len(b)] = b[:, 0]
        se = np.hypot(wa.std(), wb.std()) / np.sqrt(n)
        assert abs(resc[s - 1] / n - r1.reflectance[0, s - 1] / n) < 4 * se
