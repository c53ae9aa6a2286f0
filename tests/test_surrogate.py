"""The reflectance surrogate: label statistics and features (CPU), and on the GPU the
trained network's gain over plain Monte Carlo and the fallback to plain Monte Carlo."""
import time

import numpy as np
import pytest

from mcml_gpu_amd import batch
from mcml_gpu_amd import surrogate as SG


def test_features_span_the_unit_box():
    lo = SG.LO[None, :]
    hi = SG.HI[None, :]
    f = SG.features(np.concatenate([lo, hi, 0.5 * (lo + hi)]))
    assert np.allclose(f[0], -1) and np.allclose(f[1], 1)
    # log scale for mua and mus: the arithmetic midpoint maps above 0 there, to 0 elsewhere
    assert np.all(f[2, :6] > 0) and np.allclose(f[2, 6:], 0, atol=1e-6)
    p = batch.draw_parameters(1000, 3)
    assert np.all(np.abs(SG.features(p)) <= 1.0 + 1e-6)


def test_relative_variance_is_the_binomial_one_for_equal_weights():
    """k of n photons detected with equal weight: relative variance of the mean is
    (1 - p) / (n p) = 1/k - 1/n."""
    n, k, w = 1000, 37, 2.5e9
    rv = SG.relative_variance(k * w, k * w * w, n)
    assert rv == pytest.approx(1.0 / k - 1.0 / n, rel=1e-12)
    assert SG.relative_variance(0.0, 0.0, n) == np.inf
    # photons for 2% at every SDS: the worst SDS decides
    assert SG.photons_for_error(np.array([[100.0, 400.0, 2500.0]]))[0] == pytest.approx(2500 / 4e-4)


def test_white_dataset_layout():
    """Shapes and the identity vector: row j*m of the expansion is set j at its own mua."""
    class FakeCtx:
        def run_batch(self, sims, opts, seed):
            class R:
                reflectance = np.array([[s.number_photons * 1e6] * 3 for s in sims], np.float64)
            return R()

        def white_rescale(self, sims, res, mua, with_squares=False):
            W = np.repeat(res.reflectance[:, None, :], mua.shape[1], axis=1) * np.exp(-mua.sum(axis=2))[..., None]
            return W, W * 1e6

    params = batch.draw_parameters(5, 9)
    x, R, rv = SG.white_dataset(FakeCtx(), params, np.full(5, 1000), 4, seed=3)
    assert x.shape == (20, 14) and R.shape == (20, 3) and rv.shape == (20, 3)
    assert np.array_equal(x[::4, :3], params[:, :3].astype(np.float32).astype(np.float64))
    assert np.array_equal(x[:, 3:], np.repeat(params[:, 3:], 4, axis=0))


@pytest.mark.gpu
def test_surrogate_gains_over_plain_mc(ctx):
    """Train on a reduced cfg3-style set (384 sets x 24 absorption vectors), then on 16
    held-out sets simulated with 2e7 photons each: the network's median relative error is
    below that of plain MC given the same time per evaluation, and plain MC needs over
    1000x the time per evaluation to reach the network's error."""
    t0 = time.time()
    x, R, rv = SG.training_set(ctx, n_sets=384, m=24, seed=101, budget_lo=1e5, budget_hi=1e6)
    print("training samples", len(x), "%.1fs" % (time.time() - t0), flush=True)
    net, norm, hist = SG.fit(x, R, rv, epochs=120, batch_size=256, lr=2e-3, seed=0)
    print("loss first/last epoch %.4g %.4g, train %.0f ms" % (hist[0], hist[-1], net.train_ms), flush=True)
    hp, hR, hrv = SG.reference_set(ctx, 16, 2e7, seed=202)
    assert np.all(hR > 0)
    err_s = SG.relative_error(SG.predict_R(net, norm, hp), hR)
    med_s = float(np.median(err_s))
    # time per surrogate evaluation (batch of 1e5) and plain-MC photons per second
    q = batch.draw_parameters(100000, 303)
    SG.predict_R(net, norm, q[:1000])
    t = time.perf_counter()
    SG.predict_R(net, norm, q)
    t_eval = (time.perf_counter() - t) / len(q)
    n_mc = 2_000_000
    t = time.perf_counter()
    SG.simulate(ctx, hp[:4], np.full(4, n_mc), seed=5, replay=False)
    pps = 4 * n_mc / (time.perf_counter() - t)
    # plain MC at the surrogate's cost per evaluation
    n_eq = max(1, int(round(t_eval * pps)))
    _, _, R_eq = SG.simulate(ctx, hp, np.full(len(hp), n_eq), seed=6, replay=False)
    med_mc = float(np.median(SG.relative_error(R_eq, hR)))
    # photons plain MC needs for the surrogate's median error (per-photon relative variance
    # from the 2e7-photon runs)
    need = float(np.median(np.max(hrv * 2e7, axis=1))) / med_s ** 2
    gain = (need / pps) / t_eval
    print("surrogate median rel err %.4f; %.3g s/eval; MC %.3g photons/s; MC at equal cost (%d photons) "
          "median rel err %.3f; MC photons for the surrogate's error %.3g; gain %.3g x"
          % (med_s, t_eval, pps, n_eq, med_mc, need, gain), flush=True)
    assert med_s < 0.25
    assert med_s < med_mc
    assert gain > 1e3


@pytest.mark.gpu
def test_surrogate_meets_the_error_bound(ctx):
    """The benchmark's recipe at reduced scale: 16384 sets of 2e5 photons, each expanded to
    13 absorption vectors by White MC (212,992 samples), 40 epochs.  On 64 held-out sets
    simulated with 2e7 photons each, the pooled median relative error is within the 2 %
    bound, and so are the medians of the two nearer SDS; the farthest SDS, whose labels
    carry the most Monte Carlo noise, stays within 5 %."""
    params = batch.draw_parameters(16384, seed=31)
    t0 = time.time()
    xs, Rs, rvs = [], [], []
    for i in range(0, len(params), 1024):
        p = params[i:i + 1024]
        x, R, rv = SG.white_dataset(ctx, p, np.full(len(p), 200000), 13, seed=41 + i)
        xs.append(x), Rs.append(R), rvs.append(rv)
    x, R, rv = np.concatenate(xs), np.concatenate(Rs), np.concatenate(rvs)
    print("training samples", len(x), "%.1fs" % (time.time() - t0), flush=True)
    net, norm, hist = SG.fit(x, R, rv, epochs=40, batch_size=1024, lr=2e-3, seed=0)
    print("trained in %.1f s, loss %.4g -> %.4g" % (net.train_ms / 1e3, hist[0], hist[-1]), flush=True)
    hp, hR, hrv = SG.reference_set(ctx, 64, 2e7, seed=303)
    sur = SG.Surrogate(net, norm, bound=0.02)
    med = sur.validate(hp, hR)
    per_sds = np.median(SG.relative_error(SG.predict_R(net, norm, hp), hR), axis=0)
    print("median rel err %.4f, per SDS %s; reference rel SE per SDS %s"
          % (med, np.round(per_sds, 4), np.round(np.median(np.sqrt(hrv), axis=0), 4)), flush=True)
    assert sur.usable and med <= 0.02
    assert np.all(per_sds[:2] <= 0.02) and per_sds[2] <= 0.05
    net.close()


@pytest.mark.gpu
def test_surrogate_falls_back_to_plain_mc(ctx):
    x, R, rv = SG.training_set(ctx, n_sets=32, m=8, seed=111, budget_lo=1e5, budget_hi=2e5)
    net, norm, _ = SG.fit(x, R, rv, epochs=5, batch_size=64, seed=1)
    hp, hR, _ = SG.reference_set(ctx, 4, 1e6, seed=212)
    s = SG.Surrogate(net, norm, bound=1e-9)
    s.validate(hp, hR)
    assert not s.usable
    got, src = s.reflectance(hp, ctx=ctx, photons=100000, seed=17)
    _, _, want = SG.simulate(ctx, hp, np.full(len(hp), 100000), seed=17, replay=False)
    assert src == "mc" and np.array_equal(got, want)
    with pytest.raises(Exception):
        s.reflectance(hp)
    s.bound = 10.0
    assert s.usable
    got, src = s.reflectance(hp)
    assert src == "surrogate" and np.allclose(got, SG.predict_R(net, norm, hp))
