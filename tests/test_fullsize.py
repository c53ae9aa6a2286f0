"""Checks at the benchmark's full sizes (GPU): the 5-layer slab and its absorbance-map run at
1e8 photons, and a detection-heavy case at 1e8 photons.

The oracle cannot run 1e8 photons within a test, but a photon's random stream depends only on
(seed, set, photon index), so parts of a full-size run can be checked photon by photon:
* every photon the GPU reports as detected is run again by the oracle on its own (photon k
  alone), which must detect it by the same detectors with bit-identical weights, and the
  oracle's replay of it must give the GPU's pathlength row bit for bit;
* with -A, the absorbance maps hold the drops of the detected photons only, so the oracle's
  replays of all of them must add up to the GPU's u64 grids exactly;
* the records of the first photons of a detection-heavy full-size run must equal the oracle's
  first pass over the same photon range (no detection missed or added);
* a second run with another launch shape must return identical records and identical
  step, boundary and interaction totals (8e10 steps on the slab);
* every detected photon has travelled at least as far as its detector is from the source.
"""
import numpy as np
import pytest

import fixtures as F
import oracle_ffi as O
from mcml_gpu_amd import make_opts

SEED = 1
FULL = 100_000_000


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _confirm_each_record(res, sim, d, grids=None):
    """Run every detected photon through the oracle on its own; compare records and rows."""
    L = sim.number_layers
    keys = np.unique(res.records["key"])
    for k in keys:
        mine = res.records[res.records["key"] == k]
        rec, _ = O.first_pass(d, SEED, int(k), int(k) + 1)
        assert np.array_equal(np.sort(rec["sds"]), np.sort(mine["sds"])), int(k)
        for r in mine:
            o = rec[rec["sds"] == r["sds"]][0]
            assert _bits(o["weight"]) == _bits(r["weight"]), (int(k), int(r["sds"]))
    for s in range(1, sim.num_sds + 1):
        rows = res.pathlength_records(0, s)
        ks = np.sort(res.records["key"][res.records["sds"] == s])
        assert len(rows) == len(ks)
        for i, k in enumerate(ks):
            g = None if grids is None else (grids[0][s - 1], grids[1][s - 1])
            hit, ref = O.replay(d, SEED, int(k), s, grids=g)
            assert hit == 1
            assert np.array_equal(_bits(rows[i]), _bits(ref)), (int(k), s)
        if len(rows):
            assert rows.shape[1] == L + 2


@pytest.mark.gpu
def test_slab_full_size(ctx):
    """The benchmark workload itself: 5-layer slab, 1e8 photons, -R -P."""
    sim = F.config("cfg1")
    assert sim.number_photons == FULL
    print("full size: cfg1 slab 1e8 photons -R -P", flush=True)
    res = ctx.run(sim, make_opts(replay=True, pathlength=True), seed=SEED)
    assert res.photons == FULL
    assert len(res.records) > 0
    d = O.derive(sim)
    _confirm_each_record(res, sim, d)
    # per-SDS totals follow from the records
    for s in range(1, sim.num_sds + 1):
        w = res.records["weight"][res.records["sds"] == s].astype(np.float64)
        assert res.detected[0, s - 1] == len(w)
        assert res.reflectance[0, s - 1] == pytest.approx(float(w.sum()), rel=1e-12, abs=0)
    # every step is a boundary hit or an interaction (the step's uniform is below 1)
    assert res.steps == res.steps_boundary + res.steps_interaction
    # a detected photon started at the origin and left through detector s's footprint
    for s in range(1, sim.num_sds + 1):
        rows = res.pathlength_records(0, s)
        det = sim.detectors[s - 1]
        # fibre footprint ((x - pos) cos a)^2 + y^2 <= r^2 lies beyond pos - r / cos a
        nearest = det.pos - det.radius / np.cos(np.radians(det.angle_deg))
        if len(rows):
            travelled = rows[:, 1:sim.number_layers + 1].astype(np.float64).sum(axis=1)
            assert np.all(travelled >= nearest * (1 - 1e-6))
            assert np.all(rows[:, -1] >= 1)
    # another launch shape: the same photons, the same trajectories
    print("full size: cfg1 slab again with block 128, refill 1, boundary 1", flush=True)
    res2 = ctx.run(sim, make_opts(replay=True, pathlength=True, block_size=128, refill_min=1, boundary_min=1,
                                  chunk=256), seed=SEED)
    assert (res2.steps, res2.steps_boundary, res2.steps_interaction) == \
        (res.steps, res.steps_boundary, res.steps_interaction)
    assert np.array_equal(res2.records, res.records)
    assert np.array_equal(_bits(res2.pathlength), _bits(res.pathlength))


@pytest.mark.gpu
def test_slab_absorbance_full_size(ctx):
    """The absorbance-map workload: the slab at 1e8 photons, -R -A (-P to return the replay's
    rows as well) into 400 x 200 grids.  The grids hold the drops of the detected photons
    only, so the oracle's replays of those photons must reproduce them exactly."""
    sim = F.config("cfg2")
    assert sim.number_photons == FULL
    print("full size: cfg2 slab 1e8 photons -R -A -P", flush=True)
    res = ctx.run(sim, make_opts(replay=True, absorbance=True, pathlength=True), seed=SEED)
    assert res.photons == FULL
    assert len(res.records) > 0
    d = O.derive(sim)
    S, nz, nr = sim.num_sds, sim.grid_nz, sim.grid_nr
    arz = np.zeros((S, nz * nr), np.uint64)
    a0 = np.zeros((S, nz), np.uint64)
    _confirm_each_record(res, sim, d, grids=(arz, a0))
    assert np.array_equal(res.A_rz[0].reshape(S, -1), arz)
    assert np.array_equal(res.A0_z[0], a0)
    assert arz.sum() + a0.sum() > 0


@pytest.mark.gpu
def test_detection_heavy_full_size(ctx):
    """1e8 photons of the detection-heavy fibre case (millions of records): the records of
    photons [0, 1e5) equal the oracle's first pass over that range."""
    sim = F.fiber_busy(FULL)
    print("full size: fiber_busy 1e8 photons", flush=True)
    res = ctx.run(sim, make_opts(), seed=SEED)
    assert res.photons == FULL
    K = 100_000
    rec, _ = O.first_pass(O.derive(sim), SEED, 0, K)
    got = res.records[res.records["key"] < K]
    got = got[np.lexsort((got["sds"], got["key"]))]
    assert len(got) == len(rec) and len(rec) > 1000
    assert np.array_equal(got["key"], rec["key"])
    assert np.array_equal(got["sds"], rec["sds"])
    assert np.array_equal(_bits(got["weight"]), _bits(rec["weight"]))
    # the full run detects about FULL / K times as many photons as the range does
    ratio = len(res.records) / len(rec) / (FULL / K)
    assert 0.9 < ratio < 1.1


@pytest.mark.gpu
def test_readme_example_full_size_photon_by_photon(ctx):
    """The README example (3 layers, fibres at 0.8/1.5/3.0 cm) at its full 1e6 photons and
    seed 1: the records and step counts equal the oracle's over all 1e6 photons."""
    sim = F.config("cfg0")
    assert sim.number_photons == 1_000_000
    print("full size: README example 1e6 photons", flush=True)
    res = ctx.run(sim, make_opts(replay=True, pathlength=True), seed=SEED)
    d = O.derive(sim)
    rec, cnt = O.first_pass(d, SEED, 0, sim.number_photons)
    assert res.photons == sim.number_photons
    assert (res.steps, res.steps_boundary, res.steps_interaction) == tuple(int(c) for c in cnt)
    got = res.records[np.lexsort((res.records["sds"], res.records["key"]))]
    assert np.array_equal(got["key"], rec["key"]) and np.array_equal(got["sds"], rec["sds"])
    assert np.array_equal(_bits(got["weight"]), _bits(rec["weight"]))
    _confirm_each_record(res, sim, d)


def _group_bounds(records, nsets, S, k_limit):
    """Sorted records: start of every (set, SDS) group and the end of its photons < k_limit."""
    key = records["key"]
    order = ((key >> np.uint64(40)) << np.uint64(45)) | (records["sds"].astype(np.uint64) << np.uint64(40)) | \
        (key & np.uint64((1 << 40) - 1))
    assert np.all(order[1:] > order[:-1])  # (set, SDS, photon) order, no duplicates
    g = np.arange(nsets * S, dtype=np.uint64)
    base = ((g // np.uint64(S)) << np.uint64(45)) | ((g % np.uint64(S) + np.uint64(1)) << np.uint64(40))
    return np.searchsorted(order, base), np.searchsorted(order, base + np.uint64(k_limit))


@pytest.mark.gpu
def test_batch_512_sets_full_size(ctx):
    """The 512-set batch at its real size (budgets log-uniform 1e5-1e7, about 9.8e8 photons)
    in one persistent launch: every ticket runs once, per-(set, SDS) counts and sums follow
    from the records, the records of photons [0, 1000) of every set equal the oracle's first
    pass with that set index, and the autotuned launch shape through the HIP graph returns
    identical records and step totals."""
    from mcml_gpu_amd import batch
    sims, _, budgets = batch.cfg3_sets(512, seed=1)
    S = sims[0].num_sds
    print("full size: 512-set batch, %d photons" % int(budgets.sum()), flush=True)
    res = ctx.run_batch(sims, make_opts(), seed=SEED)
    assert res.photons == int(budgets.sum()) and res.first_pass_launches == 1
    recs = res.records
    lo, hi = _group_bounds(recs, 512, S, 1000)
    ends = np.append(lo[1:], len(recs))
    assert np.array_equal((ends - lo).reshape(512, S), res.detected)
    nonempty = ends > lo  # groups are contiguous: a non-empty group ends where the next one starts
    sums = np.zeros(len(lo))
    sums[nonempty] = np.add.reduceat(recs["weight"].astype(np.float64), lo[nonempty])
    assert np.allclose(sums.reshape(512, S), res.reflectance, rtol=1e-9, atol=0)
    K = 1000
    for j, sim in enumerate(sims):
        rec, _ = O.first_pass(O.derive(sim), SEED, 0, K, set_index=j)
        mine = np.concatenate([recs[lo[j * S + s]:hi[j * S + s]] for s in range(S)])
        mine = mine[np.lexsort((mine["sds"], mine["key"]))]
        assert len(mine) == len(rec), j
        assert np.array_equal(mine["key"], rec["key"]) and np.array_equal(mine["sds"], rec["sds"]), j
        assert np.array_equal(_bits(mine["weight"]), _bits(rec["weight"])), j
    print("full size: 512-set batch again, autotuned shape through the graph", flush=True)
    o = ctx.autotune(sims, trial_photons=2e7)
    res2 = ctx.run_batch(sims, make_opts(block_size=o.block_size, chunk=o.chunk, refill_min=o.refill_min,
                                         boundary_min=o.boundary_min, use_graph=True), seed=SEED)
    assert (res2.steps, res2.steps_boundary, res2.steps_interaction) == \
        (res.steps, res.steps_boundary, res.steps_interaction)
    assert np.array_equal(res2.records, recs)


@pytest.mark.gpu
def test_accept_all_detector_one_launch_full_size(ctx):
    """A detector that accepts every escaping photon, at 1e8 photons (about 4e7 records):
    the record buffer is sized from the most records the photons can produce, so the pass
    takes one launch; the records of photons [0, 1e5) equal the oracle's."""
    sim = F.hemisphere(FULL, mua=1.0, mus=100.0, g=0.9, n=1.4)
    print("full size: accept-all hemisphere 1e8 photons", flush=True)
    res = ctx.run(sim, make_opts(), seed=SEED)
    assert res.photons == FULL and res.first_pass_launches == 1
    assert len(res.records) > FULL // 10
    K = 100_000
    rec, _ = O.first_pass(O.derive(sim), SEED, 0, K)
    got = res.records[res.records["key"] < K]
    assert np.array_equal(got["key"], rec["key"]) and np.array_equal(_bits(got["weight"]), _bits(rec["weight"]))
