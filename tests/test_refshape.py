"""The reference launch shape (mcml_opts.reference_shape) on the GPU.

That mode runs the photon step in the reference's own launch structure (transport.cu:76-147,
341-471): one photon per thread, a fixed step budget per launch with the photon state saved
to global memory between launches, one atomic ticket per relaunched photon, and per-thread
detection buffers that the host reads back and reduces after each launch.  Photon k still
draws from stream (seed, 0, k), so its records, sums and work counters must be exactly
those of the persistent kernel (and of the oracle), whatever the step budget, including
runs with more photons than threads (relaunch tickets), budgets that cut photons across
launches, and photons whose detections overflow a thread's 20-slot buffer.
"""
import numpy as np
import pytest

import fixtures as F
import oracle_ffi as O
from mcml_gpu_amd import MCMLError, abi, make_opts

SEED = 77


def _photon_order(rec):
    return rec[np.lexsort((rec["sds"], rec["key"]))]


def _same_records(a, b):
    assert a.shape == b.shape
    assert np.array_equal(a["key"], b["key"])
    assert np.array_equal(a["sds"], b["sds"])
    assert np.array_equal(a["weight"].view(np.uint32), b["weight"].view(np.uint32))


def _same_run(r, p):
    _same_records(r.records, p.records)
    assert np.array_equal(r.detected, p.detected)
    assert np.array_equal(r.reflectance, p.reflectance)
    for f in ("photons", "steps", "steps_boundary", "steps_interaction"):
        assert getattr(r, f) == getattr(p, f), f


def stacked_rings(photons):
    """30 identical ring detectors (header.h:44) that accept every photon leaving within
    0.4 cm of the source: each detected photon yields 30 records, more than a thread's
    20-slot buffer holds."""
    dets = [(0.2, 1.0, 0.2, 0.0)] * 30
    cfg = F.base_cfg(photons, 1, 1.0, 1.0, 0, 0, F.probes((0.0, 0.0, 0.0), dets))
    return abi.make_sim(cfg, [(0.0, 5.0, 95.0, 1.0, 0.0)])


@pytest.mark.gpu
@pytest.mark.parametrize("steps", [0, 200])
def test_matches_oracle_and_persistent(ctx, steps):
    sim = F.fiber_busy(60000)
    print("refshape fiber_busy steps", steps, flush=True)
    r = ctx.run(sim, make_opts(reference_shape=True, chunk=steps), seed=SEED)
    p = ctx.run(sim, make_opts(), seed=SEED)
    _same_run(r, p)
    rec, cnt = O.first_pass(O.derive(sim), SEED, 0, sim.number_photons)
    _same_records(_photon_order(r.records), rec)
    assert r.steps == int(cnt[0])
    assert r.block_size == 32
    if steps:
        assert r.first_pass_launches > 10  # photons of ~100s of steps cut across launches
    else:
        assert r.first_pass_launches >= 1


@pytest.mark.gpu
def test_more_photons_than_threads(ctx):
    """1.2e6 short photons on 2048 x CUs threads: every thread relaunches by ticket."""
    sim = F.hemisphere(1_200_000, mua=10.0, mus=90.0)
    print("refshape hemisphere 1.2e6", flush=True)
    r = ctx.run(sim, make_opts(reference_shape=True), seed=SEED)
    p = ctx.run(sim, make_opts(), seed=SEED)
    _same_run(r, p)
    assert r.grid_size * r.block_size < sim.number_photons


@pytest.mark.gpu
@pytest.mark.parametrize("block", [32, 64, 256])
def test_buffer_overflow_kept(ctx, block):
    """Detections beyond a thread's 20 slots go to the spill list (the reference drops
    them, transport.cu:899-904); counts and weights stay those of the persistent kernel."""
    sim = stacked_rings(20000)
    print("refshape stacked rings block", block, flush=True)
    r = ctx.run(sim, make_opts(reference_shape=True, block_size=block), seed=SEED)
    p = ctx.run(sim, make_opts(), seed=SEED)
    _same_run(r, p)
    assert r.block_size == block
    det = r.detected[0]
    assert det.min() == det.max() and det[0] > 1000  # every detector sees every detection


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [dict(pathlength=True), dict(absorbance=True, pathlength=True)])
def test_replay_after_reference_shape(ctx, flags):
    """-R -P / -R -A: the replay that follows gives the persistent path's rows and grids."""
    sim = F.ring_busy(40000)
    print("refshape replay", flags, flush=True)
    r = ctx.run(sim, make_opts(replay=True, reference_shape=True, **flags), seed=SEED)
    p = ctx.run(sim, make_opts(replay=True, **flags), seed=SEED)
    _same_run(r, p)
    assert r.replayed and p.replayed
    assert np.array_equal(r.pl_offset, p.pl_offset)
    assert np.array_equal(r.pathlength.view(np.uint32), p.pathlength.view(np.uint32))
    assert np.array_equal(r.replay_reflectance, p.replay_reflectance)
    if flags.get("absorbance"):
        assert np.array_equal(r.A_rz, p.A_rz)
        assert np.array_equal(r.A0_z, p.A0_z)
    else:
        assert r.A_rz is None


@pytest.mark.gpu
def test_one_set_only(ctx):
    sims = [F.fiber_busy(1000), F.fiber_busy(2000)]
    with pytest.raises(MCMLError, match="one parameter set"):
        ctx.run_batch(sims, make_opts(reference_shape=True), seed=SEED)
    with pytest.raises(MCMLError, match="multiple of 32"):
        ctx.run(sims[0], make_opts(reference_shape=True, block_size=48), seed=SEED)
