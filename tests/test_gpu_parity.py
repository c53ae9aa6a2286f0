"""Per-photon parity of the HIP kernels with the CPU oracle (GPU tests).

Every photon's random stream is a function of (seed, set, photon index) and both sides use
the same correctly rounded arithmetic, so the HIP first pass must return exactly the
oracle's detection records (same photons, same detectors, bit-identical weights) and the
same step/boundary/interaction counts; the replay must return bit-identical per-layer
pathlengths, step counts and u64 absorbance grids.  The bar stated for floating point
(pathlength within 1e-4 relative) is checked as well, but equality is what is expected.
"""
import ctypes as C

import numpy as np
import pytest

import fixtures as F
import oracle_ffi as O
from mcml_gpu_amd import make_opts

SEED = 20240601

CASES = {
    "hemisphere_g0": lambda: F.hemisphere(20000),
    "hemisphere_g09_n14": lambda: F.hemisphere(20000, mua=1.0, mus=100.0, g=0.9, n=1.4),
    "vdh_slab": lambda: F.vdh_slab(30000),
    "fiber_busy": lambda: F.fiber_busy(60000),
    "ring_busy": lambda: F.ring_busy(60000),
    "HW8_A": lambda: F.example("HW8_A", 40000),
    "head_model": lambda: F.example("head_model", 20000),
    "cfg0": lambda: F.config("cfg0", 20000),
    "cfg1": lambda: F.config("cfg1", 4000),
    "max_stack_fiber": lambda: F.max_stack(20000),
    "max_stack_ring": lambda: F.max_stack(20000, ring=True),
}


def photon_order(rec):
    """Records in the oracle's order: by photon key, then SDS (the library returns them
    grouped by (set, SDS) in photon order, the order of the pathlength files)."""
    return rec[np.lexsort((rec["sds"], rec["key"]))]


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _same_floats(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(nan | (_bits(a) == _bits(b))))


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_first_pass_matches_oracle(ctx, name):
    sim = CASES[name]()
    n = sim.number_photons
    print("case", name, "photons", n, flush=True)
    res = ctx.run(sim, make_opts(), seed=SEED)
    d = O.derive(sim)
    rec, cnt = O.first_pass(d, SEED, 0, n)
    assert res.photons == n
    assert (res.steps, res.steps_boundary, res.steps_interaction) == tuple(int(c) for c in cnt)
    assert len(res.records) == len(rec), (len(res.records), len(rec))
    got = photon_order(res.records)
    assert np.array_equal(got["key"], rec["key"])
    assert np.array_equal(got["sds"], rec["sds"])
    assert _same_floats(got["weight"], rec["weight"])
    # grouped by (set, SDS), photon order inside a group
    grp = res.records["sds"].astype(np.int64) * (1 << 41) + res.records["key"].astype(np.int64)
    assert np.all(np.diff(grp) > 0)
    # per-SDS totals and counts follow from the records
    for s in range(1, sim.num_sds + 1):
        m = rec["sds"] == s
        assert res.detected[0, s - 1] == m.sum()
        if not np.isnan(rec["weight"][m]).any():
            assert res.reflectance[0, s - 1] == pytest.approx(float(rec["weight"][m].astype(np.float64).sum()),
                                                               rel=1e-12, abs=0)


REPLAY_CASES = ["fiber_busy", "ring_busy", "HW8_A", "head_model", "hemisphere_g09_n14", "max_stack_fiber",
                "max_stack_ring"]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["grids_in_pass", "replay_pass", "inline_pathlength"])
@pytest.mark.parametrize("name", REPLAY_CASES)
def test_replay_matches_oracle(ctx, name, mode):
    """With -A the detected photons are replayed into the absorbance maps, by their own lanes
    within the first pass or in a replay pass of their own; without -A the first pass
    records the pathlengths itself.  All must equal the oracle's replay."""
    absorbance = mode != "inline_pathlength"
    sim = CASES[name]()
    sim.number_photons = min(sim.number_photons, 8000)
    sim.grid_nr, sim.grid_nz, sim.grid_dr, sim.grid_dz = 60, 50, 0.02, 0.01
    print("case", name, "photons", sim.number_photons, "mode", mode, flush=True)
    res = ctx.run(sim, make_opts(replay=True, pathlength=True, absorbance=absorbance,
                                 separate_replay=mode == "replay_pass"), seed=SEED)
    d = O.derive(sim)
    L, S = sim.number_layers, sim.num_sds
    nz, nr = 50, 60
    for s in range(1, S + 1):
        rows = res.pathlength_records(0, s)
        keys = np.sort(res.records["key"][res.records["sds"] == s])
        assert len(rows) == len(keys)
        arz = np.zeros(nz * nr, np.uint64)
        a0 = np.zeros(nz, np.uint64)
        ref = np.zeros((len(keys), L + 2), np.float32)
        for i, k in enumerate(keys):
            hit, r = O.replay(d, SEED, int(k), s, grids=(arz, a0))
            assert hit == 1  # the replayed photon is detected again by the same SDS
            ref[i] = r
        assert _same_floats(rows, ref)
        if len(keys):
            pl_gpu, pl_ref = rows[:, 1:L + 1].astype(np.float64), ref[:, 1:L + 1].astype(np.float64)
            assert np.all(np.abs(pl_gpu - pl_ref) <= 1e-4 * np.maximum(np.abs(pl_ref), 1e-30))
        # replayed weights equal the first-pass weights of the same photons
        fw = res.records["weight"][res.records["sds"] == s]
        assert _same_floats(np.sort(rows[:, 0]), np.sort(fw))
        if absorbance:
            assert np.array_equal(res.A_rz[0, s - 1].reshape(-1), arz)
            assert np.array_equal(res.A0_z[0, s - 1], a0)
        else:
            assert res.A_rz is None


@pytest.mark.gpu
def test_batch_matches_single_sets(ctx):
    """Uneven budgets in one persistent launch == each set run by the oracle with its own
    set index."""
    sims = [F.fiber_busy(7000), F.fiber_busy(23000), F.fiber_busy(1), F.fiber_busy(15000)]
    sims[1].layers[0].mua = 0.9
    sims[3].layers[1].mus = 60.0
    res = ctx.run_batch(sims, make_opts(replay=True, pathlength=True), seed=SEED)
    assert res.photons == sum(s.number_photons for s in sims)
    allrec = []
    for j, sim in enumerate(sims):
        rec, _ = O.first_pass(O.derive(sim), SEED, 0, sim.number_photons, set_index=j)
        allrec.append(rec)
        for s in range(1, sim.num_sds + 1):
            assert res.detected[j, s - 1] == (rec["sds"] == s).sum()
    rec = np.concatenate(allrec)
    got = photon_order(res.records)
    assert np.array_equal(got["key"], rec["key"])
    assert np.array_equal(got["sds"], rec["sds"])
    assert _same_floats(got["weight"], rec["weight"])
    # replay rows of a few (set, sds) groups
    for j, sim in enumerate(sims):
        d = O.derive(sim)
        for s in range(1, sim.num_sds + 1):
            rows = res.pathlength_records(j, s)
            keys = np.sort(res.records["key"][(res.records["sds"] == s) & ((res.records["key"] >> 40) == j)])
            for i, key in enumerate(keys[:50]):
                _, r = O.replay(d, SEED, int(key & ((1 << 40) - 1)), s, set_index=j)
                assert _same_floats(rows[i], r)


@pytest.mark.gpu
def test_small_budgets_do_not_hang(ctx):
    """Fewer photons than lanes (the reference hangs here: transport.cu:442 underflows)."""
    for n in (1, 7, 63, 64, 65, 1000):
        sim = F.fiber_busy(n)
        res = ctx.run(sim, make_opts(), seed=SEED)
        assert res.photons == n
        rec, _ = O.first_pass(O.derive(sim), SEED, 0, n)
        assert np.array_equal(photon_order(res.records)["key"], rec["key"])


@pytest.mark.gpu
def test_seed_changes_streams(ctx):
    sim = F.hemisphere(5000)
    a = ctx.run(sim, make_opts(), seed=1)
    b = ctx.run(sim, make_opts(), seed=2)
    c = ctx.run(sim, make_opts(), seed=1)
    assert not np.array_equal(a.records["weight"], b.records["weight"])
    assert np.array_equal(a.records["weight"], c.records["weight"])


@pytest.mark.gpu
@pytest.mark.parametrize("block,chunk,refill,bnd", [(64, 64, 1, 1), (128, 128, 8, 4), (256, 1024, 32, 64),
                                                    (192, 64, 64, 16)])
def test_launch_shape_does_not_change_results(ctx, block, chunk, refill, bnd):
    sim = F.ring_busy(30000)
    opts = make_opts(replay=True, pathlength=True)
    ref = ctx.run(sim, opts, seed=SEED)
    got = ctx.run(sim, make_opts(replay=True, pathlength=True, block_size=block, chunk=chunk, refill_min=refill,
                                 boundary_min=bnd), seed=SEED)
    assert got.block_size == block and got.chunk == chunk and got.boundary_min == bnd
    assert np.array_equal(got.records, ref.records)
    assert got.steps == ref.steps
    assert np.array_equal(got.pathlength, ref.pathlength)


@pytest.mark.gpu
def test_graph_replays_match_eager(ctx):
    """The graph path (reset kernel + transport + read-back) gives the eager results on
    every replay, with eager work and synchronisations in between."""
    sim = F.ring_busy(20000)
    eager = ctx.run(sim, make_opts(), seed=SEED)
    other = F.fiber_busy(5000)
    for i in range(4):
        g = ctx.run(sim, make_opts(use_graph=True), seed=SEED)
        assert np.array_equal(g.records, eager.records) and g.steps == eager.steps and g.photons == eager.photons
        ctx.run(other, make_opts(), seed=i)  # eager work between replays
    g2 = ctx.run(sim, make_opts(use_graph=True), seed=SEED + 1)
    e2 = ctx.run(sim, make_opts(), seed=SEED + 1)
    assert np.array_equal(g2.records, e2.records)


@pytest.mark.gpu
def test_autotune_returns_a_valid_shape(ctx):
    from mcml_gpu_amd import batch
    sims, _, _ = batch.cfg3_sets(16, seed=3, budget_lo=1e3, budget_hi=1e4)
    o = ctx.autotune(sims, trial_photons=200000)
    assert o.block_size in (128, 256) and o.refill_min in (1, 2, 4) and o.chunk in (64, 128, 256)
    assert o.boundary_min in (1, 4)
    a = ctx.run_batch(sims, make_opts(), seed=SEED)
    b = ctx.run_batch(sims, make_opts(block_size=o.block_size, chunk=o.chunk, refill_min=o.refill_min,
                                      boundary_min=o.boundary_min, use_graph=True), seed=SEED)
    assert np.array_equal(a.records, b.records)


@pytest.mark.gpu
def test_cfg3_batch_matches_oracle_on_a_subset(ctx):
    """The cfg3 generator's batch (reduced budgets) against the oracle set by set."""
    from mcml_gpu_amd import batch
    sims, _, budgets = batch.cfg3_sets(24, seed=1, budget_lo=2e3, budget_hi=2e4)
    res = ctx.run_batch(sims, make_opts(), seed=SEED)
    assert res.photons == int(budgets.sum())
    got = res.records
    for j, sim in enumerate(sims):
        rec, _ = O.first_pass(O.derive(sim), SEED, 0, sim.number_photons, set_index=j)
        mine = photon_order(got[(got["key"] >> 40) == j])
        assert np.array_equal(mine["key"], rec["key"]) and _same_floats(mine["weight"], rec["weight"])


@pytest.mark.gpu
def test_batch_absorbance_grids_match_oracle(ctx):
    """-A over a batch: each wavefront's LDS tile of the grids' hot corner passes from one
    (set, SDS) group to the next; every group's u64 grids must still equal the oracle's."""
    sims = [F.ring_busy(3000), F.ring_busy(5000), F.fiber_busy(4000)]
    sims[1].layers[0].mua = 0.8
    sims[2] = F.ring_busy(2000)
    sims[2].layers[1].mus = 60.0
    for s in sims:
        s.grid_nr, s.grid_nz, s.grid_dr, s.grid_dz = 40, 30, 0.01, 0.004
    res = ctx.run_batch(sims, make_opts(replay=True, pathlength=True, absorbance=True), seed=SEED)
    sep = ctx.run_batch(sims, make_opts(replay=True, pathlength=True, absorbance=True, separate_replay=True),
                        seed=SEED)
    assert np.array_equal(sep.A_rz, res.A_rz) and np.array_equal(sep.A0_z, res.A0_z)
    assert np.array_equal(_bits(sep.pathlength), _bits(res.pathlength))
    for j, sim in enumerate(sims):
        d = O.derive(sim)
        for s in range(1, sim.num_sds + 1):
            arz = np.zeros(30 * 40, np.uint64)
            a0 = np.zeros(30, np.uint64)
            keys = np.sort(res.records["key"][(res.records["sds"] == s) & ((res.records["key"] >> 40) == j)])
            for key in keys:
                hit, _ = O.replay(d, SEED, int(key & ((1 << 40) - 1)), s, set_index=j, grids=(arz, a0))
                assert hit == 1
            assert np.array_equal(res.A_rz[j, s - 1].reshape(-1), arz), (j, s)
            assert np.array_equal(res.A0_z[j, s - 1], a0), (j, s)


@pytest.mark.gpu
def test_batch_rejects_mixed_absorbance_grids_and_accepts_empty_sets(ctx):
    from mcml_gpu_amd import MCMLError
    a, b = F.ring_busy(100), F.ring_busy(100)
    b.grid_nr = a.grid_nr + 1
    with pytest.raises(MCMLError):
        ctx.run_batch([a, b], make_opts(replay=True, absorbance=True), seed=SEED)
    # without -A the grids do not matter; sets with no photons are fine
    c = F.ring_busy(0)
    res = ctx.run_batch([a, b, c], make_opts(replay=True, pathlength=True), seed=SEED)
    assert res.photons == 200 and res.detected[2].sum() == 0
    res = ctx.run(F.ring_busy(0), make_opts(), seed=SEED)
    assert res.photons == 0 and len(res.records) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("pathlength", [False, True], ids=["weights", "inline_pathlength"])
def test_record_buffer_pause_and_resume(ctx, pathlength):
    """A record buffer far smaller than the run's records: the first pass pauses when it
    fills, the buffer grows and the pass resumes from the ticket counter.  Records, rows and
    step counts equal the single-launch run's and the oracle's."""
    sim = F.hemisphere(300000, mua=1.0, mus=100.0, g=0.9, n=1.4)
    base = dict(replay=pathlength, pathlength=pathlength, block_size=64, blocks_per_cu=1, chunk=64)
    one = ctx.run(sim, make_opts(**base), seed=SEED)
    assert one.first_pass_launches == 1
    small = ctx.run(sim, make_opts(record_cap=60000, **base), seed=SEED)
    print("records", len(one.records), "launches with a small buffer", small.first_pass_launches, flush=True)
    assert len(one.records) > 100000 and small.first_pass_launches >= 2
    assert np.array_equal(small.records, one.records)
    assert (small.photons, small.steps, small.steps_boundary) == (one.photons, one.steps, one.steps_boundary)
    if pathlength:
        assert np.array_equal(_bits(small.pathlength), _bits(one.pathlength))
    rec, cnt = O.first_pass(O.derive(sim), SEED, 0, sim.number_photons)
    assert small.steps == int(cnt[0])
    got = photon_order(small.records)
    assert np.array_equal(got["key"], rec["key"]) and _same_floats(got["weight"], rec["weight"])


@pytest.mark.gpu
def test_lane_statistics_variant_same_results(ctx):
    """The statistics variant of the first pass returns the same records and counters, and
    lane-blocks that account for every stepped photon step."""
    sim = F.ring_busy(200000)
    opts = dict(replay=True, pathlength=True)
    ref = ctx.run(sim, make_opts(**opts), seed=SEED)
    st = ctx.run(sim, make_opts(lane_stats=True, **opts), seed=SEED)
    assert np.array_equal(st.records, ref.records) and st.steps == ref.steps
    assert np.array_equal(_bits(st.pathlength), _bits(ref.pathlength))
    lb = st.lane_blocks.astype(np.int64)
    print("lane blocks", lb, "drain ms", st.ms_drain, "clock ms", st.ms_clock, flush=True)
    assert lb[0] > 0 and lb[0] * 12 >= st.steps  # a stepping lane takes at most 12 steps per block
    assert 0 < st.ms_drain < st.ms_clock
    assert ref.lane_blocks.sum() == 0 and ref.ms_drain == 0


@pytest.mark.gpu
def test_device_reciprocal_and_square_root_are_correctly_rounded(ctx):
    """The spin's reciprocal and square root (hardware units plus one correction) equal IEEE
    1.0f / x and sqrtf(x) on every value the self-test covers (about 2.3e9)."""
    bad = (C.c_uint64 * 2)()
    assert ctx.lib.mcml_selftest_math(ctx.h, bad) == 0
    assert list(bad) == [0, 0]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ring_busy", "max_stack_fiber"])
def test_drain_compaction_does_not_change_results(ctx, name):
    """Photons handed from sparse wavefronts to compacted follow-up launches continue exactly
    where they stopped: records, pathlength rows and step counts equal those of a pass with
    the drain compaction off, and the oracle's."""
    sim = CASES[name]()
    sim.number_photons = 600000 if name == "ring_busy" else 200000
    opts = dict(replay=True, pathlength=True)
    off = ctx.run(sim, make_opts(spill_below=-1, **opts), seed=SEED)
    assert off.drain_launches == 0
    for sb in (0, 63):
        on = ctx.run(sim, make_opts(spill_below=sb, **opts), seed=SEED)
        print(name, "spill_below", sb, "drain launches", on.drain_launches, flush=True)
        assert on.drain_launches >= 1
        assert np.array_equal(on.records, off.records)
        assert (on.photons, on.steps, on.steps_boundary, on.steps_interaction) == \
            (off.photons, off.steps, off.steps_boundary, off.steps_interaction)
        assert np.array_equal(_bits(on.pathlength), _bits(off.pathlength))
    rec, cnt = O.first_pass(O.derive(sim), SEED, 0, sim.number_photons)
    assert off.steps == int(cnt[0])
    got = photon_order(on.records)
    assert np.array_equal(got["key"], rec["key"]) and _same_floats(got["weight"], rec["weight"])


@pytest.mark.gpu
def test_absorbance_with_many_detections_takes_the_replay_pass(ctx):
    """More detections than the pass has lanes: the first pass records the rows and the replay
    pass books the grids.  Rows and grids equal those of the separate replay path (which
    the smaller cases pin to the oracle), and the oracle's replays of the first photons'
    detections land in the same cells."""
    sim = F.hemisphere(1_500_000, mua=1.0, mus=100.0, g=0.9, n=1.4)
    sim.grid_nr, sim.grid_nz, sim.grid_dr, sim.grid_dz = 40, 30, 0.01, 0.004
    opts = dict(replay=True, pathlength=True, absorbance=True)
    res = ctx.run(sim, make_opts(**opts), seed=SEED)
    print("records", len(res.records), "replay kernel ms", res.ms_replay_kernel, flush=True)
    assert len(res.records) > 524288 and res.ms_replay_kernel > 0
    sep = ctx.run(sim, make_opts(separate_replay=True, **opts), seed=SEED)
    assert np.array_equal(res.A_rz, sep.A_rz) and np.array_equal(res.A0_z, sep.A0_z)
    assert np.array_equal(_bits(res.pathlength), _bits(sep.pathlength))
    assert np.array_equal(res.records, sep.records)
