"""mcml_average_pathlength (host code of the library, no device needed) against a float64
restatement of calculate_average_pathlength (transport.cu:274-297) and the mathematics of
its perturbation estimate: ln(sum w / sum w e^{-eps L}) / eps = <L>_w - eps Var_w(L) / 2 + ...
so it sits within eps Var_w(L) of the weighted mean pathlength."""
import ctypes as C

import numpy as np

import fixtures as F
from mcml_gpu_amd import abi, derive, load_library


def test_average_pathlength_formula_and_limit():
    lib = load_library()
    sim = F.config("cfg1", 1000)
    L, S = sim.number_layers, sim.num_sds
    rng = np.random.default_rng(3)
    counts = [50, 0, 7, 120, 3]
    blocks = []
    for c in counts:
        w = rng.integers(1_000_000, 4_000_000_000, c).astype(np.float32)
        pl = rng.gamma(2.0, 0.5, (c, L)).astype(np.float32)
        steps = rng.integers(10, 1000, c).astype(np.float32)
        blocks.append(np.column_stack([w, pl, steps]).astype(np.float32))
    rows = np.ascontiguousarray(np.concatenate(blocks), np.float32)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    r = abi.Result()
    r.nsets, r.num_sds, r.num_layers, r.replayed = 1, S, L, 1
    r.pl_offset = off.ctypes.data_as(C.POINTER(C.c_uint64))
    r.pathlength = rows.ctypes.data_as(C.POINTER(C.c_float))
    out = np.zeros(S * L)
    assert lib.mcml_average_pathlength(C.byref(sim), C.byref(r), 0, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    d = derive(sim)
    inv = float(np.float32(1.0) / np.float32(d.start_weight))
    for s in range(S):
        blk = rows[off[s]:off[s + 1]].astype(np.float64)
        for l in range(1, L + 1):
            eps = 0.001 * float(np.float32(d.layer[l].mua))
            e = blk[:, 0] * inv
            with np.errstate(invalid="ignore", divide="ignore"):
                want = np.log(e.sum() / (e * np.exp(-eps * blk[:, l])).sum()) / eps
            got = out[s * L + l - 1]
            if counts[s] == 0:
                assert np.isnan(got) and np.isnan(want)
                continue
            # the log of a ratio near 1 amplifies summation-order rounding by about 1/eps
            assert abs(got - want) <= 1e-9 * abs(want), (s, l, got, want)
            mean = (blk[:, 0] * blk[:, l]).sum() / blk[:, 0].sum()
            var = (blk[:, 0] * (blk[:, l] - mean) ** 2).sum() / blk[:, 0].sum()
            assert abs(got - mean) <= eps * var + 1e-9, (got, mean, eps * var)


def test_average_pathlength_many_records_on_threads():
    """Above 1e5 records the (SDS, layer) pairs run on several host threads; every pair's
    sums still run over its records in order, so the values equal a sequential float64
    restatement (the sums taken left to right with cumsum)."""
    lib = load_library()
    sim = F.config("cfg1", 1000)
    L, S = sim.number_layers, sim.num_sds
    rng = np.random.default_rng(11)
    counts = [90000, 0, 150000, 40000, 60000]
    rows = np.column_stack([rng.integers(1_000_000, 4_000_000_000, sum(counts)).astype(np.float32),
                            rng.gamma(2.0, 0.5, (sum(counts), L)).astype(np.float32),
                            rng.integers(10, 1000, sum(counts)).astype(np.float32)])
    rows = np.ascontiguousarray(rows, np.float32)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    r = abi.Result()
    r.nsets, r.num_sds, r.num_layers, r.replayed = 1, S, L, 1
    r.pl_offset = off.ctypes.data_as(C.POINTER(C.c_uint64))
    r.pathlength = rows.ctypes.data_as(C.POINTER(C.c_float))
    out = np.zeros(S * L)
    assert lib.mcml_average_pathlength(C.byref(sim), C.byref(r), 0, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    d = derive(sim)
    inv = float(np.float32(1.0) / np.float32(d.start_weight))
    for s in range(S):
        blk = rows[off[s]:off[s + 1]].astype(np.float64)
        for l in range(1, L + 1):
            got = out[s * L + l - 1]
            if counts[s] == 0:
                assert np.isnan(got)
                continue
            eps = 0.001 * float(np.float32(d.layer[l].mua))
            e = blk[:, 0] * inv
            want = np.log(np.cumsum(e)[-1] / np.cumsum(e * np.exp(-eps * blk[:, l]))[-1]) / eps
            assert abs(got - want) <= 1e-9 * abs(want), (s, l, got, want)
