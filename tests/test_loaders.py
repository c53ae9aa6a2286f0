"""NumPy loaders (mcml_gpu_amd.loaders) read back what the writers produce (CPU)."""
import ctypes as C

import numpy as np
import pytest

import fixtures as F
import mcml_gpu_amd as M
from mcml_gpu_amd import abi, loaders
from test_io import FakeResult


def write(tmp_path, binary, calls=2):
    L, S, N = 3, 2, 1000
    cfg = F.base_cfg(N, L, 1.457, 1.457, 1, 1, F.probes((0.37, 0.075, 0), [(0.8, .12, .02, 0), (1.5, .12, .02, 0)]))
    sim = abi.make_sim(cfg, [(0.1, 0.3, 30, 1.4, 0.9), (0.2, 0.2, 20, 1.4, 0.9), (0, 0.25, 25, 1.4, 0.9)])
    sim.grid_nr, sim.grid_nz, sim.grid_dr, sim.grid_dz = 3, 4, 0.02, 0.005
    fr = FakeResult(S, L, [1.2345678e9, 5.5e6], [7, 3], [7, 3])
    opts = M.make_opts(replay=True, absorbance=True, pathlength=True, avg_pathlength=True, binary=binary)
    info = abi.SummaryInfo(1.5, 0.25, b"AMD Instinct MI355X")
    lib = M.load_library()
    for _ in range(calls):
        assert lib.mcml_write_outputs(C.byref(sim), C.byref(opts), C.byref(fr.r), 0, str(tmp_path / "output.txt").encode(),
                                      str(tmp_path).encode(), C.byref(info)) == 0
    return sim, fr


@pytest.mark.parametrize("binary", [True, False])
def test_loaders_read_the_writers(tmp_path, binary):
    sim, fr = write(tmp_path, int(binary))
    summaries = loaders.read_summaries(tmp_path / "summary.json")
    assert len(summaries) == 2 and summaries[-1]["SDS_detected_number"] == [7, 3]
    scale1 = float(0xFFFFFFFF) * 1000
    for s in (1, 2):
        rows = fr.pl[fr.off[s - 1]:fr.off[s]]
        if binary:
            got = loaders.load_binary_pathlength(tmp_path / "summary.json", s, tmp_path / ("pathlength_SDS_%d.bin" % s))
            assert np.array_equal(got, rows)
        else:
            got = loaders.load_text_pathlength(tmp_path / ("pathlength_SDS_%d.txt" % s))
            assert got.shape == (2 * len(rows), 5)  # appended twice
            assert np.allclose(got[:len(rows)], rows, rtol=1e-5)
        # A_rz' with A0_z in the r = 0 column, normalised as the writer does (io.cpp:383-469)
        if binary:
            A = loaders.load_A_arr(tmp_path, s, nz=4, nr=3)
            PI, dr, dz = float(np.float32(3.141592654)), float(np.float32(0.02)), float(np.float32(0.005))
            vol = scale1 * 2 * PI * (np.arange(3) + 0.5) * dr * dr * dz
            want = (fr.arz[s - 1].astype(np.float64) / vol).astype(np.float32).astype(np.float64)
            want[:, 0] += (fr.a0[s - 1].astype(np.float64) / vol[0]).astype(np.float32)
            assert np.allclose(A, want, rtol=1e-12)
    out = loaders.load_output(tmp_path / "output.txt")
    assert len(out) == 2 and np.allclose(out[0], np.array([1.2345678e9, 5.5e6]) / scale1, rtol=1e-5)
    avg = loaders.load_average_pathlength(tmp_path / "average_pathlength.txt", 2, 3)
    assert avg.shape == (2, 2, 3) and np.all(np.isfinite(avg))


def test_text_absorbance_map_matches_binary(tmp_path):
    (tmp_path / "b").mkdir()
    (tmp_path / "t").mkdir()
    write(tmp_path / "b", 1, calls=1)
    write(tmp_path / "t", 0, calls=1)
    for s in (1, 2):
        a = loaders.load_A_arr(tmp_path / "b", s, nz=4, nr=3)
        b = loaders.load_A_arr(tmp_path / "t", s, nz=4, nr=3, binary=False)
        assert np.allclose(a, b, rtol=1e-5)


def test_loader_checks_sizes(tmp_path):
    write(tmp_path, 1, calls=1)
    with pytest.raises(ValueError):
        loaders.load_A_arr(tmp_path, 1)  # default 500 x 1000 grid does not match
    with pytest.raises(ValueError):
        loaders.load_binary_pathlength(tmp_path / "summary.json", 3, tmp_path / "pathlength_SDS_1.bin")
