"""Input parsing and output writers of the operator surface (CPU, no GPU needed).

Writers are checked byte for byte against expectations built here from the reference's
formats (io.cpp:11-33, 281-469): C++ ostream default formatting is printf's %g with
precision 6, which Python's '%g' reproduces; binary files are little-endian float32."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import fixtures as F
import mcml_gpu_amd as M
from mcml_gpu_amd import abi

GOLD = F.GOLDEN


def test_read_examples():
    sims = M.read_input(os.path.join(GOLD, "HW8_A", "sim_setup.json"), os.path.join(GOLD, "HW8_A", "GPUMC_input.txt"))
    assert len(sims) == 1
    s = sims[0]
    assert s.number_photons == 1000000 and s.number_layers == 2 and s.num_sds == 5
    assert s.source_probe_oblique == 0 and s.detector_probe_oblique == 0
    assert s.upper_n == 1.0 and s.lower_n == 1.0
    assert [s.detectors[i].pos for i in range(5)] == pytest.approx([0.0, 0.01, 0.02, 0.05, 0.1])
    assert (s.layers[0].thickness, s.layers[0].mua, s.layers[0].mus) == pytest.approx((0.005, 37, 480))
    assert (s.layers[1].thickness, s.layers[1].mua, s.layers[1].mus, s.layers[1].g) == pytest.approx((0, 2.2, 220, 0.79))
    h = M.read_input(os.path.join(GOLD, "head_model", "sim_setup.json"),
                     os.path.join(GOLD, "head_model", "GPUMC_input.txt"))[0]
    assert h.number_layers == 5 and h.num_sds == 3 and h.number_photons == 100000000
    assert h.layers[4].mus == pytest.approx(971.1992) and h.layers[3].thickness == pytest.approx(0.4)
    assert h.source.NA == pytest.approx(0.37) and h.source.radius == pytest.approx(0.075)


def _write_case(tmp_path, cfg, line):
    j = tmp_path / "sim.json"
    t = tmp_path / "in.txt"
    j.write_text(json.dumps(cfg))
    t.write_text(line)
    return str(j), str(t)


def _cfg(**kw):
    c = F.base_cfg(1000, 2, 1.457, 1.457, 1, 1, F.probes((0.37, 0.075, 0), [(0.8, 0.12, 0.02, 0)]))
    c.update(kw)
    return c


@pytest.mark.parametrize("kw,line,msg", [
    (dict(source_probe_oblique=2), "0.1 0.1 10 1.4 0.9 0.1 10 1.4 0.9", "source_probe_oblique: 2is not 1 or 0!"),
    (dict(detector_probe_oblique=3), "0.1 0.1 10 1.4 0.9 0.1 10 1.4 0.9", "detector_probe_oblique: 3is not 1 or 0!"),
    (dict(detector_reflectance=1.5), "0.1 0.1 10 1.4 0.9 0.1 10 1.4 0.9", "Detector reflectance: 1.5 is out of range !"),
    (dict(number_layers=11), "", "Number of layer is too large!\nThis program only allow number lower than 10 !"),
    (dict(probes=F.probes((0, 0, 0), [(1, .1, .1, 0)] * 31)), "", "Number of SDS is too large!\nThis program only allow number lower than 30 !"),
    (dict(probes=dict(F.probes((0, 0, 0), [(1, .1, .1, 0)] * 2), num_SDS=3)), "",
     "Number of SDS : 3 not match the probe setting : 2 !"),
    (dict(), "0.1 0.1 10 1.4 0.9 0.1 0 1.4 0.9", "Mus for simulation 0 layer 1 Error with value 0 !"),
    (dict(), "0.1 0.1 10 1.4 0.9 0.1 10 1.0 0.9", "n for simulation 0 layer 1 Error with value 1 !"),
    (dict(), "0.1 0.1 10 1.4 0.9 0.1 10 1.4 1.5", "g for simulation 0 layer 1 Error with value 1.5 !"),
    (dict(number_simulation=2), "0.1 0.1 10 1.4 0.9 0.1 10 1.4 0.9", "Mus for simulation 1 layer 1 Error with value 0 !"),
])
def test_rejections_use_reference_messages(tmp_path, kw, line, msg):
    j, t = _write_case(tmp_path, _cfg(**kw), line)
    with pytest.raises(M.MCMLError) as e:
        M.read_input(j, t)
    assert e.value.code == abi.ERR_INPUT
    assert str(e.value).endswith(msg)


def test_only_last_layer_is_validated(tmp_path):
    # io.cpp:149-160 checks the last layer only: a zero mus in layer 1 is accepted
    j, t = _write_case(tmp_path, _cfg(), "0.1 0.1 0 1.0 5.0 0.1 10 1.4 0.9")
    s = M.read_input(j, t)[0]
    assert s.layers[0].mus == 0 and s.layers[0].g == 5.0


def test_multiple_lines(tmp_path):
    j, t = _write_case(tmp_path, _cfg(number_simulation=3),
                       "0.1 0.1 10 1.4 0.9 0.2 20 1.4 0.9\n0.2 0.2 20 1.4 0.8 0.3 30 1.4 0.8\n"
                       "0.3 0.3 30 1.4 0.7 0.4 40 1.4 0.7\n")
    sims = M.read_input(j, t)
    assert [s.layers[0].thickness for s in sims] == pytest.approx([0.1, 0.2, 0.3])
    assert [s.layers[1].mus for s in sims] == pytest.approx([20, 30, 40])


def test_grid_keys(tmp_path):
    j, t = _write_case(tmp_path, _cfg(grid_nr=200, grid_nz=400, grid_dr=0.02, grid_dz=0.005),
                       "0.1 0.1 10 1.4 0.9 0.1 10 1.4 0.9")
    s = M.read_input(j, t)[0]
    assert (s.grid_nr, s.grid_nz) == (200, 400) and s.grid_dz == pytest.approx(0.005)


def test_malformed_json(tmp_path):
    j = tmp_path / "bad.json"
    j.write_text('{"number_photons": 10, ')
    t = tmp_path / "in.txt"
    t.write_text("1")
    with pytest.raises(M.MCMLError):
        M.read_input(str(j), str(t))


# ------------------------------------------------------------------------- writers
class FakeResult:
    """An mcml_result filled from numpy arrays (kept alive by this object)."""

    def __init__(self, S, L, refl, det, recs_per_sds, nz=4, nr=3, seed=0):
        rng = np.random.default_rng(seed)
        self.keep = []
        r = abi.Result()
        r.nsets, r.num_sds, r.num_layers, r.grid_nr, r.grid_nz = 1, S, L, nr, nz

        def arr(a, ctype):
            a = np.ascontiguousarray(a)
            self.keep.append(a)
            return a.ctypes.data_as(C.POINTER(ctype))

        r.reflectance = arr(np.array(refl, np.float64), C.c_double)
        r.detected = arr(np.array(det, np.uint64), C.c_uint64)
        r.replayed = 1
        r.replay_reflectance = arr(np.array(refl, np.float64), C.c_double)
        off = np.concatenate([[0], np.cumsum(recs_per_sds)]).astype(np.uint64)
        r.pl_offset = arr(off, C.c_uint64)
        n = int(off[-1])
        pl = np.zeros((n, L + 2), np.float32)
        pl[:, 0] = rng.uniform(1e3, 4e9, n).astype(np.float32)
        pl[:, 1:L + 1] = rng.uniform(0, 3, (n, L)).astype(np.float32)
        pl[:, L + 1] = rng.integers(10, 9000, n).astype(np.float32)
        self.pl = pl
        r.pathlength = arr(pl, C.c_float)
        self.arz = rng.integers(0, 2 ** 40, (S, nz, nr)).astype(np.uint64)
        self.a0 = rng.integers(0, 2 ** 40, (S, nz)).astype(np.uint64)
        r.A_rz = arr(self.arz, C.c_uint64)
        r.A0_z = arr(self.a0, C.c_uint64)
        self.r = r
        self.off = off


def g(x):
    return "%g" % x


@pytest.mark.parametrize("binary", [0, 1])
def test_writers_byte_exact(tmp_path, binary):
    L, S, N = 3, 2, 1000
    cfg = F.base_cfg(N, L, 1.457, 1.457, 1, 1, F.probes((0.37, 0.075, 0), [(0.8, .12, .02, 0), (1.5, .12, .02, 0)]))
    sim = abi.make_sim(cfg, [(0.1, 0.3, 30, 1.4, 0.9), (0.2, 0.2, 20, 1.4, 0.9), (0, 0.25, 25, 1.4, 0.9)])
    sim.grid_nr, sim.grid_nz, sim.grid_dr, sim.grid_dz = 3, 4, 0.02, 0.005
    refl = [1.2345678e9, 5.5e6]
    fr = FakeResult(S, L, refl, [7, 3], [7, 3])
    opts = M.make_opts(replay=True, absorbance=True, pathlength=True, avg_pathlength=True, binary=binary)
    info = abi.SummaryInfo(1.5, 0.25, b"AMD Instinct MI355X")
    out = tmp_path / "output.txt"
    lib = M.load_library()
    for _ in range(2):  # text outputs append, binary ones are rewritten
        rc = lib.mcml_write_outputs(C.byref(sim), C.byref(opts), C.byref(fr.r), 0, str(out).encode(),
                                    str(tmp_path).encode(), C.byref(info))
        assert rc == 0
    scale1 = float(0xFFFFFFFF) * N
    line = "".join(g(x / scale1) + " " for x in refl) + "\n"
    assert out.read_text() == line * 2
    # pathlength files
    for s in range(S):
        rows = fr.pl[fr.off[s]:fr.off[s + 1]]
        if binary:
            assert (tmp_path / ("pathlength_SDS_%d.bin" % (s + 1))).read_bytes() == rows.astype("<f4").tobytes()
        else:
            txt = "".join("".join(g(float(v)) + "\t" for v in row) + "\n" for row in rows)
            assert (tmp_path / ("pathlength_SDS_%d.txt" % (s + 1))).read_text() == txt * 2
    # average pathlength (transport.cu:274-297)
    isw = float(np.float32(1) / np.float32(M.derive(sim).start_weight))
    vals = []
    for s in range(S):
        rows = fr.pl[fr.off[s]:fr.off[s + 1]].astype(np.float64)
        for l in range(1, L + 1):
            dmua = 0.001 * float(np.float32([0.3, 0.2, 0.25][l - 1]))
            e = rows[:, 0] * isw
            sr = 0.0
            sc = 0.0
            for k in range(len(rows)):
                sr += e[k]
                sc += e[k] * np.exp(-1 * dmua * rows[k, l])
            vals.append(np.log(sr / sc) / dmua)
    assert (tmp_path / "average_pathlength.txt").read_text() == ("".join(g(v) + "\t" for v in vals) + "\n") * 2
    # absorbance grids (io.cpp:383-469)
    PI = float(np.float32(3.141592654))
    dr, dz = float(np.float32(0.02)), float(np.float32(0.005))
    for s in range(S):
        exp = np.zeros((4, 3))
        for z in range(4):
            for r in range(3):
                exp[z, r] = float(fr.arz[s, z, r]) / (scale1 * 2 * PI * (r + 0.5) * dr * dr * dz)
        e0 = np.array([float(v) / (scale1 * 2 * PI * (0 + 0.5) * dr * dr * dz) for v in fr.a0[s]])
        if binary:
            assert (tmp_path / ("A_rz_SDS_%d.bin" % (s + 1))).read_bytes() == exp.astype("<f4").tobytes()
            assert (tmp_path / ("A0_z_SDS_%d.bin" % (s + 1))).read_bytes() == e0.astype("<f4").tobytes()
        else:
            txt = "".join("".join(g(v) + "\t" for v in row) + "\n" for row in exp)
            assert (tmp_path / ("A_rz_SDS_%d.txt" % (s + 1))).read_text() == txt * 2
            assert (tmp_path / ("A0_z_SDS_%d.txt" % (s + 1))).read_text() == "".join(g(v) + "\n" for v in e0) * 2
    # summary.json (io.cpp:345-381), appended per call
    one = ('{\n"MCML_version_K": 3.02,\n"num_photon": 1000,\n"sim_time": 1.5,\n"replay_time": 0.25,\n'
           '"sim_speed(photons/s)": %s,\n"sim_GPU": "AMD Instinct MI355X",\n"each_photon_weight": %s,\n'
           '"number_layers": 3,\n"detect_mode": "fiber",\n"num_SDS": 2,\n"SDS_detected_number": [ 7 , 3 ]\n}'
           % (g(1000 / 1.75), g(scale1)))
    assert (tmp_path / "summary.json").read_text() == one * 2


def test_output_without_replay_uses_first_pass_reflectance(tmp_path):
    cfg = F.base_cfg(10, 1, 1.0, 1.0, 0, 0, F.probes((0, 0, 0), [(0.5, .2, .1, 0)]))
    sim = abi.make_sim(cfg, [(0, 1, 10, 1.4, 0)])
    fr = FakeResult(1, 1, [3e9], [2], [2])
    fr.r.replayed = 0
    opts = M.make_opts()
    info = abi.SummaryInfo(2.0, 0.0, b"gpu")
    out = tmp_path / "o.txt"
    rc = M.load_library().mcml_write_outputs(C.byref(sim), C.byref(opts), C.byref(fr.r), 0, str(out).encode(),
                                             str(tmp_path).encode(), C.byref(info))
    assert rc == 0
    assert out.read_text() == g(3e9 / (float(0xFFFFFFFF) * 10)) + " \n"
    s = (tmp_path / "summary.json").read_text()
    assert '"detect_mode": "ring"' in s and "replay_time" not in s and '"sim_speed(photons/s)": 5,' in s
    assert not (tmp_path / "pathlength_SDS_1.txt").exists()
