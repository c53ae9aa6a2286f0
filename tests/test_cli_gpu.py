"""The MCML_GPU executable end to end on the GPU: two input lines in one batched launch, -R -P
-AP -A -B with a fixed seed.  Every file must hold what the library returns for the same
batch and seed (GPU)."""
import json
import os
import subprocess

import numpy as np
import pytest

import mcml_gpu_amd as M
from mcml_gpu_amd import loaders

SEED = 4242
LAYERS = [  # h, mua, mus, n, g per layer; two simulations
    [(0.1, 0.3, 60.0, 1.4, 0.8), (0.3, 0.05, 120.0, 1.37, 0.9), (0.0, 0.2, 80.0, 1.45, 0.85)],
    [(0.2, 0.6, 90.0, 1.38, 0.85), (0.2, 0.1, 100.0, 1.4, 0.9), (0.0, 0.3, 70.0, 1.42, 0.8)],
]


def write_inputs(tmp_path, photons):
    cfg = {"number_simulation": 2, "number_photons": photons, "number_layers": 3, "detector_reflectance": 0.35,
           "upper_n": 1.457, "lower_n": 1.457, "source_probe_oblique": 1, "detector_probe_oblique": 1,
           "grid_nr": 40, "grid_nz": 30, "grid_dr": 0.02, "grid_dz": 0.01,
           "probes": {"source": {"NA": 0.37, "radius": 0.075, "angle": 8.0}, "num_SDS": 3,
                      "detectors": [{"pos": 0.25, "NA": 0.5, "radius": 0.15, "angle": 10.0},
                                    {"pos": 0.5, "NA": 0.6, "radius": 0.2, "angle": 0.0},
                                    {"pos": 0.9, "NA": 0.6, "radius": 0.3, "angle": 25.0}]}}
    js = tmp_path / "sim_set.json"
    js.write_text(json.dumps(cfg))
    lines = []
    for layers in LAYERS:
        vals = []
        for i, (h, mua, mus, n, g) in enumerate(layers):
            if i < len(layers) - 1:
                vals.append(h)
            vals += [mua, mus, n, g]
        lines.append(" ".join(repr(v) for v in vals))
    txt = tmp_path / "input.txt"
    txt.write_text("\n".join(lines) + "\n")
    return str(js), str(txt)


@pytest.mark.gpu
def test_cli_outputs_equal_library_results(tmp_path):
    photons = 200000
    js, txt = write_inputs(tmp_path, photons)
    r = subprocess.run([M.EXE_PATH, js, txt, "output.txt", "-R", "-P", "-AP", "-A", "-B", "-S", str(SEED)],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    sims = M.read_input(js, txt)
    assert len(sims) == 2
    with M.Context(0) as ctx:
        res = ctx.run_batch(sims, M.make_opts(replay=True, pathlength=True, avg_pathlength=True, absorbance=True,
                                              binary=True), seed=SEED)
    S, L = 3, 3
    # output.txt: one line per simulation, replayed reflectance per SDS / (0xFFFFFFFF N)
    rows = [[float(v) for v in line.split()] for line in (tmp_path / "output.txt").read_text().splitlines()]
    assert len(rows) == 2
    for j in range(2):
        want = res.replay_reflectance[j] / (4294967295.0 * photons)
        assert np.allclose(rows[j], want, rtol=1e-5, atol=0)
    # pathlength_SDS_#.bin: binary files are overwritten, so they hold the last simulation
    for s in range(1, S + 1):
        got = np.fromfile(tmp_path / ("pathlength_SDS_%d.bin" % s), dtype="<f4").reshape(-1, L + 2)
        assert np.array_equal(got, res.pathlength_records(1, s))
        assert len(got) == res.detected[1, s - 1] > 0
    # average_pathlength.txt: appended, one line per simulation
    avg = [[float(v) for v in line.split()] for line in (tmp_path / "average_pathlength.txt").read_text().splitlines()]
    assert len(avg) == 2 and all(len(a) == S * L for a in avg)
    assert all(np.all(np.isfinite(a)) and min(a) >= 0 for a in avg)
    # absorbance maps: last simulation, FP32 binary, z-major
    for s in range(1, S + 1):
        a = np.fromfile(tmp_path / ("A_rz_SDS_%d.bin" % s), dtype="<f4")
        assert a.size == 30 * 40 and np.all(a >= 0)
        nz_cells = res.A_rz[1, s - 1].reshape(-1) > 0
        assert np.array_equal(a > 0, nz_cells)
    # summary.json: appended, one JSON object per simulation
    objs = loaders.read_summaries(tmp_path / "summary.json")
    assert len(objs) == 2
    assert [o["SDS_detected_number"] for o in objs] == [list(map(int, res.detected[j])) for j in range(2)]
    # the NumPy loaders read the same files
    for s in range(1, S + 1):
        got = loaders.load_binary_pathlength(tmp_path / "summary.json", s, tmp_path / ("pathlength_SDS_%d.bin" % s))
        assert np.array_equal(got, res.pathlength_records(1, s))
        amap = loaders.load_A_arr(tmp_path, s, nz=30, nr=40)
        assert amap.shape == (30, 40) and np.all(amap >= 0)


@pytest.mark.gpu
def test_cli_outputs_do_not_depend_on_batching(tmp_path):
    """Lines run in batches that fit a memory budget; with MCML_BATCH_BUDGET_BYTES=1 every
    line is a batch of its own, under a shifted seed that gives it the photon streams it has
    in one batch: every output file (summary.json, which holds timings, aside) is the same."""
    photons = 100000
    outs = []
    for budget in (None, "1"):
        d = tmp_path / ("whole" if budget is None else "split")
        d.mkdir()
        js, txt = write_inputs(d, photons)
        env = dict(os.environ)
        if budget is not None:
            env["MCML_BATCH_BUDGET_BYTES"] = budget
        r = subprocess.run([M.EXE_PATH, js, txt, "output.txt", "-R", "-P", "-AP", "-A", "-S", str(SEED)],
                           capture_output=True, text=True, cwd=d, timeout=120, env=env)
        print("budget", budget, r.stdout[-600:], r.stderr[-600:], flush=True)
        assert r.returncode == 0
        outs.append(d)
    names = sorted(f for f in os.listdir(outs[0]) if f not in ("sim_set.json", "input.txt", "summary.json"))
    assert "output.txt" in names and any(f.startswith("A_rz_SDS_") for f in names)
    assert names == sorted(f for f in os.listdir(outs[1]) if f not in ("sim_set.json", "input.txt", "summary.json"))
    for f in names:
        assert (outs[0] / f).read_bytes() == (outs[1] / f).read_bytes(), f


@pytest.mark.gpu
@pytest.mark.parametrize("binary,photons", [(True, 3000000), (False, 400000)])
def test_cli_pathlength_files_written_while_reading_back(tmp_path, binary, photons):
    """-P files are written by a thread of their own as each (set, SDS) group's rows land on
    the host, 16 MB at a time.  Three accept-all detectors on a semi-infinite medium give
    about 3 records per two photons (binary case: over 80 MB of rows, several read-back
    chunks): every file must hold its group's rows, whole and in order, as the library
    returns them for the same seed."""
    cfg = {"number_simulation": 1, "number_photons": photons, "number_layers": 3, "detector_reflectance": 0.0,
           "upper_n": 1.457, "lower_n": 1.457, "source_probe_oblique": 1, "detector_probe_oblique": 1,
           "probes": {"source": {"NA": 0.0, "radius": 0.0, "angle": 0.0}, "num_SDS": 3,
                      "detectors": [{"pos": 0.0, "NA": 1.457, "radius": 1000.0, "angle": 0.0}] * 3}}
    (tmp_path / "sim_set.json").write_text(json.dumps(cfg))
    (tmp_path / "input.txt").write_text("0.1 0.2 60.0 1.4 0.8 0.3 0.05 120.0 1.37 0.9 0.2 80.0 1.45 0.85\n")
    args = [M.EXE_PATH, "sim_set.json", "input.txt", "output.txt", "-R", "-P", "-S", str(SEED)]
    if binary:
        args.append("-B")
    print("cli: accept-all, %d photons, binary=%s" % (photons, binary), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    print(r.stdout[-1000:], r.stderr[-1000:], flush=True)
    assert r.returncode == 0
    sims = M.read_input(str(tmp_path / "sim_set.json"), str(tmp_path / "input.txt"))
    with M.Context(0) as ctx:
        res = ctx.run_batch(sims, M.make_opts(replay=True, pathlength=True, binary=binary), seed=SEED)
    L = 3
    if binary:
        assert len(res.pathlength) * (L + 2) * 4 > 3 * (16 << 20)
    for s in range(1, 4):
        want = res.pathlength_records(0, s)
        if binary:
            got = np.fromfile(tmp_path / ("pathlength_SDS_%d.bin" % s), dtype="<f4").reshape(-1, L + 2)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), s
        else:
            text = (tmp_path / ("pathlength_SDS_%d.txt" % s)).read_text()
            lines = text.split("\n")
            assert lines[-1] == "" and len(lines) - 1 == len(want)
            got = np.array(text.split(), dtype=np.float64).reshape(-1, L + 2)
            assert np.allclose(got, want, rtol=6e-6, atol=0), s
