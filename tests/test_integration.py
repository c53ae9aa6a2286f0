"""The C integration example of INTEGRATION.md section 2 (the replacement of DoOneSimulation,
transport.cu:57-272), compiled exactly as written.

CPU: the block compiles as C11 against include/mcml.h, with the reference's SimOptions
(header.h:195-203), and links against mcml_gpu_amd/libmcml.so.
GPU: the program built around it writes, for one simulation, the same files as the drop-in
executable given the same inputs and seed (summary.json aside, which holds timings).
"""
import filecmp
import json
import os
import re
import subprocess

import pytest

import mcml_gpu_amd as M

DOC = os.path.join(M.ROOT, "INTEGRATION.md")

# The reference's SimOptions (header.h:195-203), as a maintainer's code would declare it.
PRELUDE = """#include <stdbool.h>
#include <stdlib.h>
typedef struct {
    bool do_replay;
    bool do_output_A_arr;
    bool output_each_pathlength;
    bool do_output_average_pathlength;
    bool do_output_bin;
    int GPU_select_to_run;
} SimOptions;
"""

# A main() in the shape of main.cu:17-145: read the inputs, open the device once, run every
# simulation through the documented function.  argv: sim_set.json input.txt output.txt seed
MAIN = """
int main(int argc, char** argv)
{
    if (argc < 5) return 2;
    mcml_sim* sims = NULL;
    int n = 0;
    char msg[512];
    if (mcml_read_input(argv[1], argv[2], &sims, &n, msg, sizeof(msg)) != MCML_OK) {
        fprintf(stderr, "%s\\n", msg);
        return 1;
    }
    SimOptions simOpt = {true, false, true, true, true, 0};
    int err;
    mcml_ctx* ctx = mcml_create(simOpt.GPU_select_to_run, &err);
    if (!ctx) return 3;
    int rc = 0;
    for (int i = 0; i < n && rc == 0; i++)
        rc = DoOneSimulation_mcml(ctx, &sims[i], simOpt, argv[3], strtoull(argv[4], NULL, 10));
    mcml_destroy(ctx);
    mcml_free_sims(sims);
    return rc == 0 ? 0 : 4;
}
"""


def documented_block():
    """The first C block of section 2 of INTEGRATION.md."""
    text = open(DOC).read()
    sec = text[text.index("## 2. Replacing `DoOneSimulation`"):text.index("## 3.")]
    blocks = re.findall(r"```c\n(.*?)```", sec, flags=re.S)
    assert blocks, "no C block in section 2"
    return blocks[0]


def build_example(tmp_path):
    src = tmp_path / "do_one_simulation.c"
    src.write_text(PRELUDE + documented_block() + MAIN)
    exe = tmp_path / "do_one_simulation"
    libdir = os.path.dirname(M.LIB_PATH)
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(M.ROOT, "include"), str(src),
                        "-o", str(exe), "-L", libdir, "-lmcml", "-Wl,-rpath," + libdir],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def test_documented_example_compiles_and_links(tmp_path):
    M.load_library()  # the library must exist to link against
    exe = build_example(tmp_path)
    nm = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True).stdout
    for sym in ("mcml_run", "mcml_write_outputs", "mcml_result_free", "mcml_create"):
        assert sym in nm


def write_one_simulation(d, photons):
    cfg = {"number_simulation": 1, "number_photons": photons, "number_layers": 3, "detector_reflectance": 0.35,
           "upper_n": 1.457, "lower_n": 1.457, "source_probe_oblique": 1, "detector_probe_oblique": 1,
           "probes": {"source": {"NA": 0.37, "radius": 0.075, "angle": 8.0}, "num_SDS": 3,
                      "detectors": [{"pos": 0.25, "NA": 0.5, "radius": 0.15, "angle": 10.0},
                                    {"pos": 0.5, "NA": 0.6, "radius": 0.2, "angle": 0.0},
                                    {"pos": 0.9, "NA": 0.6, "radius": 0.3, "angle": 25.0}]}}
    (d / "sim_set.json").write_text(json.dumps(cfg))
    (d / "input.txt").write_text("0.1 0.3 60 1.4 0.8 0.3 0.05 120 1.37 0.9 0.2 80 1.45 0.85\n")


@pytest.mark.gpu
def test_documented_example_writes_the_executables_files(tmp_path):
    exe = build_example(tmp_path)
    a, b = tmp_path / "example", tmp_path / "cli"
    for d in (a, b):
        d.mkdir()
        write_one_simulation(d, 100000)
    seed = "777"
    print("integration example: 1e5 photons, -R -P -AP -B, seed", seed, flush=True)
    r = subprocess.run([exe, "sim_set.json", "input.txt", "output.txt", seed], cwd=a, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([M.EXE_PATH, "sim_set.json", "input.txt", "output.txt", "-R", "-P", "-AP", "-B", "-S", seed],
                       cwd=b, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    files = sorted(f for f in os.listdir(b) if f not in ("sim_set.json", "input.txt", "summary.json"))
    assert "output.txt" in files and "average_pathlength.txt" in files
    assert any(f.startswith("pathlength_SDS_") and f.endswith(".bin") for f in files)
    assert sorted(f for f in os.listdir(a) if f not in ("sim_set.json", "input.txt", "summary.json")) == files
    for f in files:
        assert filecmp.cmp(a / f, b / f, shallow=False), f
    ja = (a / "summary.json").read_text()
    jb = (b / "summary.json").read_text()
    for key in ("number_layers", "SDS_detected_number"):
        assert key in ja and key in jb
