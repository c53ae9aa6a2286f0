"""The MCML_GPU command line (main.cu:17-180): option rules and early exits (CPU).

All cases here stop before any GPU work, exactly where the reference stops."""
import os
import subprocess

import pytest

import mcml_gpu_amd as M

EXE = M.EXE_PATH
CFG = os.path.join(M.CONFIG_DIR, "cfg0_sim_set.json")
TXT = os.path.join(M.CONFIG_DIR, "cfg0_input.txt")


def run(*args, cwd=None):
    if not os.path.exists(EXE):
        pytest.fail("MCML_GPU is not built: run python -m mcml_gpu_amd.build")
    return subprocess.run([EXE, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


def test_usage_when_arguments_missing():
    r = run()
    assert r.returncode == 0 and "Usage: MCML_GPU sim_set.json input.txt output.txt" in r.stderr
    r = run(CFG, TXT)
    assert r.returncode == 0 and "Usage:" in r.stderr


def test_help_flag_anywhere():
    r = run(CFG, TXT, "out.txt", "-h")
    assert r.returncode == 0 and "-R,\t\t Replay the detected photon" in r.stderr


@pytest.mark.parametrize("flags,msg", [
    (["-P"], "-P option only work with -R option!"),
    (["-A"], "-A option only work with -R option!"),
    (["-AP"], "-AP option only work with -R option!"),
    (["-R", "-B"], None),  # -R alone implies -AP, and -B then needs -P or -A
    (["-B"], "-B option only work with -P or -A option!"),
])
def test_option_rules(tmp_path, flags, msg):
    r = run(CFG, TXT, str(tmp_path / "o.txt"), *flags, cwd=tmp_path)
    assert r.returncode == 0
    if msg is None:
        msg = "-B option only work with -P or -A option!"
    assert msg in r.stdout
    assert not (tmp_path / "o.txt").exists()


def test_unreadable_input_exits_1(tmp_path):
    r = run(CFG, str(tmp_path / "missing.txt"), str(tmp_path / "o.txt"), cwd=tmp_path)
    assert r.returncode == 1 and "Something wrong with read_simulation_data!" in r.stdout


def test_rejected_input_prints_reference_message(tmp_path):
    bad = tmp_path / "in.txt"
    bad.write_text("0.27 0.3201 35.2861 1.4 0 0.65 0.2954 30.5833 1.4 0 0.2747 130.6813 0.9 0")
    r = run(CFG, str(bad), str(tmp_path / "o.txt"), cwd=tmp_path)
    assert r.returncode == 1
    assert "n for simulation 0 layer 2 Error with value 0.9 !" in r.stdout
    assert "Something wrong with read_simulation_data!" in r.stdout


def test_gpu_selection_error(tmp_path):
    r = run(CFG, TXT, str(tmp_path / "o.txt"), "-G", "-1", cwd=tmp_path)
    assert r.returncode == 0 and "GPU selection ERROR!" in r.stdout
