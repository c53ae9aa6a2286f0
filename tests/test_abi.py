"""The C ABI: exported symbols and struct layouts (CPU; no compute calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import mcml_gpu_amd as M
from mcml_gpu_amd import abi

HEADER = os.path.join(M.ROOT, "include", "mcml.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mcml_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = M.load_library()
    names = declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


STRUCTS = {"mcml_layer": abi.Layer, "mcml_probe": abi.Probe, "mcml_sim": abi.Sim,
           "mcml_layer_derived": abi.LayerDerived, "mcml_derived": abi.Derived, "mcml_opts": abi.Opts,
           "mcml_record": abi.Record, "mcml_result": abi.Result, "mcml_summary_info": abi.SummaryInfo}


def test_ctypes_layouts_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mcml.h"', "int main(void){"]
    for cname, py in STRUCTS.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n") if l)
    for cname, py in STRUCTS.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[cname + "." + f]) == getattr(py, f).offset, (cname, f)


def test_create_without_gpu_fails_cleanly():
    """Without a usable device the library reports MCML_ERR_NODEV instead of falling back."""
    n = C.c_int(-1)
    rc = M.load_library().mcml_device_count(C.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(M.MCMLError) as e:
        M.Context(0)
    assert e.value.code == abi.ERR_NODEV


def test_invalid_arguments_rejected_before_device_use():
    lib = M.load_library()
    d = abi.Derived()
    sim = abi.Sim()
    sim.number_layers = 0
    assert lib.mcml_derive(C.byref(sim), C.byref(d)) == abi.ERR_LIMIT
    sim.number_layers, sim.num_sds = 11, 1
    assert lib.mcml_derive(C.byref(sim), C.byref(d)) == abi.ERR_LIMIT
    sim.number_layers, sim.num_sds = 2, 31
    assert lib.mcml_derive(C.byref(sim), C.byref(d)) == abi.ERR_LIMIT
    # a negative or NaN thickness (the reference reads it unchecked) would put z_min above z_max
    sim.number_layers, sim.num_sds = 3, 1
    sim.layers[0].thickness, sim.layers[1].thickness = 0.1, -0.05
    assert lib.mcml_derive(C.byref(sim), C.byref(d)) == abi.ERR_INPUT
    sim.layers[1].thickness = float("nan")
    assert lib.mcml_derive(C.byref(sim), C.byref(d)) == abi.ERR_INPUT
    sim.layers[1].thickness = 0.0  # a layer of zero thickness is accepted
    assert lib.mcml_derive(C.byref(sim), C.byref(d)) == 0
