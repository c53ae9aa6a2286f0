"""NumPy readers for the files MCML_GPU writes.

They replace the reference's MATLAB loaders (utils/fun_load_A_arr.m,
utils/load_binary_pathlength_output.m) and add readers for the text formats, so tests and
users can load results without MATLAB.  Formats are those of the writers in io.cpp.
"""
import json
import os

import numpy as np

DEFAULT_NZ = 500   # header.h:65-68 compiled grid, the reference loader's n_dz
DEFAULT_NR = 1000  # and n_dr


def read_summaries(path):
    """summary.json holds one JSON object per simulation, appended one after another
    (io.cpp:345-381); return them as a list of dicts."""
    text = open(path).read()
    dec, objs, i = json.JSONDecoder(), [], 0
    while i < len(text):
        if text[i].isspace():
            i += 1
            continue
        obj, i = dec.raw_decode(text, i)
        objs.append(obj)
    return objs


def load_binary_pathlength(summary_file, sds_index, pathlength_file, simulation=-1):
    """Records of pathlength_SDS_#.bin as an (n, L + 2) float32 array of
    [weight, PL_1..PL_L, steps] (utils/load_binary_pathlength_output.m:11-19).  The record
    count comes from the summary's SDS_detected_number (1-based sds_index).  Binary files are
    rewritten for every simulation, so they hold the last one: `simulation` picks the summary
    entry that describes the file (default: the last)."""
    s = read_summaries(summary_file)[simulation]
    if not 1 <= sds_index <= s["num_SDS"]:
        raise ValueError("SDS_index out of range!")
    L = int(s["number_layers"])
    n = int(s["SDS_detected_number"][sds_index - 1])
    data = np.fromfile(pathlength_file, dtype="<f4", count=n * (L + 2))
    if data.size != n * (L + 2):
        raise ValueError("%s holds %d values, the summary expects %d" % (pathlength_file, data.size, n * (L + 2)))
    return data.reshape(n, L + 2)


def load_text_pathlength(path):
    """pathlength_SDS_#.txt: one record per line, tab separated with a trailing tab, appended
    per simulation.  Returns an (n, L + 2) float64 array."""
    rows = [[float(v) for v in line.split("\t") if v.strip()] for line in open(path) if line.strip()]
    return np.array(rows, dtype=np.float64)


def load_A_arr(input_dir, sds, nz=DEFAULT_NZ, nr=DEFAULT_NR, binary=True):
    """Absorbance map of one SDS as an (nz, nr) array: A_rz with A0_z added to the r = 0
    column (utils/fun_load_A_arr.m:9-28).  Files are z-major with r fastest; binary files are
    float32, text files hold one z row per line (A0_z one value per line)."""
    if binary:
        a_rz = np.fromfile(os.path.join(input_dir, "A_rz_SDS_%d.bin" % sds), dtype="<f4").astype(np.float64)
        a0_z = np.fromfile(os.path.join(input_dir, "A0_z_SDS_%d.bin" % sds), dtype="<f4").astype(np.float64)
    else:
        a_rz = np.array([float(v) for line in open(os.path.join(input_dir, "A_rz_SDS_%d.txt" % sds))
                         for v in line.split("\t") if v.strip()])
        a0_z = np.array([float(line) for line in open(os.path.join(input_dir, "A0_z_SDS_%d.txt" % sds))
                         if line.strip()])
    if a_rz.size != nz * nr or a0_z.size != nz:
        raise ValueError("grid of %d x %d does not match %d A_rz and %d A0_z values" % (nz, nr, a_rz.size, a0_z.size))
    out = a_rz.reshape(nz, nr).copy()
    out[:, 0] += a0_z
    return out


def load_output(path):
    """output.txt: one line per simulation of per-SDS reflectance, space separated."""
    return [np.array([float(v) for v in line.split()]) for line in open(path) if line.strip()]


def load_average_pathlength(path, num_sds, num_layers):
    """average_pathlength.txt: one line per simulation of S * L values (SDS-major)."""
    rows = [[float(v) for v in line.split("\t") if v.strip()] for line in open(path) if line.strip()]
    return np.array(rows, dtype=np.float64).reshape(-1, num_sds, num_layers)
