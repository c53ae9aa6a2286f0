"""ctypes mirror of include/mcml.h (structs and constants only).

The layouts here must match the C header field for field; tests check sizeof() of each
struct against the library's own view.
"""
import ctypes as C

MAX_LAYERS = 10
MAX_SDS = 30

OK = 0
ERR_ARG = -1
ERR_INPUT = -2
ERR_IO = -3
ERR_DEVICE = -4
ERR_NOMEM = -5
ERR_NODEV = -6
ERR_LIMIT = -7

ERROR_NAMES = {
    OK: "ok", ERR_ARG: "invalid argument", ERR_INPUT: "input rejected", ERR_IO: "i/o error",
    ERR_DEVICE: "device error", ERR_NOMEM: "out of memory", ERR_NODEV: "no device",
    ERR_LIMIT: "size limit exceeded",
}


class Layer(C.Structure):
    _fields_ = [("thickness", C.c_float), ("mua", C.c_float), ("mus", C.c_float),
                ("n", C.c_float), ("g", C.c_float)]


class Probe(C.Structure):
    _fields_ = [("pos", C.c_float), ("NA", C.c_float), ("radius", C.c_float),
                ("angle_deg", C.c_float)]


class Sim(C.Structure):
    _fields_ = [
        ("number_photons", C.c_uint64),
        ("number_layers", C.c_int32),
        ("num_sds", C.c_int32),
        ("detector_reflectance", C.c_float),
        ("upper_n", C.c_float),
        ("lower_n", C.c_float),
        ("source_probe_oblique", C.c_int32),
        ("detector_probe_oblique", C.c_int32),
        ("source", Probe),
        ("detectors", Probe * MAX_SDS),
        ("layers", Layer * MAX_LAYERS),
        ("grid_nr", C.c_int32),
        ("grid_nz", C.c_int32),
        ("grid_dr", C.c_float),
        ("grid_dz", C.c_float),
    ]


class LayerDerived(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("z_min", "z_max", "mutr", "mua", "g", "n")]


_D = MAX_SDS + 1


class Derived(C.Structure):
    _fields_ = [
        ("number_photons", C.c_uint64),
        ("num_layers", C.c_int32),
        ("num_sds", C.c_int32),
        ("start_weight", C.c_float),
        ("detector_reflectance", C.c_float),
        ("source_oblique", C.c_int32),
        ("detector_oblique", C.c_int32),
        ("layer", LayerDerived * (MAX_LAYERS + 2)),
        ("src_radius", C.c_float),
        ("src_na", C.c_float),
        ("src_theta_max", C.c_float),
        ("src_sin_a", C.c_float),
        ("src_cos_a", C.c_float),
        ("det_pos", C.c_float * _D),
        ("det_radius", C.c_float * _D),
        ("det_r2", C.c_float * _D),
        ("det_cos_a", C.c_float * _D),
        ("det_sin_a", C.c_float * _D),
        ("det_crit", C.c_float * _D),
        ("det_cos_crit", C.c_float * _D),
        ("grid_nr", C.c_int32),
        ("grid_nz", C.c_int32),
        ("grid_dr", C.c_float),
        ("grid_dz", C.c_float),
    ]


class Opts(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "replay", "absorbance", "pathlength", "avg_pathlength", "binary",
        "block_size", "blocks_per_cu", "chunk", "refill_min", "boundary_min", "reference_shape",
        "use_graph", "skip_records")] + [("record_cap", C.c_uint64), ("lane_stats", C.c_int32),
                                         ("spill_below", C.c_int32), ("separate_replay", C.c_int32)]


class Record(C.Structure):
    _fields_ = [("key", C.c_uint64), ("sds", C.c_uint32), ("weight", C.c_float)]


class Result(C.Structure):
    _fields_ = [
        ("nsets", C.c_int32),
        ("num_sds", C.c_int32),
        ("num_layers", C.c_int32),
        ("grid_nr", C.c_int32),
        ("grid_nz", C.c_int32),
        ("reflectance", C.POINTER(C.c_double)),
        ("detected", C.POINTER(C.c_uint64)),
        ("n_records", C.c_uint64),
        ("records", C.POINTER(Record)),
        ("replayed", C.c_int32),
        ("replay_reflectance", C.POINTER(C.c_double)),
        ("pl_offset", C.POINTER(C.c_uint64)),
        ("pathlength", C.POINTER(C.c_float)),
        ("A_rz", C.POINTER(C.c_uint64)),
        ("A0_z", C.POINTER(C.c_uint64)),
        ("photons", C.c_uint64),
        ("steps", C.c_uint64),
        ("steps_boundary", C.c_uint64),
        ("steps_interaction", C.c_uint64),
        ("ms_first_pass_kernel", C.c_double),
        ("ms_replay_kernel", C.c_double),
        ("ms_dump_kernel", C.c_double),
        ("ms_first_pass_wall", C.c_double),
        ("ms_replay_wall", C.c_double),
        ("first_pass_launches", C.c_int32),
        ("block_size", C.c_int32),
        ("grid_size", C.c_int32),
        ("chunk", C.c_int32),
        ("refill_min", C.c_int32),
        ("boundary_min", C.c_int32),
        ("lane_blocks", C.c_uint64 * 4),
        ("ms_drain", C.c_double),
        ("ms_clock", C.c_double),
        ("drain_launches", C.c_int32),
    ]


class SummaryInfo(C.Structure):
    _fields_ = [("sim_time_s", C.c_double), ("replay_time_s", C.c_double),
                ("gpu_name", C.c_char_p)]


RECORD_DTYPE = [("key", "<u8"), ("sds", "<u4"), ("weight", "<f4")]


def make_sim(cfg: dict, layers) -> Sim:
    """Build an mcml_sim from a sim_set.json-shaped dict and a list of layer tuples.

    `layers` holds (h, mua, mus, n, g) per tissue layer; the last layer's h is ignored,
    exactly as input.txt's last layer has no thickness (io.cpp:141-146).
    """
    s = Sim()
    s.number_photons = int(cfg["number_photons"])
    s.number_layers = int(cfg["number_layers"])
    s.detector_reflectance = float(cfg.get("detector_reflectance", 0.0))
    s.upper_n = float(cfg["upper_n"])
    s.lower_n = float(cfg["lower_n"])
    s.source_probe_oblique = int(cfg["source_probe_oblique"])
    s.detector_probe_oblique = int(cfg["detector_probe_oblique"])
    pr = cfg["probes"]
    src = pr["source"]
    s.source = Probe(0.0, float(src["NA"]), float(src["radius"]), float(src["angle"]))
    s.num_sds = int(pr["num_SDS"])
    for i, d in enumerate(pr["detectors"]):
        s.detectors[i] = Probe(float(d["pos"]), float(d["NA"]), float(d["radius"]),
                               float(d["angle"]))
    if len(layers) != s.number_layers:
        raise ValueError("layer count does not match number_layers")
    for i, ly in enumerate(layers):
        h, mua, mus, n, g = ly
        s.layers[i] = Layer(float(h), float(mua), float(mus), float(n), float(g))
    s.grid_nr = int(cfg.get("grid_nr", 0))
    s.grid_nz = int(cfg.get("grid_nz", 0))
    s.grid_dr = float(cfg.get("grid_dr", 0.0))
    s.grid_dz = float(cfg.get("grid_dz", 0.0))
    return s
