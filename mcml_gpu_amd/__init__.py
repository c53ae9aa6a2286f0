"""mcml_gpu_amd - MI355X-native photon-transport hot path of MCML_GPU.

Python access to libmcml.so (built in-tree by `python -m mcml_gpu_amd.build`).  The
library is the product: C ABI in include/mcml.h, HIP kernels in csrc/transport.hip.
This module only marshals arguments; every photon is simulated by the HIP kernels and
there is no CPU fallback: without a built library, or without a GPU, calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libmcml.so")
EXE_PATH = os.path.join(HERE, "bin", "MCML_GPU")
CONFIG_DIR = os.path.join(ROOT, "configs")

_lib = None


class MCMLError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__("%s (%d)%s" % (abi.ERROR_NAMES.get(code, "error"), code,
                                        (": " + msg) if msg else ""))
        self.code = code


def load_library():
    """Load libmcml.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libmcml.so is not built: run `python -m mcml_gpu_amd.build`")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.mcml_create.argtypes = [C.c_int, P(C.c_int)]
    L.mcml_create.restype = C.c_void_p
    L.mcml_destroy.argtypes = [C.c_void_p]
    L.mcml_last_error.argtypes = [C.c_void_p]
    L.mcml_last_error.restype = C.c_char_p
    L.mcml_device_name.argtypes = [C.c_void_p]
    L.mcml_device_name.restype = C.c_char_p
    L.mcml_device_count.argtypes = [P(C.c_int)]
    L.mcml_derive.argtypes = [P(abi.Sim), P(abi.Derived)]
    L.mcml_run.argtypes = [C.c_void_p, P(abi.Sim), P(abi.Opts), C.c_uint64, P(abi.Result)]
    L.mcml_run_batch.argtypes = [C.c_void_p, P(abi.Sim), C.c_int, P(abi.Opts), C.c_uint64,
                                 P(abi.Result)]
    L.mcml_result_free.argtypes = [P(abi.Result)]
    L.mcml_average_pathlength.argtypes = [P(abi.Sim), P(abi.Result), C.c_int, P(C.c_double)]
    L.mcml_read_input.argtypes = [C.c_char_p, C.c_char_p, P(P(abi.Sim)), P(C.c_int), C.c_char_p,
                                  C.c_size_t]
    L.mcml_free_sims.argtypes = [P(abi.Sim)]
    L.mcml_write_outputs.argtypes = [P(abi.Sim), P(abi.Opts), P(abi.Result), C.c_int, C.c_char_p,
                                     C.c_char_p, P(abi.SummaryInfo)]
    L.mcml_main.argtypes = [C.c_int, P(C.c_char_p)]
    L.mcml_autotune.argtypes = [C.c_void_p, P(abi.Sim), C.c_int, C.c_uint64, P(abi.Opts)]
    L.mcml_white_rescale.argtypes = [C.c_void_p, P(abi.Sim), P(abi.Result), P(C.c_float), C.c_int,
                                     P(C.c_double), P(C.c_double)]
    L.mcml_selftest_math.argtypes = [C.c_void_p, P(C.c_uint64)]
    # device pointers for the two outputs
    L.mcml_run_batch_sums_noalloc.argtypes = [C.c_void_p, P(abi.Sim), C.c_int, P(abi.Opts), C.c_uint64,
                                              C.c_void_p, C.c_void_p]
    _lib = L
    return L


def derive(sim):
    d = abi.Derived()
    rc = load_library().mcml_derive(C.byref(sim), C.byref(d))
    if rc:
        raise MCMLError(rc)
    return d


def read_input(sim_set_json, input_txt):
    """read_mua_mus (io.cpp:35-256): list of abi.Sim, or MCMLError with the reference message."""
    L = load_library()
    p = C.POINTER(abi.Sim)()
    n = C.c_int(0)
    msg = C.create_string_buffer(1024)
    rc = L.mcml_read_input(sim_set_json.encode(), input_txt.encode(), C.byref(p), C.byref(n), msg, 1024)
    if rc:
        raise MCMLError(rc, msg.value.decode(errors="replace").strip())
    sims = [abi.Sim.from_buffer_copy(p[i]) for i in range(n.value)]
    L.mcml_free_sims(p)
    return sims


def load_config(name):
    """The first simulation of configs/<name>_sim_set.json + <name>_input.txt."""
    return read_input(os.path.join(CONFIG_DIR, name + "_sim_set.json"),
                      os.path.join(CONFIG_DIR, name + "_input.txt"))[0]


def make_opts(replay=False, absorbance=False, pathlength=False, avg_pathlength=False, binary=False,
              block_size=0, blocks_per_cu=0, chunk=0, refill_min=0, boundary_min=0, use_graph=False,
              skip_records=False, reference_shape=False, record_cap=0, lane_stats=False, spill_below=0,
              separate_replay=False):
    return abi.Opts(int(replay), int(absorbance), int(pathlength), int(avg_pathlength), int(binary),
                    block_size, blocks_per_cu, chunk, refill_min, boundary_min, int(reference_shape),
                    int(use_graph), int(skip_records), int(record_cap), int(lane_stats), int(spill_below),
                    int(separate_replay))


class _ResultOwner:
    """Owns one library-allocated mcml_result and frees it (mcml_result_free) once the last
    numpy view of its buffers is gone."""

    def __init__(self, lib, r):
        self.lib, self.r = lib, r

    def __del__(self):
        try:
            self.lib.mcml_result_free(C.byref(self.r))
        except Exception:
            pass


def _view(owner, ptr, count, dtype):
    """numpy view (no copy) of `count` elements of `dtype` at ptr that keeps `owner` alive."""
    dtype = np.dtype(dtype)
    if count == 0 or not ptr:
        return np.zeros(0, dtype)
    buf = (C.c_uint8 * (count * dtype.itemsize)).from_address(C.cast(ptr, C.c_void_p).value)
    buf._owner = owner
    return np.frombuffer(buf, dtype=dtype, count=count)


class RunResult:
    """An mcml_result as numpy arrays.  The large buffers (records, pathlength rows,
    absorbance grids) are views of the library's memory, released when the last of them is
    dropped; the per-SDS sums are small copies."""

    def __init__(self, r: abi.Result, lib=None):
        owner = _ResultOwner(lib, r) if lib is not None else None
        self.nsets, self.num_sds, self.num_layers = r.nsets, r.num_sds, r.num_layers
        g = r.nsets * r.num_sds
        self.reflectance = np.ctypeslib.as_array(r.reflectance, (g,)).copy().reshape(r.nsets, r.num_sds)
        self.detected = np.ctypeslib.as_array(r.detected, (g,)).copy().reshape(r.nsets, r.num_sds)

        def big(ptr, count, dtype):
            if owner is not None:
                return _view(owner, ptr, count, dtype)
            return _view(None, ptr, count, dtype).copy()

        n = r.n_records if r.records else 0
        self.records = big(r.records, n, abi.RECORD_DTYPE)
        self.replayed = bool(r.replayed)
        if self.replayed:
            self.replay_reflectance = np.ctypeslib.as_array(r.replay_reflectance, (g,)).copy().reshape(
                r.nsets, r.num_sds)
            self.pl_offset = np.ctypeslib.as_array(r.pl_offset, (g + 1,)).copy()
            nrec = int(self.pl_offset[-1])
            w = r.num_layers + 2
            self.pathlength = big(r.pathlength, nrec * w, np.float32).reshape(nrec, w)
            if r.A_rz:
                self.A_rz = big(r.A_rz, g * r.grid_nz * r.grid_nr, np.uint64).reshape(
                    r.nsets, r.num_sds, r.grid_nz, r.grid_nr)
                self.A0_z = big(r.A0_z, g * r.grid_nz, np.uint64).reshape(r.nsets, r.num_sds, r.grid_nz)
            else:
                self.A_rz = self.A0_z = None
        for f in ("photons", "steps", "steps_boundary", "steps_interaction", "ms_first_pass_kernel",
                  "ms_replay_kernel", "ms_dump_kernel", "ms_first_pass_wall", "ms_replay_wall",
                  "first_pass_launches", "block_size", "grid_size", "chunk", "refill_min", "boundary_min",
                  "ms_drain", "ms_clock", "drain_launches"):
            setattr(self, f, getattr(r, f))
        self.lane_blocks = np.array(list(r.lane_blocks), dtype=np.uint64)

    def pathlength_records(self, set_index, sds):
        """[weight, PL_1..PL_L, steps] rows of one (set, 1-based SDS) in photon order."""
        g = set_index * self.num_sds + (sds - 1)
        return self.pathlength[self.pl_offset[g]:self.pl_offset[g + 1]]


class Context:
    """One GPU; wraps mcml_create / mcml_run / mcml_run_batch / mcml_destroy."""

    def __init__(self, device=0):
        self.lib = load_library()
        err = C.c_int(0)
        self.h = self.lib.mcml_create(device, C.byref(err))
        if not self.h:
            raise MCMLError(err.value, "mcml_create(%d)" % device)

    @property
    def device_name(self):
        return self.lib.mcml_device_name(self.h).decode()

    def _check(self, rc):
        if rc:
            raise MCMLError(rc, self.lib.mcml_last_error(self.h).decode(errors="replace"))

    def run_raw(self, sim, opts=None, seed=1):
        """Run and return the raw abi.Result (caller frees with free_raw)."""
        r = abi.Result()
        o = opts if opts is not None else make_opts()
        self._check(self.lib.mcml_run(self.h, C.byref(sim), C.byref(o), seed, C.byref(r)))
        return r

    def free_raw(self, r):
        self.lib.mcml_result_free(C.byref(r))

    def run(self, sim, opts=None, seed=1):
        return self._wrap(self.run_raw(sim, opts, seed))

    def _wrap(self, r):
        """RunResult owning r (freed when its last array is dropped)."""
        try:
            return RunResult(r, lib=self.lib)
        except BaseException:
            self.lib.mcml_result_free(C.byref(r))
            raise

    def run_batch(self, sims, opts=None, seed=1):
        arr = (abi.Sim * len(sims))(*sims)
        r = abi.Result()
        o = opts if opts is not None else make_opts()
        self._check(self.lib.mcml_run_batch(self.h, arr, len(sims), C.byref(o), seed, C.byref(r)))
        return self._wrap(r)

    def autotune(self, sims, trial_photons=10_000_000, opts=None):
        """Launch shape picked by mcml_autotune for these sets (returns a filled Opts)."""
        if isinstance(sims, abi.Sim):
            sims = [sims]
        arr = (abi.Sim * len(sims))(*sims)
        o = opts if opts is not None else make_opts()
        self._check(self.lib.mcml_autotune(self.h, arr, len(sims), int(trial_photons), C.byref(o)))
        return o

    def white_rescale(self, sims, result, mua_new, with_squares=False):
        """White Monte Carlo: detected weight per (set, mua' vector, SDS) that the replayed
        photons of `result` would carry at absorption `mua_new` [nsets, m, L] (1/cm);
        same units as RunResult.reflectance.  with_squares also returns the sums of the
        squared rescaled weights."""
        if isinstance(sims, abi.Sim):
            sims = [sims]
        mua_new = np.ascontiguousarray(mua_new, dtype=np.float32)
        nsets, m, L = mua_new.shape
        assert nsets == result.nsets and L == result.num_layers and result.replayed
        r = abi.Result()
        r.nsets, r.num_sds, r.num_layers, r.replayed = result.nsets, result.num_sds, result.num_layers, 1
        off = np.ascontiguousarray(result.pl_offset, dtype=np.uint64)
        rows = np.ascontiguousarray(result.pathlength, dtype=np.float32)
        r.pl_offset = off.ctypes.data_as(C.POINTER(C.c_uint64))
        r.pathlength = rows.ctypes.data_as(C.POINTER(C.c_float))
        out = np.zeros((nsets, m, result.num_sds), np.float64)
        sq = np.zeros_like(out) if with_squares else None
        arr = (abi.Sim * len(sims))(*sims)
        dp = C.POINTER(C.c_double)
        self._check(self.lib.mcml_white_rescale(self.h, arr, C.byref(r), mua_new.ctypes.data_as(C.POINTER(C.c_float)),
                                                m, out.ctypes.data_as(dp), sq.ctypes.data_as(dp) if with_squares else None))
        return (out, sq) if with_squares else out

    def close(self):
        if self.h:
            self.lib.mcml_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
