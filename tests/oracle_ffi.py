"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests and the CPU-baseline
leg of bench.py.  The oracle is the checker only; nothing in the product imports this."""
import ctypes as C
import os
import subprocess

import numpy as np

from mcml_gpu_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None

_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        build_oracle()
    L = C.CDLL(ORACLE_SO)
    D = C.POINTER(abi.Derived)
    L.oracle_derive.argtypes = [C.POINTER(abi.Sim), D]
    L.oracle_derive.restype = C.c_int
    L.oracle_first_pass.argtypes = [D, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64,
                                    C.c_void_p, C.c_int64, _u64p, C.c_int]
    L.oracle_first_pass.restype = C.c_int64
    L.oracle_replay.argtypes = [D, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int, _fp,
                                C.c_void_p, C.c_void_p]
    L.oracle_replay.restype = C.c_int
    L.oracle_uniforms.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, _fp, C.c_int]
    L.oracle_stream_state.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64,
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
    L.oracle_multiplier_table.argtypes = [C.POINTER(C.c_uint32)]
    L.oracle_multiplier_table.restype = C.POINTER(C.c_uint32)
    _u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
    L.oracle_neglog_u32_n.argtypes = [_u32p, _fp, C.c_int64]
    L.oracle_sincos_u32_n.argtypes = [_u32p, _fp, _fp, C.c_int64]
    L.oracle_tables.argtypes = [C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_int),
                                C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_int)]
    L.oracle_sincos_turns_n.argtypes = [_fp, _fp, _fp, C.c_int64]
    L.oracle_acos_n.argtypes = [_fp, _fp, C.c_int64]
    L.oracle_spin_n.argtypes = [C.c_float, _fp, C.c_uint64, _fp, _fp, C.c_int64]
    L.oracle_reflect_fraction.argtypes = [D, C.c_int, C.c_int, _fp, C.c_uint64, C.c_int64, _fp]
    L.oracle_reflect_fraction.restype = C.c_double
    L.oracle_launch_n.argtypes = [D, C.c_uint64, C.c_int64, _fp]
    L.oracle_roulette_mean.argtypes = [C.c_float, C.c_uint64, C.c_int64]
    L.oracle_roulette_mean.restype = C.c_double
    L.oracle_steps_n.argtypes = [D, C.c_int, C.c_uint64, _fp, C.c_int64]
    L.oracle_detect_state.argtypes = [D, _fp, C.c_int, _fp]
    L.oracle_detect_state.restype = C.c_uint32
    L.oracle_ring_weight.argtypes = [C.c_float] * 4
    L.oracle_ring_weight.restype = C.c_float
    L.oracle_trace.argtypes = [D, C.c_uint64, C.c_uint32, C.c_uint64, C.c_int, _fp, C.c_int64, _fp,
                               C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
    L.oracle_trace.restype = C.c_int64
    L.oracle_bin.argtypes = [D, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    _lib = L
    return L


def derive(sim):
    d = abi.Derived()
    rc = lib().oracle_derive(C.byref(sim), C.byref(d))
    if rc != 0:
        raise ValueError("oracle_derive failed: %d" % rc)
    return d


def first_pass(d, seed, k0, k1, set_index=0, nthreads=0):
    """Records (structured array: key, sds, weight) and counters [steps, boundary, interaction]."""
    cap = max(4096, (k1 - k0) // 8)
    while True:
        counts = np.zeros(3, dtype=np.uint64)
        rec = np.zeros(cap, dtype=abi.RECORD_DTYPE)
        n = lib().oracle_first_pass(C.byref(d), seed, set_index, k0, k1, rec.ctypes.data, cap, counts,
                                    nthreads)
        if n < 0:
            raise RuntimeError("oracle_first_pass failed: %d" % n)
        if n <= cap:
            return rec[:n], counts
        cap = int(n)


def replay(d, seed, k, sds, set_index=0, grids=None):
    """Replay photon k for detector sds.  grids = (A_rz[nz*nr] u64, A0_z[nz] u64) or None."""
    L = d.num_layers
    rec = np.zeros(L + 2, dtype=np.float32)
    a_rz = grids[0].ctypes.data if grids is not None else None
    a0 = grids[1].ctypes.data if grids is not None else None
    hit = lib().oracle_replay(C.byref(d), seed, set_index, k, sds, rec, a_rz, a0)
    return hit, rec


TRACE_COLS = 6  # x, y, z, drop, first, layer


def trace(d, seed, k, sds, set_index=0, grids=None):
    """Replay photon k for detector sds and return (hit, record, interactions), where
    interactions is an [n, 6] array of (x, y, z, drop, first, layer) per drop."""
    L = d.num_layers
    rec = np.zeros(L + 2, dtype=np.float32)
    cap = 4096
    while True:
        buf = np.zeros(cap * TRACE_COLS, dtype=np.float32)
        a_rz = grids[0].ctypes.data if grids is not None else None
        a0 = grids[1].ctypes.data if grids is not None else None
        if grids is not None:
            grids[0][:] = 0
            grids[1][:] = 0
        hit = C.c_int()
        n = lib().oracle_trace(C.byref(d), seed, set_index, k, sds, buf, cap, rec, a_rz, a0, C.byref(hit))
        if n <= cap:
            return hit.value, rec, buf.reshape(cap, TRACE_COLS)[:n]
        cap = int(n)


def bin_of(d, x, y, z):
    zb, rb = C.c_int(), C.c_int()
    lib().oracle_bin(C.byref(d), x, y, z, C.byref(zb), C.byref(rb))
    return zb.value, rb.value
