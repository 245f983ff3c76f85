"""ctypes binding of libmcx_hip.so (the C ABI declared in include/mcx_hip.h).

The library is the product: every sweep, exchange and weight update runs in
its HIP kernels.  There is no CPU fallback -- if the library is missing or no
GPU is visible, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG), "libmcx_hip.so")

# status codes
MCX_OK, MCX_EINVAL, MCX_EBOUNDS, MCX_EDOMAIN, MCX_ENOMEM, MCX_EHIP, MCX_ESTATE = range(7)
ISING, BLUME_CAPEL = 0, 1
METROPOLIS, HEATBATH, MULTICANONICAL, WANGLANDAU, GLAUBER = 0, 1, 2, 3, 4
MUCA_PAIR_SPIN2 = 5  # Boltzmann pair term + multicanonical sum s^2 (Blume-Capel walkers)
INIT_UP, INIT_DOWN, INIT_ZERO, INIT_RANDOM = 0, 1, 2, 3
RUN_RECORD, RUN_PLAIN = 1, 2
STATE_SERIES = 1  # MCX_STATE_SERIES: snapshots carry the recorded series
N_MOMENTS = 6


class MCXError(RuntimeError):
    """ErrorException analogue: HIP failures, out-of-memory, bad state."""


class MCXArgumentError(ValueError):
    """ArgumentError analogue (MCX_EINVAL)."""


class MCXBoundsError(IndexError):
    """BoundsError analogue (MCX_EBOUNDS)."""


class MCXAssertionError(AssertionError):
    """AssertionError analogue (MCX_EDOMAIN)."""


class MCXStateError(MCXError):
    """Handle used in a state that does not allow the call (MCX_ESTATE)."""


_EXC = {MCX_EINVAL: MCXArgumentError, MCX_EBOUNDS: MCXBoundsError,
        MCX_EDOMAIN: MCXAssertionError, MCX_ENOMEM: MemoryError, MCX_EHIP: MCXError,
        MCX_ESTATE: MCXStateError}


class BatchDesc(C.Structure):
    _fields_ = [("model", C.c_int32), ("dim", C.c_int32), ("n_chains", C.c_int32),
                ("L_uniform", C.c_int32), ("L", C.POINTER(C.c_int32)), ("seed", C.c_uint64),
                ("chain_keys", C.POINTER(C.c_uint32)), ("D_uniform", C.c_double),
                ("D", C.POINTER(C.c_double)), ("series_capacity", C.c_int64)]


class LaunchShape(C.Structure):
    _fields_ = [("threads", C.c_int32), ("chains_per_group", C.c_int32),
                ("lanes_per_word", C.c_int32), ("pad", C.c_int32)]


class ServerDesc(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_jobs", C.c_int32), ("n_streams", C.c_int32),
                ("quantum_ms", C.c_double), ("tile_sites", C.c_int64),
                ("band_sweeps", C.c_int32), ("pad_", C.c_int32)]


class JobDesc(C.Structure):
    _fields_ = [("model", C.c_int32), ("dim", C.c_int32), ("L", C.c_int32), ("alg", C.c_int32),
                ("beta", C.c_double), ("D", C.c_double), ("seed", C.c_uint64),
                ("key", C.c_uint32), ("nsweeps", C.c_int64), ("tag", C.c_uint64)]


class JobResult(C.Structure):
    _fields_ = [("job_id", C.c_uint64), ("tag", C.c_uint64), ("status", C.c_int32),
                ("L", C.c_int32), ("nsweeps", C.c_int64), ("accepted", C.c_int64),
                ("E_pair", C.c_int32), ("M", C.c_int32), ("Q", C.c_int32),
                ("slo_met", C.c_int32),
                ("moments", C.c_double * N_MOMENTS), ("latency_ms", C.c_double),
                ("queue_ms", C.c_double), ("device_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("pending", C.c_int64),
                ("active", C.c_int64), ("quanta", C.c_int64), ("attempts", C.c_double),
                ("busy_ms", C.c_double), ("wall_ms", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("max_ms", C.c_double), ("attempts_per_ns", C.c_double),
                ("busy_attempts_per_ns", C.c_double)]


_lib = None
_P = C.c_void_p
_i32, _i64, _u32, _u64, _f64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double
_pi8, _pi32, _pi64 = C.POINTER(C.c_int8), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
_pu32, _pu64, _pf64 = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_double)

# name -> argtypes; every function returns int32 status unless noted
SIGNATURES = {
    "mcx_ctx_create": [_i32, C.POINTER(_P)],
    "mcx_ctx_destroy": [_P],
    "mcx_version": [_pi32, _pi32],
    "mcx_batch_create": [_P, C.POINTER(BatchDesc), C.POINTER(_P)],
    "mcx_batch_destroy": [_P],
    "mcx_batch_num_sites": [_P, _i32, _pi64],
    "mcx_batch_set_betas": [_P, _pf64],
    "mcx_batch_set_tables": [_P, _i32, _pf64, _i32],
    "mcx_batch_table_layout": [_P, _i32, _pi32],
    "mcx_batch_init_spins": [_P, _i32],
    "mcx_batch_set_spins": [_P, _i32, _pi8, C.c_size_t],
    "mcx_batch_get_spins": [_P, _i32, _pi8, C.c_size_t],
    "mcx_batch_run": [_P, _i32, _i64, _u32],
    "mcx_batch_last_run_ms": [_P, _pf64],
    "mcx_batch_set_launch_shape": [_P, _i32, C.POINTER(LaunchShape)],
    "mcx_batch_get_launch_shape": [_P, _i32, C.POINTER(LaunchShape)],
    "mcx_batch_get_observables": [_P, _pi32, _pi32, _pi32],
    "mcx_batch_series_length": [_P, _pi64],
    "mcx_batch_get_series": [_P, _i32, _i64, _i64, _pi32, _pi32, _pi32],
    "mcx_batch_reset_series": [_P],
    "mcx_batch_get_moments": [_P, _pf64],
    "mcx_batch_get_counters": [_P, _pi64, _pi64],
    "mcx_batch_reset_counters": [_P],
    "mcx_batch_get_sweeps": [_P, _pu64],
    "mcx_batch_set_sweeps": [_P, _pu64],
    "mcx_batch_set_streams": [_P, _u64, _pu32],
    "mcx_batch_tau_int": [_P, _i32, _i64, _i64, _i32, _i32, _f64, _pf64],
    "mcx_batch_tau_int_noalloc": [_P, _i32, _i64, _i64, _i32, _i32, _f64, _pf64],
    "mcx_batch_state_size": [_P, _u32, C.POINTER(C.c_size_t)],
    "mcx_batch_save_state": [_P, _u32, _P, C.c_size_t],
    "mcx_batch_load_state": [_P, _P, C.c_size_t],
}
OPTIONAL = {}  # filled by modules that extend the ABI (pt, gen, server)


def declare(name, argtypes):
    """Register a C-ABI function signature (called at import by the modules)."""
    OPTIONAL[name] = argtypes
    if _lib is not None:
        f = getattr(_lib, name)
        f.argtypes = argtypes
        f.restype = C.c_int32


def load():
    """Load libmcx_hip.so.  Raises MCXError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MCXError(f"libmcx_hip.so not built at {LIB_PATH}; run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        for name, argt in list(SIGNATURES.items()) + list(OPTIONAL.items()):
            f = getattr(lib, name)
            f.argtypes = argt
            f.restype = C.c_int32
        lib.mcx_last_error.argtypes = []
        lib.mcx_last_error.restype = C.c_char_p
        _lib = lib
    return _lib


def call(name, *args):
    """Call a C-ABI function and map a non-zero status to the matching exception."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != MCX_OK:
        msg = lib.mcx_last_error().decode(errors="replace")
        raise _EXC.get(st, MCXError)(msg if msg.startswith(name) else f"{name}: {msg}")


def exported_symbols():
    return sorted(list(SIGNATURES) + list(OPTIONAL) + ["mcx_last_error"])


class Context:
    """A device context (mcx_ctx): one HIP device, one stream."""

    def __init__(self, device: int = 0):
        self._h = _P()
        call("mcx_ctx_create", device, C.byref(self._h))
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            call("mcx_ctx_destroy", self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
