"""ctypes bindings to libsipgpu.so, the gfx950 SIP contract/permute library.

The drop-in functions mirror the reference's tensor-kernel interface
(src/sip/tensor_algebra/tensor_ops_c_prototypes.h): same names without the
trailing underscores, same argument meaning, same ierr codes, host arrays in
and out.  The native classes wrap device blocks, streams, events and planned
batches.

There is no CPU fallback: if the library is missing this module raises on
first use, and every compute call goes to the GPU.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SIPGPU_LIB selects another build of the library (A/B timing of variants).
LIB_PATH = os.environ.get("SIPGPU_LIB") or os.path.join(_HERE, "libsipgpu.so")
MAX_RANK = 6

ERRORS = {
    -101: "SIPGPU_EPATTERN", -102: "SIPGPU_ERANK", -103: "SIPGPU_EPERM",
    -104: "SIPGPU_ESIZE", -105: "SIPGPU_EDEVICE", -106: "SIPGPU_ENOMEM",
    -107: "SIPGPU_EARG", -108: "SIPGPU_EHAZARD", -109: "SIPGPU_EBUDGET",
    -110: "SIPGPU_ENOTFOUND", -111: "SIPGPU_ECAPTURE",
}

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

EXPORTED = [
    "tensor_size_by_shape_", "get_contraction_ptrn_", "tensor_block_copy__",
    "tensor_block_contract__", "sipgpu_last_error", "sipgpu_version", "sipgpu_init",
    "sipgpu_device_count", "sipgpu_device_synchronize", "sipgpu_malloc", "sipgpu_free",
    "sipgpu_memcpy_h2d", "sipgpu_memcpy_d2h", "sipgpu_memcpy_d2d", "sipgpu_memset",
    "sipgpu_stream_create", "sipgpu_stream_destroy", "sipgpu_stream_synchronize",
    "sipgpu_event_create", "sipgpu_event_destroy", "sipgpu_event_record",
    "sipgpu_event_synchronize", "sipgpu_event_elapsed_ms", "sipgpu_permute",
    "sipgpu_contract", "sipgpu_batch_plan", "sipgpu_batch_execute", "sipgpu_batch_destroy",
    "sipgpu_batch_info", "sipgpu_plan_tables", "sipgpu_plan_permute", "sipgpu_measure_peaks",
    "sipgpu_measured_clock_mhz", "sipgpu_batch_plan_split", "sipgpu_batch_split_info",
    "sipgpu_fill_random", "sipgpu_energy_denominator", "sipgpu_energy_sum", "sipgpu_batch_kind",
    "sipgpu_batch_execute_stages", "sipgpu_graph_begin", "sipgpu_graph_end",
    "sipgpu_graph_launch", "sipgpu_graph_nodes", "sipgpu_graph_destroy", "sipgpu_batch_tune",
    "sipgpu_batch_plan_choice", "sipgpu_tune_save", "sipgpu_tune_load", "sipgpu_tune_clear",
    "sipgpu_batch_plan_split_tol", "tensor_block_init__", "tensor_block_scale__",
    "tensor_block_norm2__", "tensor_block_slice__", "tensor_block_insert__", "tensor_block_add__",
    "sipgpu_block_init", "sipgpu_block_scale", "sipgpu_block_add", "sipgpu_block_norm2",
    "sipgpu_block_slice", "sipgpu_block_insert", "sipgpu_plan_slice", "sipgpu_cache_remove",
    "sipgpu_cache_scope_enter", "sipgpu_cache_scope_add", "sipgpu_cache_scope_leave",
    "sipgpu_cache_set_persistent", "sipgpu_cache_end_program", "sipgpu_block_add_many",
    "sipgpu_energy_sum_parts", "sipgpu_stream_wait_event",
    "sipgpu_cache_create", "sipgpu_cache_destroy", "sipgpu_cache_insert", "sipgpu_cache_acquire",
    "sipgpu_cache_release", "sipgpu_cache_read", "sipgpu_cache_flush", "sipgpu_cache_evict",
    "sipgpu_cache_get_stats", "sipgpu_server_create", "sipgpu_server_submit",
    "sipgpu_server_wait", "sipgpu_server_replay", "sipgpu_server_get_stats",
    "sipgpu_server_destroy", "sipgpu_contract_noalloc", "sipgpu_server_submit_permute",
    "sipgpu_server_lanes", "sipgpu_worker_create", "sipgpu_worker_destroy",
    "sipgpu_worker_push_selector", "sipgpu_worker_stack_depth", "sipgpu_worker_contract",
    "sipgpu_worker_permute", "sipgpu_worker_contract_to_scalar", "sipgpu_worker_read",
    "sipgpu_worker_write", "sipgpu_worker_flush", "sipgpu_worker_get_stats",
    "sipgpu_permute_rhs_to_lhs", "sipgpu_triples_create", "sipgpu_triples_info",
    "sipgpu_triples_pack", "sipgpu_triples_run", "sipgpu_triples_energy", "sipgpu_triples_tuple",
    "sipgpu_triples_destroy", "sipgpu_permute_batch_plan", "sipgpu_permute_batch_execute",
    "sipgpu_permute_batch_info", "sipgpu_permute_batch_destroy", "sipgpu_measure_lds",
]


class SipGpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "error"), code, msg))
        self.code = code


class Contraction(ctypes.Structure):
    """sipgpu_contraction of include/sipgpu.h."""
    _fields_ = [("l", _vp), ("r", _vp), ("d", _vp), ("lrank", ctypes.c_int),
                ("rrank", ctypes.c_int), ("drank", ctypes.c_int),
                ("ldims", ctypes.c_int * MAX_RANK), ("rdims", ctypes.c_int * MAX_RANK),
                ("ddims", ctypes.c_int * MAX_RANK), ("ptrn", ctypes.c_int * (2 * MAX_RANK)),
                ("accumulate", ctypes.c_int)]


class Permutation(ctypes.Structure):
    """sipgpu_permutation of include/sipgpu.h."""
    _fields_ = [("in_", _vp), ("out", _vp), ("rank", ctypes.c_int),
                ("dims", ctypes.c_int * MAX_RANK), ("dim_transp", ctypes.c_int * (MAX_RANK + 1))]


def permutation(dims, dim_transp, din, dout):
    """A Permutation entry: block `din` (extents `dims`) into `dout`, with
    dim_transp = [sign, new position of each old dimension (1-based)]."""
    e = Permutation()
    e.in_, e.out, e.rank = din, dout, len(dims)
    for i, x in enumerate(dims):
        e.dims[i] = int(x)
    for i, x in enumerate(dim_transp):
        e.dim_transp[i] = int(x)
    return e


class Request(ctypes.Structure):
    """sipgpu_request of include/sipgpu.h."""
    _fields_ = [("ptrn", ctypes.c_int * (2 * MAX_RANK)), ("lrank", ctypes.c_int),
                ("rrank", ctypes.c_int), ("drank", ctypes.c_int),
                ("ldims", ctypes.c_int * MAX_RANK), ("rdims", ctypes.c_int * MAX_RANK),
                ("ddims", ctypes.c_int * MAX_RANK), ("lkey", ctypes.c_ulonglong),
                ("rkey", ctypes.c_ulonglong), ("dkey", ctypes.c_ulonglong),
                ("accumulate", ctypes.c_int), ("moduli", ctypes.c_int)]


class TuneResult(ctypes.Structure):
    """sipgpu_tune_result of include/sipgpu.h."""
    _fields_ = [("candidates", ctypes.c_int), ("best_ms", ctypes.c_double),
                ("default_ms", ctypes.c_double), ("path", ctypes.c_int), ("config", ctypes.c_int),
                ("merge", ctypes.c_int), ("tail", ctypes.c_int), ("tile", ctypes.c_int)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class CacheStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_ulonglong) for n in
                ("hits", "misses", "evictions", "spills", "bytes_h2d", "bytes_d2h",
                 "resident_bytes", "budget_bytes", "blocks")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class WorkerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_ulonglong) for n in
                ("contractions", "permutes", "scalars", "reads", "writes", "waits", "flushes",
                 "queued")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_ulonglong), ("requests", ctypes.c_ulonglong),
                ("max_batch", ctypes.c_ulonglong), ("flops", ctypes.c_double)]

    def as_dict(self):
        return {n: (float(getattr(self, n)) if n == "flops" else int(getattr(self, n)))
                for n, _ in self._fields_}


_lib = None


def lib():
    """Load libsipgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsipgpu.so is not built (%s); run aces4_amd/build.py"
                               % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.tensor_size_by_shape_.restype = ctypes.c_longlong
        L.sipgpu_last_error.restype = ctypes.c_char_p
        L.sipgpu_version.restype = ctypes.c_char_p
        L.sipgpu_malloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
        L.sipgpu_free.argtypes = [_vp]
        for f in ("sipgpu_memcpy_h2d", "sipgpu_memcpy_d2h", "sipgpu_memcpy_d2d"):
            getattr(L, f).argtypes = [_vp, _vp, ctypes.c_size_t]
        L.sipgpu_memset.argtypes = [_vp, ctypes.c_int, ctypes.c_size_t]
        L.sipgpu_stream_create.argtypes = [ctypes.POINTER(_vp)]
        L.sipgpu_stream_destroy.argtypes = [_vp]
        L.sipgpu_stream_synchronize.argtypes = [_vp]
        L.sipgpu_event_create.argtypes = [ctypes.POINTER(_vp)]
        L.sipgpu_event_destroy.argtypes = [_vp]
        L.sipgpu_event_record.argtypes = [_vp, _vp]
        L.sipgpu_stream_wait_event.argtypes = [_vp, _vp]
        L.sipgpu_event_synchronize.argtypes = [_vp]
        L.sipgpu_event_elapsed_ms.argtypes = [_vp, _vp, ctypes.POINTER(ctypes.c_float)]
        L.sipgpu_permute.argtypes = [_vp, ctypes.c_int, _ip, _ip, _vp, _vp]
        L.sipgpu_permute_batch_plan.argtypes = [ctypes.POINTER(Permutation), ctypes.c_int,
                                                ctypes.POINTER(_vp)]
        L.sipgpu_permute_batch_execute.argtypes = [_vp, _vp]
        L.sipgpu_permute_batch_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_longlong),
                                                ctypes.POINTER(ctypes.c_int)]
        L.sipgpu_permute_batch_destroy.argtypes = [_vp]
        L.sipgpu_contract.argtypes = [_vp, _ip, _vp, ctypes.c_int, _ip, _vp, ctypes.c_int, _ip,
                                      _vp, ctypes.c_int, _ip, ctypes.c_int]
        L.sipgpu_contract_noalloc.argtypes = [_vp, _ip, _vp, ctypes.c_int, _ip, _vp, ctypes.c_int, _ip,
                                      _vp, ctypes.c_int, _ip, ctypes.c_int]
        L.sipgpu_batch_plan.argtypes = [ctypes.POINTER(Contraction), ctypes.c_int,
                                        ctypes.POINTER(_vp)]
        L.sipgpu_batch_execute.argtypes = [_vp, _vp]
        L.sipgpu_batch_plan_split.argtypes = [ctypes.POINTER(Contraction), ctypes.c_int,
                                              ctypes.c_int, ctypes.POINTER(_vp)]
        L.sipgpu_batch_split_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_int),
                                              ctypes.POINTER(ctypes.c_int)]
        L.sipgpu_batch_kind.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
        L.sipgpu_batch_execute_stages.argtypes = [_vp, ctypes.c_int, _vp]
        L.sipgpu_batch_tune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(_vp), ctypes.POINTER(TuneResult)]
        L.sipgpu_batch_plan_choice.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.sipgpu_batch_plan_split_tol.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                                  ctypes.POINTER(_vp)]
        L.sipgpu_tune_save.argtypes = [ctypes.c_char_p]
        L.sipgpu_tune_load.argtypes = [ctypes.c_char_p]
        L.sipgpu_graph_begin.argtypes = [_vp]
        L.sipgpu_graph_end.argtypes = [_vp, ctypes.POINTER(_vp)]
        L.sipgpu_graph_launch.argtypes = [_vp, _vp]
        L.sipgpu_graph_nodes.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
        L.sipgpu_graph_destroy.argtypes = [_vp]
        L.sipgpu_batch_destroy.argtypes = [_vp]
        L.sipgpu_batch_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]
        L.sipgpu_plan_tables.argtypes = [_ip, ctypes.c_int, _ip, ctypes.c_int, _ip,
                                         ctypes.c_int, _ip, ctypes.POINTER(ctypes.c_longlong),
                                         _ip, _ip, _ip, _ip, _ip, _ip, _ip]
        L.sipgpu_plan_permute.argtypes = [ctypes.c_int, _ip, _ip, _ip]
        L.sipgpu_measured_clock_mhz.restype = ctypes.c_double
        L.sipgpu_measure_lds.argtypes = [ctypes.POINTER(ctypes.c_double)]
        _u64 = ctypes.c_ulonglong
        L.sipgpu_cache_create.argtypes = [ctypes.c_size_t, ctypes.POINTER(_vp)]
        L.sipgpu_cache_destroy.argtypes = [_vp]
        L.sipgpu_cache_insert.argtypes = [_vp, _u64, ctypes.c_size_t, _vp, ctypes.c_int, _u64,
                                          _u64, ctypes.c_double, ctypes.c_double]
        L.sipgpu_cache_acquire.argtypes = [_vp, _vp, _u64, ctypes.c_int, ctypes.POINTER(_vp)]
        L.sipgpu_cache_release.argtypes = [_vp, _vp, _u64]
        L.sipgpu_cache_read.argtypes = [_vp, _u64, _vp]
        L.sipgpu_cache_flush.argtypes = [_vp, _u64, ctypes.c_int]
        L.sipgpu_cache_evict.argtypes = [_vp, _u64]
        L.sipgpu_cache_get_stats.argtypes = [_vp, ctypes.POINTER(CacheStats)]
        L.sipgpu_cache_remove.argtypes = [_vp, _u64]
        L.sipgpu_cache_scope_enter.argtypes = [_vp]
        L.sipgpu_cache_scope_add.argtypes = [_vp, _u64]
        L.sipgpu_cache_scope_leave.argtypes = [_vp]
        L.sipgpu_cache_set_persistent.argtypes = [_vp, _u64, ctypes.c_int]
        L.sipgpu_cache_end_program.argtypes = [_vp, ctypes.POINTER(ctypes.c_ulonglong)]
        L.sipgpu_server_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp)]
        L.sipgpu_server_submit.argtypes = [_vp, ctypes.POINTER(Request), ctypes.POINTER(_u64)]
        L.sipgpu_server_wait.argtypes = [_vp, _u64, ctypes.POINTER(ctypes.c_double)]
        L.sipgpu_server_replay.argtypes = [_vp, ctypes.POINTER(Request),
                                           ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double)]
        L.sipgpu_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
        L.sipgpu_server_destroy.argtypes = [_vp]
        L.sipgpu_server_submit_permute.argtypes = [_vp, ctypes.c_int, _ip, _ip, _u64, _u64,
                                                   ctypes.POINTER(_u64)]
        L.sipgpu_server_lanes.argtypes = [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_u64)]
        L.sipgpu_worker_create.argtypes = [_vp, _vp, ctypes.POINTER(_vp)]
        L.sipgpu_worker_destroy.argtypes = [_vp]
        L.sipgpu_worker_push_selector.argtypes = [_vp, _u64, ctypes.c_int, _ip, _ip]
        L.sipgpu_worker_stack_depth.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
        L.sipgpu_worker_contract.argtypes = [_vp, _u64, ctypes.c_int, _ip, _ip, ctypes.c_int,
                                             ctypes.c_int, ctypes.POINTER(_u64)]
        L.sipgpu_worker_permute.argtypes = [_vp, _ip, ctypes.POINTER(_u64)]
        L.sipgpu_worker_contract_to_scalar.argtypes = [_vp, ctypes.POINTER(ctypes.c_double)]
        L.sipgpu_worker_read.argtypes = [_vp, _u64, _vp]
        L.sipgpu_worker_write.argtypes = [_vp, _u64, ctypes.c_int, _ip, _vp]
        L.sipgpu_worker_flush.argtypes = [_vp, ctypes.c_int]
        L.sipgpu_worker_get_stats.argtypes = [_vp, ctypes.POINTER(WorkerStats)]
        L.sipgpu_permute_rhs_to_lhs.argtypes = [ctypes.c_int, _ip, _ip, _ip]
        L.sipgpu_triples_create.argtypes = [ctypes.c_int, _ip, ctypes.c_int, _ip, _vp, _vp, _vp,
                                            _vp, _vp, ctypes.POINTER(_vp)]
        L.sipgpu_triples_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_longlong),
                                          ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_size_t)]
        L.sipgpu_triples_pack.argtypes = [_vp, _vp]
        L.sipgpu_triples_run.argtypes = [_vp, ctypes.c_longlong, ctypes.c_longlong, _vp]
        L.sipgpu_triples_energy.argtypes = [_vp, ctypes.c_double, _vp, _vp]
        L.sipgpu_triples_tuple.argtypes = [_vp, _ip, ctypes.c_double, _vp, _vp, _vp]
        L.sipgpu_triples_destroy.argtypes = [_vp]
        _pp = ctypes.POINTER(ctypes.c_void_p)
        L.sipgpu_energy_denominator.argtypes = [_vp, ctypes.c_int, _ip, _vp, _pp, _ip]
        L.sipgpu_energy_sum.argtypes = [_vp, ctypes.c_int, _ip, _vp, _vp, _pp, _ip,
                                        ctypes.c_double, _vp]
        L.sipgpu_energy_sum_parts.argtypes = [_vp, ctypes.c_int, _ip, _vp, _pp, ctypes.c_int, _pp,
                                              _ip, ctypes.c_double, _vp]
        L.sipgpu_fill_random.argtypes = [_vp, _vp, ctypes.c_longlong, ctypes.c_ulonglong,
                                         ctypes.c_ulonglong, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double]
        L.sipgpu_measure_peaks.argtypes = [ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
        L.tensor_block_norm2__.restype = ctypes.c_double
        L.sipgpu_block_init.argtypes = [_vp, _vp, ctypes.c_int, _ip, ctypes.c_double]
        L.sipgpu_block_scale.argtypes = [_vp, _vp, ctypes.c_int, _ip, ctypes.c_double]
        L.sipgpu_block_add.argtypes = [_vp, _vp, _vp, ctypes.c_int, _ip, ctypes.c_double]
        L.sipgpu_block_norm2.argtypes = [_vp, _vp, ctypes.c_int, _ip, _vp]
        L.sipgpu_block_add_many.argtypes = [_vp, _vp, ctypes.POINTER(_vp), ctypes.c_int,
                                            ctypes.c_int, _ip]
        for f in ("sipgpu_block_slice", "sipgpu_block_insert"):
            getattr(L, f).argtypes = [_vp, ctypes.c_int, _vp, _ip, _vp, _ip, _ip]
        L.sipgpu_plan_slice.argtypes = [ctypes.c_int, _ip, _ip, _ip, _ip, _ip,
                                        ctypes.POINTER(ctypes.c_longlong),
                                        ctypes.POINTER(ctypes.c_longlong)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SipGpuError(rc, lib().sipgpu_last_error().decode())
    return rc


def _ints(seq, n=None):
    vals = [int(x) for x in seq]
    if n is not None:
        vals = vals + [0] * (n - len(vals))
    a = np.array(vals if vals else [0], dtype=np.int32)
    return a


def _iptr(a):
    return a.ctypes.data_as(_ip)


# --------------------------------------------------------------- drop-in ----

def tensor_size_by_shape(dims):
    d = _ints(dims)
    ierr = ctypes.c_int(0)
    s = lib().tensor_size_by_shape_(ctypes.byref(ctypes.c_int(len(dims))), _iptr(d),
                                    ctypes.byref(ierr))
    return s, ierr.value


def get_contraction_ptrn(drank, lrank, rrank, aces):
    """ACES index labels [d | l | r] -> (+k / -k pattern, ierr)."""
    a = _ints(aces)
    out = np.zeros(max(lrank + rrank, 1), dtype=np.int32)
    ierr = ctypes.c_int(0)
    lib().get_contraction_ptrn_(ctypes.byref(ctypes.c_int(drank)),
                                ctypes.byref(ctypes.c_int(lrank)),
                                ctypes.byref(ctypes.c_int(rrank)), _iptr(a), _iptr(out),
                                ctypes.byref(ierr))
    return out[:lrank + rrank].copy(), ierr.value


def tensor_block_copy(dims, dim_transp, x):
    """Permute a host block on the GPU; returns (out, ierr)."""
    d = _ints(dims)
    t = _ints(dim_transp)
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.empty_like(x)
    ierr = ctypes.c_int(0)
    lib().tensor_block_copy__(ctypes.byref(ctypes.c_int(8)),
                              ctypes.byref(ctypes.c_int(len(dims))), _iptr(d), _iptr(t),
                              x.ctypes.data_as(_dp), out.ctypes.data_as(_dp), ctypes.byref(ierr))
    return out, ierr.value


def _byref_i(v):
    return ctypes.byref(ctypes.c_int(v))


def _byref_d(v):
    return ctypes.byref(ctypes.c_double(v))


def _host(x):
    return np.ascontiguousarray(x, dtype=np.float64).ravel().copy()


def tensor_block_init(dims, t, sval, rank=None):
    """Fill a host block on the GPU; returns (block, ierr).  `rank` overrides
    len(dims) (to pass a negative rank)."""
    d = _ints(dims)
    t = _host(t)
    ierr = ctypes.c_int(0)
    lib().tensor_block_init__(_byref_i(8), t.ctypes.data_as(_dp),
                              _byref_i(len(dims) if rank is None else rank), _iptr(d),
                              _byref_d(sval), ctypes.byref(ierr))
    return t, ierr.value


def tensor_block_scale(dims, t, factor, rank=None):
    d = _ints(dims)
    t = _host(t)
    ierr = ctypes.c_int(0)
    lib().tensor_block_scale__(_byref_i(8), t.ctypes.data_as(_dp),
                               _byref_i(len(dims) if rank is None else rank), _iptr(d),
                               _byref_d(factor), ctypes.byref(ierr))
    return t, ierr.value


def tensor_block_norm2(dims, t, rank=None):
    d = _ints(dims)
    t = _host(t)
    ierr = ctypes.c_int(0)
    v = lib().tensor_block_norm2__(_byref_i(8), t.ctypes.data_as(_dp),
                                   _byref_i(len(dims) if rank is None else rank), _iptr(d),
                                   ctypes.byref(ierr))
    return v, ierr.value


def tensor_block_add(dims, t0, t1, factor, rank=None):
    d = _ints(dims)
    t0 = _host(t0)
    t1 = _host(t1)
    ierr = ctypes.c_int(0)
    lib().tensor_block_add__(_byref_i(8), _byref_i(len(dims) if rank is None else rank),
                             _iptr(d), t0.ctypes.data_as(_dp), t1.ctypes.data_as(_dp),
                             _byref_d(factor), ctypes.byref(ierr))
    return t0, ierr.value


def tensor_block_slice(tens, tens_ext, slice_ext, ext_beg, out=None, rank=None):
    """Extract a slice of a host block on the GPU; returns (slice, ierr)."""
    te, se, eb = _ints(tens_ext), _ints(slice_ext), _ints(ext_beg)
    tens = np.ascontiguousarray(tens, dtype=np.float64).ravel()
    n = max(int(np.prod(slice_ext)) if len(slice_ext) else 1, 1)
    out = np.full(n, np.nan) if out is None else _host(out)
    ierr = ctypes.c_int(0)
    lib().tensor_block_slice__(_byref_i(8), _byref_i(len(tens_ext) if rank is None else rank),
                               tens.ctypes.data_as(_dp), _iptr(te), out.ctypes.data_as(_dp),
                               _iptr(se), _iptr(eb), ctypes.byref(ierr))
    return out, ierr.value


def tensor_block_insert(tens, tens_ext, slc, slice_ext, ext_beg, rank=None):
    """Insert a slice into a host block on the GPU; returns (block, ierr)."""
    te, se, eb = _ints(tens_ext), _ints(slice_ext), _ints(ext_beg)
    tens = _host(tens)
    slc = np.ascontiguousarray(slc, dtype=np.float64).ravel()
    ierr = ctypes.c_int(0)
    lib().tensor_block_insert__(_byref_i(8), _byref_i(len(tens_ext) if rank is None else rank),
                                tens.ctypes.data_as(_dp), _iptr(te), slc.ctypes.data_as(_dp),
                                _iptr(se), _iptr(eb), ctypes.byref(ierr))
    return tens, ierr.value


def tensor_block_contract(ptrn, l, ldims, r, rdims, ddims, d_init=None):
    """D = L (x)_ptrn R on the GPU with the reference's assign semantics."""
    p = _ints(ptrn)
    ld, rd, dd = _ints(ldims), _ints(rdims), _ints(ddims)
    l = np.ascontiguousarray(l, dtype=np.float64).ravel()
    r = np.ascontiguousarray(r, dtype=np.float64).ravel()
    dsize = int(np.prod(ddims)) if len(ddims) else 1
    d = (np.full(dsize, np.nan) if d_init is None
         else np.ascontiguousarray(d_init, dtype=np.float64).ravel().copy())
    ierr = ctypes.c_int(0)
    lib().tensor_block_contract__(ctypes.byref(ctypes.c_int(8)), _iptr(p), l.ctypes.data_as(_dp),
                                  ctypes.byref(ctypes.c_int(len(ldims))), _iptr(ld),
                                  r.ctypes.data_as(_dp), ctypes.byref(ctypes.c_int(len(rdims))),
                                  _iptr(rd), d.ctypes.data_as(_dp),
                                  ctypes.byref(ctypes.c_int(len(ddims))), _iptr(dd),
                                  ctypes.byref(ierr))
    return d, ierr.value


def plan_tables(ptrn, ldims, rdims, ddims):
    """Host-only: (M, N, K), o2n maps and the six offset tables of a shape."""
    p = _ints(ptrn)
    ld, rd, dd = _ints(ldims), _ints(rdims), _ints(ddims)
    mnk = (ctypes.c_longlong * 3)()
    o2n = np.zeros(len(ldims) + len(rdims) + len(ddims) + 1, np.int32)
    _check(lib().sipgpu_plan_tables(_iptr(p), len(ldims), _iptr(ld), len(rdims), _iptr(rd),
                                    len(ddims), _iptr(dd), mnk, _iptr(o2n),
                                    None, None, None, None, None, None))
    M, N, K = mnk[0], mnk[1], mnk[2]
    t = [np.zeros(max(n, 1), np.int32) for n in (K, M, K, N, M, N)]
    _check(lib().sipgpu_plan_tables(_iptr(p), len(ldims), _iptr(ld), len(rdims), _iptr(rd),
                                    len(ddims), _iptr(dd), mnk, _iptr(o2n),
                                    *[_iptr(a) for a in t]))
    lr, rr = len(ldims), len(rdims)
    maps = dict(lo2n=list(o2n[:lr]), ro2n=list(o2n[lr:lr + rr]),
                do2n=list(o2n[lr + rr:lr + rr + len(ddims)]))
    names = ("koffL", "moffL", "koffR", "noffR", "moffD", "noffD")
    return (M, N, K), maps, dict(zip(names, t))


def plan_permute(dims, dim_transp):
    """Host-only: (canonical rank, leading tile dims, row mode, plain copy)."""
    info = np.zeros(4, np.int32)
    _check(lib().sipgpu_plan_permute(len(dims), _iptr(_ints(dims)), _iptr(_ints(dim_transp)),
                                     _iptr(info)))
    return tuple(int(x) for x in info)


# ---------------------------------------------------------------- native ----

def init(device=0):
    _check(lib().sipgpu_init(device))


def device_count():
    n = ctypes.c_int(0)
    _check(lib().sipgpu_device_count(ctypes.byref(n)))
    return n.value


def measure_peaks(copy_bytes=1 << 30):
    """Measured (FP64 MFMA TFLOP/s, streaming-copy GB/s) of the device."""
    m, c = ctypes.c_double(0), ctypes.c_double(0)
    _check(lib().sipgpu_measure_peaks(ctypes.byref(m), ctypes.byref(c), copy_bytes))
    return m.value, c.value


def fill_random(ptr, n, seed, stream, dist="uniform", a=0.0, b=1.0, s=None):
    """Seeded data in HBM at device address `ptr` (n doubles): the same
    SplitMix64 sequence as workloads.uniform / workloads.normal."""
    d = {"uniform": 0, "normal": 1}[dist]
    _check(lib().sipgpu_fill_random(None if s is None else s.handle, ptr, int(n), int(seed),
                                    int(stream), d, float(a), float(b)))


def _energy_arrays(dims, eps, signs):
    r = len(dims)
    d = (ctypes.c_int * r)(*[int(x) for x in dims])
    e = (ctypes.c_void_p * r)(*[int(x) for x in eps])
    sg_ = (ctypes.c_int * r)(*[int(x) for x in signs])
    return r, d, e, sg_


def energy_denominator(dims, x, eps, signs, s=None):
    """x /= D in place (device block x, device energy vectors eps[d])."""
    r, d, e, sg_ = _energy_arrays(dims, eps, signs)
    _check(lib().sipgpu_energy_denominator(None if s is None else s.handle, r, d, x, e, sg_))


def energy_sum(dims, x, y, eps, signs, factor, out, s=None):
    """*out += factor * sum (x / D) * y   (out: device address of one double)."""
    r, d, e, sg_ = _energy_arrays(dims, eps, signs)
    _check(lib().sipgpu_energy_sum(None if s is None else s.handle, r, d, x, y, e, sg_,
                                   float(factor), out))


def energy_sum_parts(dims, t, parts, eps, signs, factor, out, s=None):
    """t = ((t + parts[0]) + ...) stored, then *out += factor * sum (t / D) * t."""
    r, d, e, sg_ = _energy_arrays(dims, eps, signs)
    arr = (ctypes.c_void_p * max(len(parts), 1))(*[int(p) for p in parts])
    _check(lib().sipgpu_energy_sum_parts(None if s is None else s.handle, r, d, t, arr, len(parts),
                                         e, sg_, float(factor), out))


def measured_clock_mhz():
    """Shader clock (MHz) seen by the last measure_peaks MFMA loop."""
    return lib().sipgpu_measured_clock_mhz()


def measure_lds():
    """Measured LDS read bandwidth, bytes per shader clock per CU."""
    v = ctypes.c_double(0)
    _check(lib().sipgpu_measure_lds(ctypes.byref(v)))
    return v.value


def synchronize():
    _check(lib().sipgpu_device_synchronize())


class DeviceArray:
    """A plain HBM allocation of float64 (or raw bytes) owned by Python."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        self.nbytes = self.n * self.dtype.itemsize
        p = _vp()
        _check(lib().sipgpu_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, x):
        x = np.ascontiguousarray(x)
        a = cls(x.size, x.dtype)
        a.upload(x)
        return a

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype).ravel()
        assert x.size == self.n
        _check(lib().sipgpu_memcpy_h2d(self.ptr, x.ctypes.data, self.nbytes))

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        _check(lib().sipgpu_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def download_range(self, offset, count):
        """Elements [offset, offset + count) as a host array."""
        assert 0 <= offset and offset + count <= self.n
        out = np.empty(int(count), dtype=self.dtype)
        _check(lib().sipgpu_memcpy_d2h(out.ctypes.data, self.at(offset),
                                       int(count) * self.dtype.itemsize))
        return out

    def upload_range(self, offset, x):
        x = np.ascontiguousarray(x, dtype=self.dtype).ravel()
        assert 0 <= offset and offset + x.size <= self.n
        _check(lib().sipgpu_memcpy_h2d(self.at(offset), x.ctypes.data, x.nbytes))

    def zero(self):
        self.fill_bytes(0)

    def fill_bytes(self, value):
        _check(lib().sipgpu_memset(self.ptr, value, self.nbytes))

    def at(self, offset):
        """Device address of element `offset`."""
        return self.ptr + int(offset) * self.dtype.itemsize

    def free(self):
        if self.ptr:
            _check(lib().sipgpu_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = _vp()
        _check(lib().sipgpu_stream_create(ctypes.byref(p)))
        self.handle = p.value

    def synchronize(self):
        _check(lib().sipgpu_stream_synchronize(self.handle))

    def wait(self, event):
        """Later work on this stream waits for `event` (no host wait)."""
        _check(lib().sipgpu_stream_wait_event(self.handle, event.handle))

    def destroy(self):
        if self.handle:
            _check(lib().sipgpu_stream_destroy(self.handle))
            self.handle = None


class Graph:
    """A captured sequence of library work, replayed by `launch`:

        g = Graph.capture(lambda s: (batch.execute(s), ...), stream)
        g.launch(stream)
    """

    def __init__(self, handle):
        self.handle = handle
        n = ctypes.c_int(0)
        _check(lib().sipgpu_graph_nodes(self.handle, ctypes.byref(n)))
        self.nodes = n.value

    @staticmethod
    def begin(stream=None):
        _check(lib().sipgpu_graph_begin(stream.handle if stream else None))

    @staticmethod
    def end(stream=None):
        p = _vp()
        _check(lib().sipgpu_graph_end(stream.handle if stream else None, ctypes.byref(p)))
        return Graph(p.value)

    @staticmethod
    def capture(fn, stream=None):
        """Runs fn(stream) under capture; the capture is ended (and the
        error re-raised) if fn raises."""
        Graph.begin(stream)
        try:
            fn(stream)
        except Exception:
            try:
                Graph.end(stream).destroy()
            except SipGpuError:
                pass
            raise
        return Graph.end(stream)

    def launch(self, stream=None):
        _check(lib().sipgpu_graph_launch(self.handle, stream.handle if stream else None))

    def destroy(self):
        if self.handle:
            _check(lib().sipgpu_graph_destroy(self.handle))
            self.handle = None


class Event:
    def __init__(self):
        p = _vp()
        _check(lib().sipgpu_event_create(ctypes.byref(p)))
        self.handle = p.value

    def record(self, stream=None):
        _check(lib().sipgpu_event_record(self.handle, stream.handle if stream else None))

    def synchronize(self):
        _check(lib().sipgpu_event_synchronize(self.handle))

    def elapsed_ms(self, later):
        ms = ctypes.c_float(0)
        _check(lib().sipgpu_event_elapsed_ms(self.handle, later.handle, ctypes.byref(ms)))
        return ms.value


def block_init(t, dims, value, stream=None, rank=None):
    """Device block `t` (address) = value, enqueued on `stream`."""
    _check(lib().sipgpu_block_init(stream.handle if stream else None, t,
                                   len(dims) if rank is None else rank, _iptr(_ints(dims)),
                                   float(value)))


def block_scale(t, dims, factor, stream=None, rank=None):
    _check(lib().sipgpu_block_scale(stream.handle if stream else None, t,
                                    len(dims) if rank is None else rank, _iptr(_ints(dims)),
                                    float(factor)))


def block_add(t0, t1, dims, factor, stream=None, rank=None):
    _check(lib().sipgpu_block_add(stream.handle if stream else None, t0, t1,
                                  len(dims) if rank is None else rank, _iptr(_ints(dims)),
                                  float(factor)))


def block_add_many(t, parts, dims, stream=None):
    """Device block t = ((t + parts[0]) + parts[1]) + ... (addresses)."""
    arr = (_vp * max(len(parts), 1))(*parts)
    _check(lib().sipgpu_block_add_many(stream.handle if stream else None, t, arr, len(parts),
                                       len(dims), _iptr(_ints(dims))))


def block_norm2(t, dims, result, stream=None, rank=None):
    """*result (device double) = sum of squares of device block `t`."""
    _check(lib().sipgpu_block_norm2(stream.handle if stream else None, t,
                                    len(dims) if rank is None else rank, _iptr(_ints(dims)),
                                    result))


def block_slice(tens, tens_ext, slc, slice_ext, ext_beg, stream=None):
    _check(lib().sipgpu_block_slice(stream.handle if stream else None, len(tens_ext), tens,
                                    _iptr(_ints(tens_ext)), slc, _iptr(_ints(slice_ext)),
                                    _iptr(_ints(ext_beg))))


def block_insert(tens, tens_ext, slc, slice_ext, ext_beg, stream=None):
    _check(lib().sipgpu_block_insert(stream.handle if stream else None, len(tens_ext), tens,
                                     _iptr(_ints(tens_ext)), slc, _iptr(_ints(slice_ext)),
                                     _iptr(_ints(ext_beg))))


def plan_slice(tens_ext, slice_ext, ext_beg):
    """Host-only slice plan: (canonical extents, tensor strides, base, count,
    span, tensor elements)."""
    cr = ctypes.c_int(0)
    ext = np.zeros(8, np.int32)
    st = (ctypes.c_longlong * 8)()
    info = (ctypes.c_longlong * 4)()
    _check(lib().sipgpu_plan_slice(len(tens_ext), _iptr(_ints(tens_ext)), _iptr(_ints(slice_ext)),
                                   _iptr(_ints(ext_beg)), ctypes.byref(cr), _iptr(ext), st, info))
    r = cr.value
    return [int(x) for x in ext[:r]], [int(x) for x in st[:r]], info[0], info[1], info[2], info[3]


def stream_wait(event, stream=None):
    """Work enqueued on `stream` (None: the library's stream) after this
    call waits for `event`."""
    _check(lib().sipgpu_stream_wait_event(stream.handle if stream else None, event.handle))


def permute(rank_dims, dim_transp, din, dout, stream=None):
    """Permute device block `din` into `dout` (device addresses)."""
    d = _ints(rank_dims)
    t = _ints(dim_transp)
    _check(lib().sipgpu_permute(stream.handle if stream else None, len(rank_dims), _iptr(d),
                                _iptr(t), din, dout))


class PermuteBatch:
    """Permutes run as one launch (sipgpu_permute_batch_*): a list of
    Permutation entries (see `permutation`) over device blocks."""

    def __init__(self, entries):
        arr = (Permutation * len(entries))(*entries)
        p = _vp()
        _check(lib().sipgpu_permute_batch_plan(arr, len(entries), ctypes.byref(p)))
        self.handle = p.value
        el, wg = ctypes.c_longlong(0), ctypes.c_int(0)
        _check(lib().sipgpu_permute_batch_info(self.handle, ctypes.byref(el), ctypes.byref(wg)))
        self.elements, self.workgroups = el.value, wg.value

    def execute(self, stream=None):
        _check(lib().sipgpu_permute_batch_execute(self.handle, stream.handle if stream else None))

    def destroy(self):
        if self.handle:
            _check(lib().sipgpu_permute_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def contract(ptrn, l, ldims, r, rdims, d, ddims, accumulate=False, stream=None):
    """Contract device blocks (addresses) asynchronously."""
    p = _ints(ptrn)
    ld, rd, dd = _ints(ldims), _ints(rdims), _ints(ddims)
    _check(lib().sipgpu_contract(stream.handle if stream else None, _iptr(p), l, len(ldims),
                                 _iptr(ld), r, len(rdims), _iptr(rd), d, len(ddims), _iptr(dd),
                                 1 if accumulate else 0))


def contract_noalloc(ptrn, l, ldims, r, rdims, d, ddims, accumulate=False, stream=None):
    """sipgpu_contract_noalloc: contract() with the library's static scratch
    (no device allocation; the shape must have run through contract() once)."""
    p = _ints(ptrn)
    ld, rd, dd = _ints(ldims), _ints(rdims), _ints(ddims)
    _check(lib().sipgpu_contract_noalloc(stream.handle if stream else None, _iptr(p), l,
                                         len(ldims), _iptr(ld), r, len(rdims), _iptr(rd), d,
                                         len(ddims), _iptr(dd), 1 if accumulate else 0))


def make_contraction(ptrn, l, ldims, r, rdims, d, ddims, accumulate=False):
    c = Contraction()
    c.l, c.r, c.d = l, r, d
    c.lrank, c.rrank, c.drank = len(ldims), len(rdims), len(ddims)
    for i, v in enumerate(ldims):
        c.ldims[i] = int(v)
    for i, v in enumerate(rdims):
        c.rdims[i] = int(v)
    for i, v in enumerate(ddims):
        c.ddims[i] = int(v)
    for i, v in enumerate(ptrn):
        c.ptrn[i] = int(v)
    c.accumulate = 1 if accumulate else 0
    return c


STAGE_PREPARE, STAGE_PRODUCT, STAGE_FINISH, STAGE_ALL = 1, 2, 4, 7
PATHS = {0: "tables", 1: "packed", 2: "split"}


def tune(contractions, moduli=None):
    """Autotunes a batch (sipgpu_batch_tune): returns the fastest plan as a
    Batch and the tuning result as a dict."""
    arr = (Contraction * len(contractions))(*contractions)
    p = _vp()
    res = TuneResult()
    _check(lib().sipgpu_batch_tune(arr, len(contractions), int(moduli or 0), ctypes.byref(p),
                                   ctypes.byref(res)))
    return Batch(contractions, moduli=moduli, _handle=p.value), res.as_dict()


def plan_choice(contractions, path, moduli=None, config=-1, merge=-1, tail=-1, tile=-1):
    """A batch planned one given way (sipgpu_batch_plan_choice)."""
    arr = (Contraction * len(contractions))(*contractions)
    p = _vp()
    code = {v: k for k, v in PATHS.items()}[path]
    _check(lib().sipgpu_batch_plan_choice(arr, len(contractions), int(moduli or 0), code,
                                          int(config), int(merge), int(tail), int(tile),
                                          ctypes.byref(p)))
    return Batch(contractions, moduli=moduli, _handle=p.value)


def tune_save(path):
    _check(lib().sipgpu_tune_save(str(path).encode()))


def tune_load(path):
    _check(lib().sipgpu_tune_load(str(path).encode()))


def tune_clear():
    _check(lib().sipgpu_tune_clear())


class Batch:
    """A planned batch of contractions: one grouped FP64-MFMA launch, or with
    `moduli` set, the split-precision (INT8 / CRT) path over that many moduli."""

    def __init__(self, contractions, moduli=None, _handle=None, tol=None):
        # the batch reads and writes the blocks these entries address: holding
        # the list holds whatever it keeps alive (workloads.contractions_for
        # attaches the workload that owns the device arrays)
        self.contractions = contractions
        arr = (Contraction * len(contractions))(*contractions)
        p = _vp()
        if _handle is not None:
            p.value = _handle
        elif tol is not None:  # split path, moduli set by the error target
            _check(lib().sipgpu_batch_plan_split_tol(arr, len(contractions), float(tol),
                                                     ctypes.byref(p)))
        elif moduli is None:
            _check(lib().sipgpu_batch_plan(arr, len(contractions), ctypes.byref(p)))
        else:
            _check(lib().sipgpu_batch_plan_split(arr, len(contractions), int(moduli),
                                                 ctypes.byref(p)))
        self.handle = p.value
        nm, bits = ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().sipgpu_batch_split_info(self.handle, ctypes.byref(nm), ctypes.byref(bits)))
        self.moduli, self.bits = nm.value, bits.value
        k = ctypes.c_int(0)
        _check(lib().sipgpu_batch_kind(self.handle, ctypes.byref(k)))
        self.kind = {0: "tables", 1: "packed", 2: "split"}[k.value]
        fl = ctypes.c_double(0)
        nt, ntl, npr = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _check(lib().sipgpu_batch_info(self.handle, ctypes.byref(fl), ctypes.byref(nt),
                                       ctypes.byref(ntl), ctypes.byref(npr)))
        self.flops, self.ntasks, self.ntiles, self.npairs = fl.value, nt.value, ntl.value, npr.value

    def execute(self, stream=None):
        _check(lib().sipgpu_batch_execute(self.handle, stream.handle if stream else None))

    def execute_stages(self, stages, stream=None):
        """Runs the stages in the bit mask `stages` (STAGE_PREPARE, STAGE_PRODUCT,
        STAGE_FINISH) on `stream`."""
        _check(lib().sipgpu_batch_execute_stages(self.handle, int(stages),
                                                 stream.handle if stream else None))

    def destroy(self):
        if self.handle:
            _check(lib().sipgpu_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


CACHE_INIT = {"undefined": 0, "zero": 1, "uniform": 2, "normal": 3}
CACHE_MODE = {"r": 1, "w": 2, "rw": 3}


class Cache:
    """Device block cache: blocks named by integer keys in an HBM arena of
    `budget_bytes`, evicted least-recently-released first, spilled to host."""

    def __init__(self, budget_bytes):
        p = _vp()
        _check(lib().sipgpu_cache_create(int(budget_bytes), ctypes.byref(p)))
        self.handle = p.value
        self._homes = {}

    def insert(self, key, elems, host=None, init="undefined", seed=0, stream=0, a=0.0, b=1.0):
        ptr = None
        if host is not None:
            host = np.ascontiguousarray(host, dtype=np.float64)
            assert host.size == elems
            self._homes[key] = host  # the cache writes spills back into it
            ptr = host.ctypes.data
        _check(lib().sipgpu_cache_insert(self.handle, key, int(elems), ptr, CACHE_INIT[init],
                                         int(seed), int(stream), float(a), float(b)))

    def acquire(self, key, mode="r", s=None):
        p = _vp()
        _check(lib().sipgpu_cache_acquire(self.handle, None if s is None else s.handle, key,
                                          CACHE_MODE[mode], ctypes.byref(p)))
        return p.value

    def release(self, key, s=None):
        _check(lib().sipgpu_cache_release(self.handle, None if s is None else s.handle, key))

    def read(self, key, elems):
        out = np.empty(int(elems))
        _check(lib().sipgpu_cache_read(self.handle, key, out.ctypes.data))
        return out

    def flush(self, key=None):
        _check(lib().sipgpu_cache_flush(self.handle, 0 if key is None else key,
                                        1 if key is None else 0))

    def evict(self, key):
        _check(lib().sipgpu_cache_evict(self.handle, key))

    def remove(self, key):
        """Forget a block without writing it back."""
        _check(lib().sipgpu_cache_remove(self.handle, key))
        self._homes.pop(key, None)

    def scope_enter(self):
        _check(lib().sipgpu_cache_scope_enter(self.handle))

    def scope_add(self, key):
        _check(lib().sipgpu_cache_scope_add(self.handle, key))

    def scope_leave(self):
        _check(lib().sipgpu_cache_scope_leave(self.handle))

    def set_persistent(self, key, persistent=True):
        _check(lib().sipgpu_cache_set_persistent(self.handle, key, 1 if persistent else 0))

    def end_program(self):
        """Remove every block not marked persistent; returns how many were kept."""
        kept = ctypes.c_ulonglong(0)
        _check(lib().sipgpu_cache_end_program(self.handle, ctypes.byref(kept)))
        return kept.value

    def stats(self):
        st = CacheStats()
        _check(lib().sipgpu_cache_get_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def destroy(self):
        if self.handle:
            _check(lib().sipgpu_cache_destroy(self.handle))
            self.handle = None


def make_request(ptrn, lkey, ldims, rkey, rdims, dkey, ddims, accumulate=False, moduli=0):
    """A server request; `moduli` > 0 runs it on the split-precision path."""
    r = Request()
    r.moduli = int(moduli)
    r.lkey, r.rkey, r.dkey = lkey, rkey, dkey
    r.lrank, r.rrank, r.drank = len(ldims), len(rdims), len(ddims)
    for i, v in enumerate(ldims):
        r.ldims[i] = int(v)
    for i, v in enumerate(rdims):
        r.rdims[i] = int(v)
    for i, v in enumerate(ddims):
        r.ddims[i] = int(v)
    for i, v in enumerate(ptrn):
        r.ptrn[i] = int(v)
    r.accumulate = 1 if accumulate else 0
    return r


class Server:
    """Continuous-batching contraction server over a Cache."""

    def __init__(self, cache, max_batch=256):
        p = _vp()
        _check(lib().sipgpu_server_create(cache.handle, int(max_batch), ctypes.byref(p)))
        self.handle = p.value

    def submit(self, req):
        t = ctypes.c_ulonglong(0)
        _check(lib().sipgpu_server_submit(self.handle, ctypes.byref(req), ctypes.byref(t)))
        return t.value

    def submit_permute(self, dims, dim_transp, src_key, dst_key):
        """Queues tensor_block_copy__(dims, dim_transp) of block src into dst."""
        d, t = _ints(dims), _ints(dim_transp)
        tk = ctypes.c_ulonglong(0)
        _check(lib().sipgpu_server_submit_permute(self.handle, len(dims), _iptr(d), _iptr(t),
                                                  src_key, dst_key, ctypes.byref(tk)))
        return tk.value

    def wait(self, ticket):
        lat = ctypes.c_double(0)
        _check(lib().sipgpu_server_wait(self.handle, ticket, ctypes.byref(lat)))
        return lat.value

    def lanes(self):
        """(device streams, batches that waited on a conflicting batch of another stream)"""
        n, w = ctypes.c_int(0), ctypes.c_ulonglong(0)
        _check(lib().sipgpu_server_lanes(self.handle, ctypes.byref(n), ctypes.byref(w)))
        return n.value, int(w.value)

    def replay(self, reqs, arrival_s):
        n = len(reqs)
        arr = (Request * n)(*reqs)
        at = np.ascontiguousarray(arrival_s, dtype=np.float64)
        lat = np.zeros(n)
        wall = ctypes.c_double(0)
        _check(lib().sipgpu_server_replay(
            self.handle, arr, at.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n,
            lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(wall)))
        return lat, wall.value

    def stats(self):
        st = ServerStats()
        _check(lib().sipgpu_server_get_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def destroy(self):
        if self.handle:
            _check(lib().sipgpu_server_destroy(self.handle))
            self.handle = None


def permute_rhs_to_lhs(lhs_ids, rhs_ids):
    """Interpreter::permute_rhs_to_lhs's 0-based old-to-new vector (host
    only): six entries, -1 beyond the rank.  Raises SIPGPU_EPERM for an
    illegal transpose."""
    lhs = _ints(list(lhs_ids) + [-1] * (MAX_RANK - len(lhs_ids)))
    rhs = _ints(list(rhs_ids) + [-1] * (MAX_RANK - len(rhs_ids)))
    out = _ints([0] * MAX_RANK)
    _check(lib().sipgpu_permute_rhs_to_lhs(len(lhs_ids), _iptr(lhs), _iptr(rhs), _iptr(out)))
    return [int(x) for x in out]


class Worker:
    """The SIP worker's contraction / permute dispatch over a Server and its
    Cache, driven in the interpreter's terms: push block selectors, then
    contract (pops right, then left), permute (pops the left-hand side, then
    the right-hand side) or contract to a scalar."""

    def __init__(self, cache, server):
        p = _vp()
        _check(lib().sipgpu_worker_create(cache.handle, server.handle, ctypes.byref(p)))
        self.handle = p.value

    def push(self, key, index_ids, extents):
        ids, ext = _ints(index_ids), _ints(extents)
        _check(lib().sipgpu_worker_push_selector(self.handle, key, len(index_ids), _iptr(ids),
                                                 _iptr(ext)))

    def depth(self):
        d = ctypes.c_int(0)
        _check(lib().sipgpu_worker_stack_depth(self.handle, ctypes.byref(d)))
        return d.value

    def contract(self, dkey, dindex_ids, dextents, accumulate=False, moduli=0):
        ids, ext = _ints(dindex_ids), _ints(dextents)
        t = ctypes.c_ulonglong(0)
        _check(lib().sipgpu_worker_contract(self.handle, dkey, len(dindex_ids), _iptr(ids),
                                            _iptr(ext), 1 if accumulate else 0, int(moduli),
                                            ctypes.byref(t)))
        return t.value

    def permute(self, compiler_perm=None):
        t = ctypes.c_ulonglong(0)
        cp = None
        if compiler_perm is not None:
            cp = _ints(list(compiler_perm) + [-1] * (MAX_RANK - len(compiler_perm)))
        _check(lib().sipgpu_worker_permute(self.handle, _iptr(cp) if cp is not None else None,
                                           ctypes.byref(t)))
        return t.value

    def contract_to_scalar(self):
        v = ctypes.c_double(0)
        _check(lib().sipgpu_worker_contract_to_scalar(self.handle, ctypes.byref(v)))
        return v.value

    def read(self, key, elems):
        out = np.empty(int(elems))
        _check(lib().sipgpu_worker_read(self.handle, key, out.ctypes.data))
        return out

    def write(self, key, extents, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        ext = _ints(extents)
        assert host.size == int(np.prod(extents))
        _check(lib().sipgpu_worker_write(self.handle, key, len(extents), _iptr(ext),
                                         host.ctypes.data))

    def flush(self, writeback=False):
        _check(lib().sipgpu_worker_flush(self.handle, 1 if writeback else 0))

    def stats(self):
        st = WorkerStats()
        _check(lib().sipgpu_worker_get_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def destroy(self):
        if self.handle:
            _check(lib().sipgpu_worker_destroy(self.handle))
            self.handle = None
