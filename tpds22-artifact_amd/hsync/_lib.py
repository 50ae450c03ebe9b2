"""ctypes binding of libhs.so (the C ABI declared in include/hs.h).

The library is built in-tree by ``make -C tpds22-artifact_amd`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared library is
missing, importing this module raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libhs.so")

HS_OK = 0
HS_ERR_INVALID = 1
HS_ERR_HIP = 2
HS_ERR_RESIDENCY = 3
HS_ERR_ABORTED = 4
HS_ERR_TIMEOUT = 5
HS_ERR_NOT_INIT = 6
HS_ERR_NOMEM = 7
HS_ERR_CHECK = 8

# hs_barrier_kind
BAR_FLAT, BAR_TWO_LEVEL, BAR_XCD, BAR_XCD_FLAT_RELEASE, BAR_TREE, BAR_LF, BAR_COOP, BAR_NAIVE_ALL = range(8)
BARRIER_NAMES = {
    "flat": BAR_FLAT, "two_level": BAR_TWO_LEVEL, "xcd": BAR_XCD,
    "xcd_flat_release": BAR_XCD_FLAT_RELEASE, "tree": BAR_TREE, "lf": BAR_LF,
    "coop": BAR_COOP, "naive_all": BAR_NAIVE_ALL,
}
# hs_mutex_kind / hs_sem_kind
MUTEX_SPIN, MUTEX_EBO, MUTEX_FA, MUTEX_SLEEP = range(4)
SEM_SPIN, SEM_EBO, SEM_SPIN_PRIORITY, SEM_EBO_PRIORITY, SEM_XCD_QUEUE = range(5)


class HSError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = _lib.hs_status_string(status).decode() if _lib is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


class ArenaDesc(C.Structure):
    _fields_ = [("max_blocks", C.c_int), ("num_mutexes", C.c_int), ("num_semaphores", C.c_int),
                ("num_groups", C.c_int), ("page_align", C.c_int), ("spin_limit", C.c_uint32)]


class DeviceInfo(C.Structure):
    _fields_ = [("num_cus", C.c_int), ("num_xcds", C.c_int), ("max_threads_per_block", C.c_int),
                ("lds_per_cu", C.c_int), ("warp_size", C.c_int), ("coop_launch", C.c_int),
                ("hbm_bytes", C.c_longlong), ("arch", C.c_char * 64)]


_P = C.c_void_p
_I = C.c_int
_U32 = C.c_uint32
_U64 = C.c_uint64
_PU32 = C.POINTER(C.c_uint32)
_PU64 = C.POINTER(C.c_uint64)
_PI = C.POINTER(C.c_int)

# name -> (restype, argtypes); every symbol declared in include/hs.h
SIGNATURES = {
    "hs_arena_create": (_I, [C.POINTER(ArenaDesc), C.POINTER(_P)]),
    "hs_arena_destroy": (_I, [_P]),
    "hs_arena_reset": (_I, [_P, _P]),
    "hs_arena_set_abort": (_I, [_P, _I]),
    "hs_wait_value_u32": (_I, [_P, _P, _U32, _P]),
    "hs_arena_device_status": (_I, [_P, _PI]),
    "hs_arena_sync": (_I, [_P, _P]),
    "hs_arena_episode_slots": (_I, [_P, _PU32, _I]),
    "hs_arena_device_view": (_I, [_P, _P, C.c_size_t]),
    "hs_mutex_create": (_I, [_P, _I, C.c_uint, C.POINTER(C.c_uint)]),
    "hs_sem_create": (_I, [_P, _I, C.c_uint, C.c_uint, C.POINTER(C.c_uint)]),
    "hsLocksInit": (_I, [_I, _I, _I, _I, _I]),
    "hsLocksDestroy": (_I, []),
    "hsMutexCreateSpin": (_I, [C.POINTER(C.c_uint), _I]),
    "hsMutexCreateEBO": (_I, [C.POINTER(C.c_uint), _I]),
    "hsMutexCreateFA": (_I, [C.POINTER(C.c_uint), _I]),
    "hsMutexCreateSleep": (_I, [C.POINTER(C.c_uint), _I]),
    "hsSemaphoreCreateSpin": (_I, [C.POINTER(C.c_uint), _I, C.c_uint, _I]),
    "hsSemaphoreCreateEBO": (_I, [C.POINTER(C.c_uint), _I, C.c_uint, _I]),
    "hsLocksArena": (_P, []),
    "hs_device_info_get": (_I, [C.POINTER(DeviceInfo)]),
    "hs_max_resident_wgs": (_I, [_I, _I, _PI]),
    "hs_probe_xcc_ids": (_I, [_I, _I, _PI]),
    "hs_set_xcd_groups": (_I, [_I]),
    "hs_barrier_fanin": (_I, [_P, _I, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "hs_persistent_launch": (_I, [_P, _P, _I, _I, _I, C.c_size_t, _P, _P]),
    "hs_sync_bench": (_I, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _P]),
    "hs_barrier_bench": (_I, [_P, _I, _P, _I, _I, _I, _I, _I, _P, _P]),
    "hs_sem_bench": (_I, [_P, _I, C.c_uint, _P, _I, _I, _I, _I, _I, _P, _P]),
    "hs_reduce_i32": (_I, [_P, _P, _U64, _P, _P, _I, _I, _P]),
    "hs_reduce_i32_stamped": (_I, [_P, _P, _U64, _P, _P, _I, _I, _P, _P]),
    "hs_reduce_partials_size": (_I, [_I, _U64, _PU64]),
    "hs_scan_u32": (_I, [_P, _P, _P, _U64, _I, _P, _I, _I, _P]),
    "hs_scan_scratch_size": (_I, [_I, _U64, _PU64]),
    "hs_scan_u32_noalloc": (_I, [_P, _P, _P, _U64, _I, _I, _I, _P]),
    "hs_scan_set_round_bytes": (_I, [_U64]),
    "hs_scan_get_round_bytes": (_I, [_PU64]),
    "hs_graph_capture_begin": (_I, [_P]),
    "hs_graph_capture_end": (_I, [_P, C.POINTER(_P)]),
    "hs_graph_replay": (_I, [_P, _P]),
    "hs_graph_destroy": (_I, [_P]),
    "hs_tune_reduce_i32": (_I, [_P, _P, _U64, _I, _P, _P]),
    "hs_tune_barrier": (_I, [_P, _P, _I, _I, _I, _I, _I, _P, _P]),
    "hs_tune_scan_u32": (_I, [_P, _P, _P, _U64, _I, _I, _P, _P]),
    "hs_tune_scan_batch_u32": (_I, [_P, _P, _P, _P, _U32, _I, _I, _P, _P]),
    "hs_tune_bfs": (_I, [_P, _U32, _U64, _P, _P, C.c_int32, _P, _I, _P, _P]),
    "hs_tune_msbfs": (_I, [_P, _U32, _U64, _P, _P, _P, _P, _P, _I, _P, _I, _P, _I, _P, _P]),
    "hs_tune_mssp": (_I, [_P, _U32, _U64, _P, _P, _P, _P, _I, C.c_int32, _P, _I, _P, _P]),
    "hs_tune_pagerank_f32": (_I, [_P, _U32, _U64, _P, _P, _P, C.c_float, C.c_float, _I, _P, _I, _P, _P]),
    "hs_batch_work_size": (_I, [_P, _U32, _PU64]),
    "hs_scan_batch_u32": (_I, [_P, _P, _P, _P, _U32, _I, _P, _I, _P]),
    "hs_reduce_batch_i32": (_I, [_P, _P, _P, _P, _U32, _P, _I, _P]),
    "hs_batch_status": (_I, [_P, _P]),
    "hs_serve_start": (_I, [_I, _U32, _U32, C.POINTER(_P)]),
    "hs_serve_submit": (_I, [_P, _U32, _U32, _U32, _U64, _U64, _U64, _U64, _U64, _U64, _PU32]),
    "hs_serve_submit_pagerank": (_I, [_P, _U32, _U32, C.c_float, _U64, _U64, _U64, _U64, _U64, _U64, _PU32]),
    "hs_serve_poll": (_I, [_P, _U32, _PU32, _PU64, _PU64]),
    "hs_serve_phases": (_I, [_P, _U32, _PU32]),
    "hs_serve_stop": (_I, [_P, _I]),
    "hs_mlp_param_count": (_I, []),
    "hs_mlp_scratch_floats": (_I, []),
    "hs_mlp_init": (_I, [_P, _U64, _P]),
    "hs_mlp_forward": (_I, [_P, _P, _I, _P, _P]),
    "hs_mlp_grad": (_I, [_P, _P, _P, _P, _I, C.c_float, _P, _P, _P, _P]),
    "hs_mlp_train_step": (_I, [_P, _P, _P, _P, _I, C.c_float, C.c_float, _I, _P, _P, _P]),
    "hs_mlp_actions": (_I, []),
    "hs_mlp_grad_cost": (_I, [_P, _P, _P, _I, C.c_float, _P, _P, _P, _P]),
    "hs_mlp_train_step_cost": (_I, [_P, _P, _P, _I, C.c_float, C.c_float, _I, _P, _P, _P]),
    "hs_gen_uniform_u32": (_I, [_P, _U64, _U64, _U32, _U32, _P]),
    "hs_gen_uniform_u32_host": (_I, [_P, _U64, _U64, _U32, _U32]),
    "hs_rmat_csr_host": (_I, [_I, _I, C.c_double, C.c_double, C.c_double, _U64, _P, _P, _PU64]),
    "hs_gen_weights_host": (_I, [_P, _U64, _U64, _U32, _U32]),
    "hs_gr_read_header": (_I, [C.c_char_p, _PU64, _PU64, _PU64]),
    "hs_gr_read": (_I, [C.c_char_p, _P, _P, _P]),
    "hs_gr_write": (_I, [C.c_char_p, _U32, _U64, _P, _P, _P]),
    "hs_csr_transpose_host": (_I, [_U32, _U64, _P, _P, _P, _P]),
    "hs_status_string": (C.c_char_p, [_I]),
    "hs_version": (C.c_char_p, []),
}
# graph entry points (include/hs_graph.h)
GRAPH_SIGNATURES = {
    "hs_bfs_work_size": (_I, [_U32, _U64, _PU64]),
    "hs_bfs": (_I, [_P, _U32, _U64, _P, _P, C.c_int32, _P, _P, _I, _I, _P, _P]),
    "hs_bfs_lb": (_I, [_P, _U32, _U64, _P, _P, C.c_int32, _P, _P, _I, _I, _I, _P, _P]),
    "hs_msbfs_work_size": (_I, [_U32, _U64, _I, _PU64]),
    "hs_msbfs": (_I, [_P, _U32, _U64, _P, _P, _P, _P, _P, _I, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "hs_sssp_work_size": (_I, [_U32, _U64, _PU64]),
    "hs_sssp": (_I, [_P, _U32, _U64, _P, _P, _P, C.c_int32, C.c_int32, _P, _P, _I, _I, _P, _P]),
    "hs_mssp_work_size": (_I, [_U32, _U64, _I, _PU64]),
    "hs_mssp": (_I, [_P, _U32, _U64, _P, _P, _P, _P, _I, C.c_int32, _P, _P, _I, _I, _P, _P]),
    "hs_pagerank_work_size": (_I, [_U32, _U64, _PU64]),
    "hs_pagerank_f32": (_I, [_P, _U32, _U64, _P, _P, _P, C.c_float, C.c_float, _I, _P, _P, _I, _I,
                             _P, _P]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load libhs.so and bind every entry point; raise ImportError if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"libhs.so not built at {path}: run `make -C tpds22-artifact_amd` "
                          "or __graft_entry__.build(); there is no fallback path")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 (SONAME
    # libamdhip64.so.7) and loads it by file name.  Importing torch first makes
    # libhs's dependency on libamdhip64.so.7 resolve to that same runtime;
    # loading libhs first would map a second runtime that then sees no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    _check_single_hip_runtime()
    for name, (res, args) in {**SIGNATURES, **GRAPH_SIGNATURES}.items():
        fn = getattr(lib, name, None)
        if fn is None:
            if name in GRAPH_SIGNATURES:
                continue
            raise ImportError(f"libhs.so lacks symbol {name}")
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check_single_hip_runtime() -> None:
    try:
        maps = open("/proc/self/maps").read()
    except OSError:
        return
    paths = {line.split()[-1] for line in maps.splitlines()
             if "libamdhip64" in line and line.split()[-1].startswith("/")}
    real = {os.path.realpath(p) for p in paths}
    if len(real) > 1:
        raise ImportError("two HIP runtimes are mapped in this process (%s); import torch "
                          "before hsync" % ", ".join(sorted(real)))


def check(status: int, what: str = "") -> None:
    if status != HS_OK:
        raise HSError(status, what)
