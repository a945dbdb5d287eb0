"""ctypes binding of libbrigjoin (include/bj_api.h).

Thin plumbing used by the tests, bench.py and __graft_entry__: it loads the
in-tree HIP library and exposes the C ABI with numpy / torch buffers.  All the
work happens in the library; there is no Python or CPU compute path here, and
loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbrigjoin.so")
DRIVER_PATH = os.path.join(HERE, "bin", "bj_driver")

BJ_OK = 0
ERRORS = {-1: "BJ_ERR_INVALID", -2: "BJ_ERR_HIP", -3: "BJ_ERR_NOMEM", -4: "BJ_ERR_IO",
          -5: "BJ_ERR_SIZE", -6: "BJ_ERR_NODEVICE", -7: "BJ_ERR_INTERNAL"}
PHASES = ["h2d", "r_partition", "bf_build", "s_filter", "s_partition", "probe", "total"]
KERNELS = ["hist_r", "scan_r", "scatter1_r", "scatter2_r", "bf_build", "bf_filter", "hist_s",
           "scan_s", "scatter1_s", "scatter2_s", "plan", "probe"]

# Every function the header declares (checked by the CPU tests).
EXPORTS = [
    "bj_create", "bj_destroy", "bj_set_stream", "bj_params_default", "bj_strerror",
    "bj_last_error", "bj_version", "bj_join", "bj_copy_rows_to_host", "bj_gen_build",
    "bj_gen_probe", "bj_write_output", "bj_read_output", "bj_free_host",
    "bj_cpu_sanity_count", "bj_main_args_default", "bj_gpu_main", "bj_plan", "bj_partition",
]


class BjPlanInfo(ctypes.Structure):
    _fields_ = [
        ("bits1", ctypes.c_int32), ("bits2", ctypes.c_int32), ("bf", ctypes.c_int32),
        ("bf_seg_bits", ctypes.c_int32), ("bf_k", ctypes.c_int32), ("bf_words", ctypes.c_uint64),
        ("tile", ctypes.c_int32), ("capacity", ctypes.c_uint64), ("workspace_bytes", ctypes.c_uint64),
        ("bf_stage", ctypes.c_int32),
    ]


class BjParams(ctypes.Structure):
    _fields_ = [
        ("r_size", ctypes.c_int64), ("s_size", ctypes.c_int64),
        ("partition_size", ctypes.c_int32), ("second_partition_size", ctypes.c_int32),
        ("radix_bits", ctypes.c_int32), ("use_bf", ctypes.c_int32),
        ("bf_bits_per_key", ctypes.c_double), ("true_false_ratio", ctypes.c_double),
        ("num_bits", ctypes.c_int64), ("num_hashes", ctypes.c_int32),
        ("answer_scale", ctypes.c_double), ("tile", ctypes.c_int32),
        ("inputs_on_device", ctypes.c_int32), ("collect_phase_times", ctypes.c_int32),
        ("s_chunks", ctypes.c_int32), ("bf_stage", ctypes.c_int32),
        ("pass1_bits", ctypes.c_int32),
    ]


class BjResult(ctypes.Structure):
    _fields_ = [
        ("match_count", ctypes.c_uint64), ("rows_written", ctypes.c_uint64),
        ("capacity", ctypes.c_uint64), ("truncated", ctypes.c_int32),
        ("s_survivors", ctypes.c_uint64),
        ("key", ctypes.c_void_p), ("r_val", ctypes.c_void_p), ("s_val", ctypes.c_void_p),
        ("phase_ms", ctypes.c_double * 7),
        ("kernel_ms", ctypes.c_double * 12),
        ("bits1", ctypes.c_int32), ("bits2", ctypes.c_int32), ("bf_seg_bits", ctypes.c_int32),
        ("bf_k", ctypes.c_int32), ("bf_words", ctypes.c_uint64), ("tile", ctypes.c_int32),
        ("heavy_items", ctypes.c_uint64), ("bf_stage", ctypes.c_int32),
        ("pass1_bits", ctypes.c_int32),
    ]

    def phases(self) -> dict:
        return {n: float(self.phase_ms[i]) for i, n in enumerate(PHASES)}

    def kernels(self) -> dict:
        return {n: float(self.kernel_ms[i]) for i, n in enumerate(KERNELS)}


class BjMainArgs(ctypes.Structure):
    _fields_ = [
        ("r_path", ctypes.c_char_p), ("s_path", ctypes.c_char_p), ("out_path", ctypes.c_char_p),
        ("r_size", ctypes.c_int64), ("s_size", ctypes.c_int64),
        ("sanity_check", ctypes.c_int32), ("chunk_size", ctypes.c_int32),
        ("use_bf", ctypes.c_int32), ("use_nvlink", ctypes.c_int32),
        ("use_warp_compact", ctypes.c_int32), ("true_false_ratio", ctypes.c_double),
        ("s_chunks", ctypes.c_int32), ("kernel_fusion", ctypes.c_int32),
        ("redist_skew_threshold", ctypes.c_double), ("override_num_bits", ctypes.c_int64),
        ("override_num_hashes", ctypes.c_int32), ("override_partition_rs_split", ctypes.c_int32),
        ("override_partition_size", ctypes.c_int32),
        ("override_second_partition_size", ctypes.c_int32),
        ("fused_probe_breakdown", ctypes.c_int32), ("read_val", ctypes.c_int32),
        ("r_shuffle_enable", ctypes.c_int32), ("r_replicate_ratio", ctypes.c_double),
        ("answer_scale", ctypes.c_double), ("radix_bits", ctypes.c_int32),
        ("bf_bits_per_key", ctypes.c_double), ("val_mode", ctypes.c_int32),
        ("device", ctypes.c_int32), ("verbose", ctypes.c_int32),
    ]


_lib = None
_u32p = ctypes.POINTER(ctypes.c_uint32)


class BjError(RuntimeError):
    def __init__(self, code: int, where: str):
        self.code = code
        detail = lib().bj_last_error().decode(errors="replace")
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({detail})")


def lib():
    """Load the in-tree HIP library; raise if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} not built: run __graft_entry__.build() or `make -C brig-join_amd`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.bj_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
        L.bj_destroy.argtypes = [vp]
        L.bj_set_stream.argtypes = [vp, vp]
        L.bj_params_default.argtypes = [ctypes.POINTER(BjParams)]
        L.bj_params_default.restype = None
        L.bj_strerror.restype = ctypes.c_char_p
        L.bj_strerror.argtypes = [ctypes.c_int]
        L.bj_last_error.restype = ctypes.c_char_p
        L.bj_last_error.argtypes = []
        L.bj_join.argtypes = [vp, ctypes.POINTER(BjParams), vp, vp, vp, vp,
                              ctypes.POINTER(BjResult)]
        L.bj_copy_rows_to_host.argtypes = [vp, ctypes.POINTER(BjResult), vp, vp, vp,
                                           ctypes.c_uint64]
        L.bj_gen_build.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, vp, vp]
        L.bj_gen_probe.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, vp, vp]
        L.bj_write_output.argtypes = [ctypes.c_char_p, ctypes.c_uint64, vp, vp, vp]
        L.bj_read_output.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(_u32p), ctypes.POINTER(_u32p),
                                     ctypes.POINTER(_u32p)]
        L.bj_free_host.argtypes = [vp]
        L.bj_free_host.restype = None
        L.bj_cpu_sanity_count.restype = ctypes.c_longlong
        L.bj_cpu_sanity_count.argtypes = [vp, ctypes.c_int64, vp, ctypes.c_int64]
        L.bj_main_args_default.argtypes = [ctypes.POINTER(BjMainArgs)]
        L.bj_main_args_default.restype = None
        L.bj_gpu_main.argtypes = [ctypes.POINTER(BjMainArgs), ctypes.POINTER(BjResult)]
        L.bj_plan.argtypes = [ctypes.POINTER(BjParams), ctypes.POINTER(BjPlanInfo)]
        L.bj_partition.argtypes = [vp, ctypes.POINTER(BjParams), vp, vp, ctypes.c_int64, vp, vp, vp]
        _lib = L
    return _lib


def check(code: int, where: str) -> None:
    if code != BJ_OK:
        raise BjError(code, where)


def default_params(**kw) -> BjParams:
    p = BjParams()
    lib().bj_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def plan(params: BjParams) -> BjPlanInfo:
    info = BjPlanInfo()
    check(lib().bj_plan(ctypes.byref(params), ctypes.byref(info)), "bj_plan")
    return info


def _ptr(x):
    """Device/host address of a torch tensor or numpy array (or an int)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        assert x.dtype == np.uint32 and x.flags["C_CONTIGUOUS"]
        return x.ctypes.data
    # torch tensor
    assert x.is_contiguous()
    return x.data_ptr()


class Context:
    """One library context (device memory, stream, events) on one GPU."""

    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        check(lib().bj_create(ctypes.byref(self.h), device), "bj_create")

    def close(self):
        if self.h:
            lib().bj_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle: int | None):
        check(lib().bj_set_stream(self.h, stream_handle), "bj_set_stream")

    def join(self, params: BjParams, r_key, r_val, s_key, s_val) -> BjResult:
        res = BjResult()
        rc = lib().bj_join(self.h, ctypes.byref(params), _ptr(r_key), _ptr(r_val),
                           _ptr(s_key), _ptr(s_val), ctypes.byref(res))
        check(rc, "bj_join")
        return res

    def partition(self, params: BjParams, key, val, n: int, out_key, out_val, out_start) -> None:
        check(lib().bj_partition(self.h, ctypes.byref(params), _ptr(key), _ptr(val), n,
                                 _ptr(out_key), _ptr(out_val), _ptr(out_start)), "bj_partition")

    def rows_to_host(self, res: BjResult, n: int | None = None):
        n = res.rows_written if n is None else n
        k = np.empty(max(n, 1), np.uint32)
        a = np.empty(max(n, 1), np.uint32)
        b = np.empty(max(n, 1), np.uint32)
        check(lib().bj_copy_rows_to_host(self.h, ctypes.byref(res), k.ctypes.data, a.ctypes.data,
                                         b.ctypes.data, n), "bj_copy_rows_to_host")
        return k[:n], a[:n], b[:n]

    def gen_build(self, seed: int, n_r: int, d_key, d_val) -> None:
        check(lib().bj_gen_build(self.h, seed, n_r, _ptr(d_key), _ptr(d_val)), "bj_gen_build")

    def gen_probe(self, seed_r: int, seed_s: int, n_r: int, n_s: int, m_s: int, d_key, d_val,
                  zipf_s: float = 0.0) -> None:
        check(lib().bj_gen_probe(self.h, seed_r, seed_s, n_r, n_s, m_s, zipf_s, _ptr(d_key),
                                 _ptr(d_val)), "bj_gen_probe")


def write_output(path: str, key, r_val, s_val) -> None:
    key = np.ascontiguousarray(key, np.uint32).copy()
    r_val = np.ascontiguousarray(r_val, np.uint32).copy()
    s_val = np.ascontiguousarray(s_val, np.uint32).copy()
    check(lib().bj_write_output(path.encode(), len(key), key.ctypes.data, r_val.ctypes.data,
                                s_val.ctypes.data), "bj_write_output")


def read_output(path: str):
    n = ctypes.c_uint64()
    a, b, c = _u32p(), _u32p(), _u32p()
    check(lib().bj_read_output(path.encode(), ctypes.byref(n), ctypes.byref(a), ctypes.byref(b),
                               ctypes.byref(c)), "bj_read_output")
    try:
        cnt = n.value
        out = [np.ctypeslib.as_array(p, shape=(max(cnt, 1),))[:cnt].copy() for p in (a, b, c)]
    finally:
        for p in (a, b, c):
            lib().bj_free_host(ctypes.cast(p, ctypes.c_void_p))
    return tuple(out)


def cpu_sanity_count(r_key: np.ndarray, s_key: np.ndarray) -> int:
    r = np.ascontiguousarray(r_key, np.uint32)
    s = np.ascontiguousarray(s_key, np.uint32)
    return int(lib().bj_cpu_sanity_count(r.ctypes.data, len(r), s.ctypes.data, len(s)))


def main_args(**kw) -> BjMainArgs:
    a = BjMainArgs()
    lib().bj_main_args_default(ctypes.byref(a))
    for k, v in kw.items():
        if isinstance(v, str):
            v = v.encode()
        setattr(a, k, v)
    return a


def gpu_main(args: BjMainArgs) -> BjResult:
    res = BjResult()
    check(lib().bj_gpu_main(ctypes.byref(args), ctypes.byref(res)), "bj_gpu_main")
    return res
