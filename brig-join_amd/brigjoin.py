"""ctypes binding of libbrigjoin (include/bj_api.h).

Thin plumbing used by the tests, bench.py and __graft_entry__: it loads the
in-tree HIP library and exposes the C ABI with numpy / torch buffers.  All the
work happens in the library; there is no Python or CPU compute path here, and
loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbrigjoin.so")
DRIVER_PATH = os.path.join(HERE, "bin", "bj_driver")
GENDATA_PATH = os.path.join(HERE, "bin", "bj_gendata")

BJ_OK = 0
ERRORS = {-1: "BJ_ERR_INVALID", -2: "BJ_ERR_HIP", -3: "BJ_ERR_NOMEM", -4: "BJ_ERR_IO",
          -5: "BJ_ERR_SIZE", -6: "BJ_ERR_NODEVICE", -7: "BJ_ERR_INTERNAL"}
PHASES = ["h2d", "r_partition", "bf_build", "s_filter", "s_partition", "probe", "total"]
PROBE_PHASES = ["items", "clear", "build", "count", "write"]
KERNELS = ["hist_r", "scan_r", "scatter1_r", "scatter2_r", "bf_build", "bf_filter", "hist_s",
           "scan_s", "scatter1_s", "scatter2_s", "plan", "probe"]

# Every function the header declares (checked by the CPU tests).
EXPORTS = [
    "bj_create", "bj_destroy", "bj_set_stream", "bj_params_default", "bj_strerror",
    "bj_last_error", "bj_version", "bj_join", "bj_copy_rows_to_host", "bj_gen_build",
    "bj_gen_probe", "bj_write_output", "bj_read_output", "bj_free_host",
    "bj_cpu_sanity_count", "bj_main_args_default", "bj_gpu_main", "bj_plan", "bj_partition",
    "bj_join_spec_default", "bj_batch_plan", "bj_batch_create", "bj_batch_run", "bj_batch_destroy",
    "bj_autotune", "bj_tune_get", "bj_tune_set",
    "bj_server_opts_default", "bj_server_create", "bj_server_add_relation", "bj_server_request_bytes",
    "bj_server_submit", "bj_server_poll", "bj_server_wait", "bj_server_copy_rows", "bj_server_release",
    "bj_server_drain",
    "bj_server_stats_get", "bj_server_destroy", "bj_gen_refdist_host", "bj_gen_refdist",
    "bj_join_noalloc",
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
        ("pass1_bits", ctypes.c_int32), ("s_pass1_bits", ctypes.c_int32),
        ("scatter_wg", ctypes.c_int32), ("probe_breakdown", ctypes.c_int32),
        ("chunk_probe", ctypes.c_int32),
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
        ("pass1_bits", ctypes.c_int32), ("s_pass1_bits", ctypes.c_int32),
        ("s_chunks", ctypes.c_int32),
        ("probe_phase_ms", ctypes.c_double * 5),
    ]

    def phases(self) -> dict:
        return {n: float(self.phase_ms[i]) for i, n in enumerate(PHASES)}

    def probe_phases(self) -> dict:
        return {n: float(self.probe_phase_ms[i]) for i, n in enumerate(PROBE_PHASES)}

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


class BjJoinSpec(ctypes.Structure):
    _fields_ = [
        ("r_key", ctypes.c_void_p), ("r_val", ctypes.c_void_p),
        ("s_key", ctypes.c_void_p), ("s_val", ctypes.c_void_p),
        ("r_size", ctypes.c_int64), ("s_size", ctypes.c_int64),
        ("radix_bits", ctypes.c_int32), ("tile", ctypes.c_int32),
        ("bf_bits_per_key", ctypes.c_double), ("num_hashes", ctypes.c_int32),
        ("answer_scale", ctypes.c_double), ("match_hint", ctypes.c_double),
    ]


class BjJoinOut(ctypes.Structure):
    _fields_ = [
        ("match_count", ctypes.c_uint64), ("rows_written", ctypes.c_uint64),
        ("capacity", ctypes.c_uint64), ("s_survivors", ctypes.c_uint64),
        ("truncated", ctypes.c_int32), ("radix_bits", ctypes.c_int32),
        ("tile", ctypes.c_int32), ("bf_k", ctypes.c_int32),
        ("bf_stage", ctypes.c_int32), ("bf_pass_bits", ctypes.c_int32),
        ("bf_bits_per_key", ctypes.c_double),
        ("key", ctypes.c_void_p), ("r_val", ctypes.c_void_p), ("s_val", ctypes.c_void_p),
    ]


class BjTuneEntry(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("lg", ctypes.c_int32), ("sr", ctypes.c_int32),
        ("sel", ctypes.c_int32), ("radix_bits", ctypes.c_int32), ("tile", ctypes.c_int32),
        ("bf_bits_per_key", ctypes.c_double), ("ms", ctypes.c_double),
    ]


class BjServerOpts(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_int32), ("arena_bytes", ctypes.c_uint64), ("use_bf", ctypes.c_int32),
        ("bf_bits_per_key", ctypes.c_double), ("num_hashes", ctypes.c_int32),
        ("radix_bits", ctypes.c_int32), ("answer_scale", ctypes.c_double),
    ]


class BjRequestResult(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("relation", ctypes.c_int32),
        ("s_size", ctypes.c_uint64), ("match_count", ctypes.c_uint64),
        ("rows_written", ctypes.c_uint64), ("capacity", ctypes.c_uint64),
        ("s_survivors", ctypes.c_uint64), ("truncated", ctypes.c_int32),
        ("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_end", ctypes.c_double),
        ("device_ms", ctypes.c_double),
        ("key", ctypes.c_void_p), ("r_val", ctypes.c_void_p), ("s_val", ctypes.c_void_p),
    ]


class BjServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("released", ctypes.c_uint64),
        ("max_in_flight", ctypes.c_int32), ("arena_bytes", ctypes.c_uint64),
        ("arena_peak_bytes", ctypes.c_uint64), ("relation_bytes", ctypes.c_uint64),
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
        L.bj_join_spec_default.argtypes = [ctypes.POINTER(BjJoinSpec)]
        L.bj_join_spec_default.restype = None
        L.bj_batch_plan.argtypes = [vp, ctypes.c_int, ctypes.POINTER(BjJoinSpec),
                                    ctypes.POINTER(BjJoinOut), ctypes.POINTER(ctypes.c_uint64)]
        L.bj_batch_create.argtypes = [vp, ctypes.c_int, ctypes.POINTER(BjJoinSpec), ctypes.c_int,
                                      ctypes.POINTER(vp)]
        L.bj_batch_run.argtypes = [vp, ctypes.POINTER(BjJoinOut), ctypes.POINTER(ctypes.c_double)]
        L.bj_batch_destroy.argtypes = [vp]
        L.bj_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(BjJoinSpec),
                                  ctypes.POINTER(ctypes.c_int)]
        L.bj_tune_get.argtypes = [vp, ctypes.POINTER(BjTuneEntry), ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int)]
        L.bj_tune_set.argtypes = [vp, ctypes.POINTER(BjTuneEntry), ctypes.c_int]
        L.bj_server_opts_default.argtypes = [ctypes.POINTER(BjServerOpts)]
        L.bj_server_opts_default.restype = None
        L.bj_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(BjServerOpts), ctypes.POINTER(vp)]
        L.bj_server_add_relation.argtypes = [vp, vp, vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]
        L.bj_server_request_bytes.argtypes = [vp, ctypes.c_int32, ctypes.c_int64,
                                              ctypes.POINTER(ctypes.c_uint64)]
        L.bj_server_submit.argtypes = [vp, ctypes.c_int32, vp, vp, ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_uint64)]
        L.bj_server_poll.argtypes = [vp, ctypes.POINTER(BjRequestResult), ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int)]
        L.bj_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(BjRequestResult)]
        L.bj_server_copy_rows.argtypes = [vp, ctypes.POINTER(BjRequestResult), vp, vp, vp, ctypes.c_uint64]
        L.bj_server_release.argtypes = [vp, ctypes.c_uint64]
        L.bj_server_drain.argtypes = [vp]
        L.bj_server_stats_get.argtypes = [vp, ctypes.POINTER(BjServerStats)]
        L.bj_server_destroy.argtypes = [vp]
        L.bj_join_noalloc.argtypes = [vp, ctypes.POINTER(BjParams), vp, vp, vp, vp, vp, vp, vp,
                                      ctypes.c_uint64, ctypes.POINTER(BjResult)]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.bj_gen_refdist_host.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64,
                                          ctypes.c_double, vp, vp, i64p]
        L.bj_gen_refdist.argtypes = [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64,
                                     ctypes.c_double, vp, vp, i64p]
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

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter shutdown the HIP runtime may already be going away:
        # leave the memory to the end of the process instead of freeing it
        # (the check is bound at definition: module globals such as sys may
        # already be cleared when this runs)
        if _finalizing():
            return
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

    def join_noalloc(self, params: BjParams, r_key, r_val, s_key, s_val, out_key, out_rval, out_sval,
                     out_cap: int) -> BjResult:
        """bj_join_noalloc: device inputs, rows into the caller's device arrays."""
        res = BjResult()
        rc = lib().bj_join_noalloc(self.h, ctypes.byref(params), _ptr(r_key), _ptr(r_val), _ptr(s_key),
                                   _ptr(s_val), _ptr(out_key), _ptr(out_rval), _ptr(out_sval), out_cap,
                                   ctypes.byref(res))
        check(rc, "bj_join_noalloc")
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

    def gen_refdist(self, r_size: int, s_size: int, answer_pct: float, d_r_key, d_s_key, seed: int = 42,
                    max_pct: float = -1.0) -> int:
        """The reference datagen's key distribution on the device; returns the match count."""
        k = ctypes.c_int64()
        check(lib().bj_gen_refdist(self.h, r_size, s_size, answer_pct, seed, max_pct, _ptr(d_r_key),
                                   _ptr(d_s_key), ctypes.byref(k)), "bj_gen_refdist")
        return int(k.value)


def gen_refdist_host(r_size: int, s_size: int, answer_pct: float, seed: int = 42, max_pct: float = -1.0):
    """The reference datagen's keys in its exact order (host); returns (r_key, s_key, matches)."""
    rk = np.empty(max(r_size, 1), np.uint32)
    sk = np.empty(max(s_size, 1), np.uint32)
    k = ctypes.c_int64()
    check(lib().bj_gen_refdist_host(r_size, s_size, answer_pct, seed, max_pct, rk.ctypes.data, sk.ctypes.data,
                                    ctypes.byref(k)), "bj_gen_refdist_host")
    return rk[:r_size], sk[:s_size], int(k.value)


def join_spec(r_key=None, r_val=None, s_key=None, s_val=None, r_size=0, s_size=0, **kw) -> BjJoinSpec:
    s = BjJoinSpec()
    lib().bj_join_spec_default(ctypes.byref(s))
    s.r_key, s.r_val, s.s_key, s.s_val = (_ptr(x) for x in (r_key, r_val, s_key, s_val))
    s.r_size, s.s_size = r_size, s_size
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    return s


def _spec_array(specs):
    arr = (BjJoinSpec * max(len(specs), 1))()
    for i, s in enumerate(specs):
        arr[i] = s
    return arr


def batch_plan(specs, ctx: "Context | None" = None):
    """Configuration and device footprint of a batch (no device needed without ctx)."""
    arr = _spec_array(specs)
    outs = (BjJoinOut * max(len(specs), 1))()
    ws = ctypes.c_uint64()
    check(lib().bj_batch_plan(ctx.h if ctx else None, len(specs), arr, outs, ctypes.byref(ws)),
          "bj_batch_plan")
    return [outs[i] for i in range(len(specs))], int(ws.value)


class Batch:
    """Many joins in one launch sequence (captured in a hipGraph when use_graph)."""

    def __init__(self, ctx: Context, specs, use_graph: bool = True):
        self.ctx = ctx
        self.n = len(specs)
        self._specs = _spec_array(specs)
        self.h = ctypes.c_void_p()
        check(lib().bj_batch_create(ctx.h, self.n, self._specs, 1 if use_graph else 0,
                                    ctypes.byref(self.h)), "bj_batch_create")

    def run(self):
        outs = (BjJoinOut * max(self.n, 1))()
        ms = ctypes.c_double()
        check(lib().bj_batch_run(self.h, outs, ctypes.byref(ms)), "bj_batch_run")
        return [outs[i] for i in range(self.n)], float(ms.value)

    def rows_to_host(self, out: BjJoinOut):
        """Rows of one join's result (copied through a BjResult view)."""
        r = BjResult()
        r.rows_written = out.rows_written
        r.key, r.r_val, r.s_val = out.key, out.r_val, out.s_val
        return self.ctx.rows_to_host(r)

    def close(self):
        if self.h:
            lib().bj_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter shutdown the HIP runtime may already be going away:
        # leave the memory to the end of the process instead of freeing it
        # (the check is bound at definition: module globals such as sys may
        # already be cleared when this runs)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def autotune(ctx: Context, specs) -> int:
    arr = _spec_array(specs)
    n = ctypes.c_int()
    check(lib().bj_autotune(ctx.h, len(specs), arr, ctypes.byref(n)), "bj_autotune")
    return n.value


def tune_get(ctx: Context):
    n = ctypes.c_int()
    check(lib().bj_tune_get(ctx.h, None, 0, ctypes.byref(n)), "bj_tune_get")
    arr = (BjTuneEntry * max(n.value, 1))()
    check(lib().bj_tune_get(ctx.h, arr, n.value, ctypes.byref(n)), "bj_tune_get")
    return [arr[i] for i in range(n.value)]


def tune_set(ctx: Context, entries) -> None:
    arr = (BjTuneEntry * max(len(entries), 1))()
    for i, e in enumerate(entries):
        arr[i] = e
    check(lib().bj_tune_set(ctx.h, arr, len(entries)), "bj_tune_set")


def server_opts(**kw) -> BjServerOpts:
    o = BjServerOpts()
    lib().bj_server_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Server:
    """Request loop over device-resident relations (continuous batching)."""

    def __init__(self, device: int = 0, **opts):
        self.h = ctypes.c_void_p()
        o = server_opts(**opts)
        check(lib().bj_server_create(device, ctypes.byref(o), ctypes.byref(self.h)), "bj_server_create")
        self._ctx = None

    def add_relation(self, r_key, r_val, n_r: int) -> int:
        rid = ctypes.c_int32()
        check(lib().bj_server_add_relation(self.h, _ptr(r_key), _ptr(r_val), n_r, ctypes.byref(rid)),
              "bj_server_add_relation")
        return rid.value

    def request_bytes(self, rel: int, n_s: int) -> int:
        b = ctypes.c_uint64()
        check(lib().bj_server_request_bytes(self.h, rel, n_s, ctypes.byref(b)), "bj_server_request_bytes")
        return b.value

    def submit(self, rel: int, s_key, s_val, n_s: int) -> int:
        rid = ctypes.c_uint64()
        check(lib().bj_server_submit(self.h, rel, _ptr(s_key), _ptr(s_val), n_s, ctypes.byref(rid)),
              "bj_server_submit")
        return rid.value

    def poll(self, cap: int = 256):
        arr = (BjRequestResult * cap)()
        n = ctypes.c_int()
        check(lib().bj_server_poll(self.h, arr, cap, ctypes.byref(n)), "bj_server_poll")
        return [arr[i] for i in range(n.value)]

    def wait(self, rid: int) -> BjRequestResult:
        r = BjRequestResult()
        check(lib().bj_server_wait(self.h, rid, ctypes.byref(r)), "bj_server_wait")
        return r

    def release(self, rid: int) -> None:
        check(lib().bj_server_release(self.h, rid), "bj_server_release")

    def drain(self) -> None:
        check(lib().bj_server_drain(self.h), "bj_server_drain")

    def stats(self) -> BjServerStats:
        st = BjServerStats()
        check(lib().bj_server_stats_get(self.h, ctypes.byref(st)), "bj_server_stats_get")
        return st

    def rows_to_host(self, r: BjRequestResult):
        """Rows of a completed request, copied to the host (before release)."""
        n = int(r.rows_written)
        k, a, b = (np.empty(max(n, 1), np.uint32) for _ in range(3))
        check(lib().bj_server_copy_rows(self.h, ctypes.byref(r), k.ctypes.data, a.ctypes.data,
                                        b.ctypes.data, n), "bj_server_copy_rows")
        return k[:n], a[:n], b[:n]

    def close(self):
        if self.h:
            lib().bj_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter shutdown the HIP runtime may already be going away:
        # leave the memory to the end of the process instead of freeing it
        # (the check is bound at definition: module globals such as sys may
        # already be cleared when this runs)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


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
