"""CPU oracle for the radix hash join -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  The product (the HIP library and its C-ABI)
never imports, links or calls it.

Wraps two shared libraries with ctypes:
  oracle/lib/libbjoracle.so   the C restatement in bj_oracle.c (built by `make`)
  oracle/_ref/libref_sanity.so  the reference's own src/sanity_check.cpp compiled
                              from its source (built by `make ref`; optional)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbjoracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_sanity.so")
REF_GEN_PATH = os.path.join(HERE, "_ref", "gen_dataset")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_lib = None
_ref = None


def build() -> None:
    """Compile the oracle's C restatement (and oracle/_ref when the reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_lowbias32.restype = ctypes.c_uint32
        L.orc_lowbias32.argtypes = [ctypes.c_uint32]
        L.orc_perm_keys.argtypes = [ctypes.c_uint64, _u32p]
        L.orc_permw.restype = ctypes.c_uint32
        L.orc_permw.argtypes = [ctypes.c_uint32, ctypes.c_int, _u32p]
        L.orc_gen_r.argtypes = [ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p]
        L.orc_gen_s.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                ctypes.c_uint64, ctypes.c_uint64, _u32p, _u32p]
        L.orc_permw_defects.restype = ctypes.c_uint64
        L.orc_permw_defects.argtypes = [ctypes.c_int, ctypes.c_uint64]
        L.orc_sanity_count.restype = ctypes.c_longlong
        L.orc_sanity_count.argtypes = [_i32p, ctypes.c_longlong, _i32p, ctypes.c_longlong,
                                       ctypes.c_int]
        for name in ("orc_cpu_join",):
            f = getattr(L, name)
            f.restype = ctypes.c_longlong
            f.argtypes = [_u32p, _u32p, ctypes.c_longlong, _u32p, _u32p, ctypes.c_longlong,
                          _u32p, _u32p, _u32p, ctypes.c_longlong, ctypes.c_int]
        L.orc_nested_loop.restype = ctypes.c_longlong
        L.orc_nested_loop.argtypes = [_u32p, _u32p, ctypes.c_longlong, _u32p, _u32p,
                                      ctypes.c_longlong, _u32p, _u32p, _u32p,
                                      ctypes.c_longlong]
        L.orc_radix_hash.restype = ctypes.c_uint32
        L.orc_radix_hash.argtypes = [ctypes.c_uint32]
        _u64p = ctypes.POINTER(ctypes.c_uint64)
        L.orc_bloom_build.argtypes = [_u32p, ctypes.c_longlong, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.c_int, _u64p]
        L.orc_bloom_query.restype = ctypes.c_int
        L.orc_bloom_query.argtypes = [_u64p, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.c_int]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        R.ref_sanity_count.restype = ctypes.c_longlong
        R.ref_sanity_count.argtypes = [_i32p, ctypes.c_longlong, _i32p, ctypes.c_longlong,
                                       ctypes.POINTER(ctypes.c_double)]
        _ref = R
    return _ref


def _p32(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def _pi32(a: np.ndarray):
    a = a.view(np.int32)
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_i32p)


# ---------------------------------------------------------------- generator
def gen_r(seed: int, n_r: int):
    key = np.empty(n_r, np.uint32)
    val = np.empty(n_r, np.uint32)
    lib().orc_gen_r(seed, n_r, _p32(key), _p32(val))
    return key, val


def match_count(n_s: int, ratio: float) -> int:
    """Closed-form number of matching S tuples (= join size) of the generator."""
    return int(min(n_s, max(0, round(ratio * n_s))))


def gen_s(seed_r: int, seed_s: int, n_r: int, n_s: int, m_s: int):
    key = np.empty(n_s, np.uint32)
    val = np.empty(n_s, np.uint32)
    lib().orc_gen_s(seed_r, seed_s, n_r, n_s, m_s, _p32(key), _p32(val))
    return key, val


def permw_defects(w: int, seed: int) -> int:
    return int(lib().orc_permw_defects(w, seed))


# ---------------------------------------------------------------- joins
def sanity_count(r_key: np.ndarray, s_key: np.ndarray, threads: int = 0) -> int:
    """Faithful restatement of run_sanity_check: the count only."""
    r = np.ascontiguousarray(r_key, dtype=np.uint32)
    s = np.ascontiguousarray(s_key, dtype=np.uint32)
    return int(lib().orc_sanity_count(_pi32(r), len(r), _pi32(s), len(s), threads))


def ref_sanity_count(r_key: np.ndarray, s_key: np.ndarray):
    """The reference's own run_sanity_check (compiled into oracle/_ref): (count, seconds)."""
    r = np.ascontiguousarray(r_key, dtype=np.uint32)
    s = np.ascontiguousarray(s_key, dtype=np.uint32)
    sec = ctypes.c_double(0.0)
    c = ref().ref_sanity_count(_pi32(r), len(r), _pi32(s), len(s), ctypes.byref(sec))
    return int(c), float(sec.value)


def cpu_join(r_key, r_val, s_key, s_val, cap: int | None = None, threads: int = 0):
    """Row-emitting CPU join.  Returns (count, key, r_val, s_val) with min(count,cap) rows."""
    r_key = np.ascontiguousarray(r_key, dtype=np.uint32)
    r_val = np.ascontiguousarray(r_val, dtype=np.uint32)
    s_key = np.ascontiguousarray(s_key, dtype=np.uint32)
    s_val = np.ascontiguousarray(s_val, dtype=np.uint32)
    L = lib()
    if cap is None:
        # count first (cap=0 writes nothing), then allocate exactly
        z = np.empty(1, np.uint32)
        cap = L.orc_cpu_join(_p32(r_key), _p32(r_val), len(r_key), _p32(s_key), _p32(s_val),
                             len(s_key), _p32(z), _p32(z), _p32(z), 0, threads)
        if cap < 0:
            raise MemoryError("oracle cpu_join allocation failed")
    ok = np.empty(max(cap, 1), np.uint32)
    orv = np.empty(max(cap, 1), np.uint32)
    osv = np.empty(max(cap, 1), np.uint32)
    n = L.orc_cpu_join(_p32(r_key), _p32(r_val), len(r_key), _p32(s_key), _p32(s_val),
                       len(s_key), _p32(ok), _p32(orv), _p32(osv), cap, threads)
    if n < 0:
        raise MemoryError("oracle cpu_join allocation failed")
    w = min(n, cap)
    return int(n), ok[:w], orv[:w], osv[:w]


def nested_loop(r_key, r_val, s_key, s_val):
    r_key = np.ascontiguousarray(r_key, dtype=np.uint32)
    r_val = np.ascontiguousarray(r_val, dtype=np.uint32)
    s_key = np.ascontiguousarray(s_key, dtype=np.uint32)
    s_val = np.ascontiguousarray(s_val, dtype=np.uint32)
    cap = len(r_key) * len(s_key)
    ok = np.empty(max(cap, 1), np.uint32)
    orv = np.empty(max(cap, 1), np.uint32)
    osv = np.empty(max(cap, 1), np.uint32)
    n = lib().orc_nested_loop(_p32(r_key), _p32(r_val), len(r_key), _p32(s_key), _p32(s_val),
                              len(s_key), _p32(ok), _p32(orv), _p32(osv), cap)
    return int(n), ok[:n], orv[:n], osv[:n]


def sort_rows(key, r_val, s_val):
    """Canonical order: by (key, r_val, s_val).  Returns an (n,3) uint32 array."""
    key = np.asarray(key, np.uint32)
    r_val = np.asarray(r_val, np.uint32)
    s_val = np.asarray(s_val, np.uint32)
    packed = (key.astype(np.uint64) << np.uint64(32)) | r_val.astype(np.uint64)
    order = np.lexsort((s_val, packed))
    return np.stack([key[order], r_val[order], s_val[order]], axis=1)


# ---------------------------------------------------------------- bloom addressing
def radix_hash(key: int) -> int:
    return int(lib().orc_radix_hash(key))


def bloom_build(keys: np.ndarray, seg_bits: int, wps: int, k: int) -> np.ndarray:
    """Register-blocked filter: 2^seg_bits segments of wps 64-bit blocks."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    f = np.zeros((1 << seg_bits) * wps, np.uint64)
    lib().orc_bloom_build(_p32(keys), len(keys), seg_bits, wps, k,
                          f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return f


def bloom_query(filt: np.ndarray, keys: np.ndarray, seg_bits: int, wps: int, k: int) -> np.ndarray:
    L = lib()
    fp = filt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    return np.array([L.orc_bloom_query(fp, int(x), seg_bits, wps, k) for x in keys], dtype=bool)
