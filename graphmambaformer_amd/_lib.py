"""ctypes binding of libgmfa.so (declarations in include/gmfa.h).

The library is built in-tree by ``graphmambaformer_amd.build`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared object is
missing or fails to load, every entry point raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (load torch's HIP runtime before libgmfa binds to it)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgmfa.so")

GMFA_OK = 0
GMFA_E_INVALID = -1
GMFA_E_RUNTIME = -2
GMFA_E_UNSUPPORTED = -3
GMFA_E_AGAIN = -4
GMFA_NOALLOC_SCRATCH_BYTES = 256 << 20
GMFA_SERVE_TRUNCATED = 1
GMFA_SERVE_NO_PAGES = 2

SW_TRACEBACK = 1
SW_NMASK = 2
SW_ROWKEYS = 4
SW_CIGAR = 8

OP_I, OP_D, OP_S, OP_EQ, OP_X = 1, 2, 4, 7, 8
OP_CHAR = {OP_I: "I", OP_D: "D", OP_S: "S", OP_EQ: "=", OP_X: "X"}

_P = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64


class SwParams(ctypes.Structure):
    _fields_ = [("match", _i32), ("mismatch", _i32), ("gap_open", _i32),
                ("gap_extend", _i32), ("half_band", _i32), ("flags", ctypes.c_uint32),
                ("cells_per_lane", _i32), ("lanes_per_pair", _i32),
                ("split_rows", _i32), ("split_slices", _i32), ("prio_rows", _i32),
                ("max_waves", _i32)]


class SwBatch(ctypes.Structure):
    _fields_ = [("n_pairs", _i32), ("max_qlen", _i32), ("qseq", _P), ("tseq", _P),
                ("qmask", _P), ("tmask", _P), ("q_words", _i64), ("t_words", _i64),
                ("q_off", _P), ("t_off", _P), ("qlen", _P), ("tlen", _P),
                ("band_offset", _P), ("order", _P)]


class SwOut(ctypes.Structure):
    _fields_ = [("score", _P), ("query_end", _P), ("target_end", _P), ("tb", _P),
                ("tb_off", _P), ("rowkey", _P), ("rowkey_off", _P), ("work_counter", _P),
                ("cigar_off", _P), ("cigar", _P), ("n_cigar", _P), ("read_start", _P),
                ("ref_start", _P), ("counts", _P)]


class SwTile(ctypes.Structure):
    _fields_ = [("cells_per_lane", _i32), ("lanes_per_pair", _i32), ("lanes_stored", _i32),
                ("half_band", _i32)]


class KmerIndex(ctypes.Structure):
    _fields_ = [("keys", _P), ("offsets", _P), ("positions", _P), ("dir", _P),
                ("n_keys", _i64), ("k", _i32), ("dir_bits", _i32)]


class KmerReads(ctypes.Structure):
    _fields_ = [("seq", _P), ("mask", _P), ("n_words", _i64), ("read_off", _P),
                ("read_len", _P), ("pos_off", _P), ("tile_read", _P), ("tile_start", _P),
                ("n_reads", _i32), ("n_tiles", _i32), ("window", _i32), ("reserved", _i32)]


class ServeClass(ctypes.Structure):
    _fields_ = [("batch", SwBatch), ("params", SwParams), ("out", SwOut),
                ("slab_words", _i64), ("n_slots", _i32), ("reserved", _i32),
                ("arrival_s", _P), ("complete_s", _P),
                ("page_blocks", _i32), ("n_pages", _i32), ("page_owner", _P),
                ("page_wait_s", ctypes.c_double)]


class ServeReport(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("completed", _i64), ("timed_out", _i32),
                ("waves_gave_up", _i32)]


class ServeClassCfg(ctypes.Structure):
    _fields_ = [("params", SwParams), ("max_qlen", _i32), ("ring", _i32), ("n_slots", _i32),
                ("cigar_cap", _i32), ("tb", _P), ("slab_words", _i64), ("stage", _P),
                ("head", _P), ("page_blocks", _i32), ("n_pages", _i32), ("page_owner", _P),
                ("page_wait_s", ctypes.c_double)]


class ServeResultRec(ctypes.Structure):
    _fields_ = [("ticket", _i64), ("cls", _i32), ("status", _i32), ("score", _i32),
                ("query_end", _i32), ("target_end", _i32), ("read_start", _i32),
                ("ref_start", _i32), ("n_cigar", _i32), ("counts", _i32 * 4),
                ("cigar_off", _i64), ("claim_tick", _i64), ("done_tick", _i64)]


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", _i64), ("returned", _i64), ("in_flight", _i64),
                ("launches", _i64), ("ticks_per_s", ctypes.c_double)]


# Every exported symbol with its ctypes signature (must match include/gmfa.h).
SIGNATURES = {
    "gmfa_last_error": (ctypes.c_char_p, []),
    "gmfa_abi_version": (ctypes.c_int, []),
    "gmfa_device_info": (ctypes.c_int, [_P, _P, ctypes.c_char_p, _i32]),
    "gmfa_sw_resolve_tile": (ctypes.c_int, [_i32, _i32, _i32, ctypes.POINTER(SwTile)]),
    "gmfa_sw_list_tiles": (ctypes.c_int, [_P, _P, _i32]),
    "gmfa_sw_tb_bytes": (_i64, [_i32, ctypes.POINTER(SwTile)]),
    "gmfa_sw_split_variants": (ctypes.c_int, [_i32, _i32, _P, _P, _i32]),
    "gmfa_sw_resident_waves": (ctypes.c_int, [ctypes.POINTER(SwParams), ctypes.POINTER(_i32)]),
    "gmfa_pack_codes": (ctypes.c_int, [_P, _i64, _i64, _P, _P, _i64, _P, _P, _P, _P]),
    "gmfa_pack_flat": (ctypes.c_int, [_P, _i64, _i32, _P, _P, _P, _P]),
    "gmfa_sw_forward": (ctypes.c_int, [ctypes.POINTER(SwBatch), ctypes.POINTER(SwParams),
                                       ctypes.POINTER(SwOut), _P]),
    "gmfa_sw_forward_noalloc": (ctypes.c_int, [ctypes.POINTER(SwBatch),
                                               ctypes.POINTER(SwParams),
                                               ctypes.POINTER(SwOut), _P]),
    "gmfa_sw_traceback": (ctypes.c_int, [ctypes.POINTER(SwBatch), ctypes.POINTER(SwParams),
                                         _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "gmfa_sw_xdrop": (ctypes.c_int, [_P, _P, _P, _P, _P, _i32, _i32, _i32, _i32, _P, _P, _P,
                                     _P, _P]),
    "gmfa_chain_dp_f64": (ctypes.c_int, [_P, _P, _P, _P, _i32, _i32, _i64, _P, _P, _P, _P,
                                         _P]),
    "gmfa_chain_dp_i32": (ctypes.c_int, [_P, _P, _P, _P, _i32, _i32, _i64, _i32, _i32, _P,
                                         _P, _P]),
    "gmfa_chain_backtrack": (_i64, [_i64, _P, _P, _i64, ctypes.c_double, _P, _P, _P]),
    "gmfa_cigar_bam": (ctypes.c_int, [_i64] + [_P] * 14),
    "gmfa_wfa_distance": (ctypes.c_int, [ctypes.POINTER(SwBatch), _i32, _i32, _P, _P, _P]),
    "gmfa_wfa_align": (ctypes.c_int, [ctypes.POINTER(SwBatch)] + [_P] * 9),
    "gmfa_chain_sort": (ctypes.c_int, [_P, _i32, _i64] + [_P] * 13 + [_P]),
    "gmfa_chain_backtrack_device": (ctypes.c_int, [_P, _i32, _i64] + [_P] * 7 +
                                    [_i64, ctypes.c_double] + [_P] * 8 + [_P]),
    "gmfa_chain_select": (ctypes.c_int, [_i32, _P, _P, _P, _P, _P, ctypes.c_double,
                                         ctypes.c_double, _i32, _P, _P, _P, _P, _P]),
    "gmfa_kmer_lookup": (ctypes.c_int, [ctypes.POINTER(KmerIndex), _P, _i64, _P, _P, _P]),
    "gmfa_kmer_seed": (ctypes.c_int, [ctypes.POINTER(KmerIndex), ctypes.POINTER(KmerReads), _P,
                                      _P, _P, _P, _P, _P]),
    "gmfa_kmer_emit": (ctypes.c_int, [ctypes.POINTER(KmerIndex), ctypes.POINTER(KmerReads), _P,
                                      _P, _P, _P, _P, _P, _P]),
    "gmfa_serve_run": (ctypes.c_int, [ctypes.POINTER(ServeClass), _i32, ctypes.c_double,
                                      ctypes.POINTER(ServeReport)]),
    "gmfa_serve_stage_words": (_i32, [_i32]),
    "gmfa_serve_open": (ctypes.c_int, [_P, _P, _i64, ctypes.POINTER(ServeClassCfg), _i32,
                                       ctypes.c_double, ctypes.POINTER(_P)]),
    "gmfa_serve_submit": (ctypes.c_int, [_P, _i32, _P, _i32, _i32, _i64, _i32, _i32,
                                         ctypes.POINTER(_i64)]),
    "gmfa_serve_submit_many": (ctypes.c_int, [_P, _i32, _P, _P, _P, _P, _i32, _P, _P, _P, _P,
                                              ctypes.POINTER(_i32)]),
    "gmfa_serve_poll": (ctypes.c_int, [_P, ctypes.POINTER(ServeResultRec), _i32, _P, _i64,
                                       ctypes.POINTER(_i32)]),
    "gmfa_serve_session_stats": (ctypes.c_int, [_P, ctypes.POINTER(ServeStats)]),
    "gmfa_serve_close": (ctypes.c_int, [_P, ctypes.POINTER(ServeReport)]),
}

_lib = None
_lock = threading.Lock()


class LibraryMissing(RuntimeError):
    """libgmfa.so is not built / not loadable."""


def lib():
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise LibraryMissing(
                    f"{LIB_PATH} is missing: build it with `python -m graphmambaformer_amd.build`")
            try:
                L = ctypes.CDLL(LIB_PATH)
            except OSError as exc:
                raise LibraryMissing(f"cannot load {LIB_PATH}: {exc}") from exc
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            if L.gmfa_abi_version() != 2:
                raise LibraryMissing("libgmfa ABI version mismatch")
            _lib = L
    return _lib


def check(rc: int, what: str = "") -> int:
    if rc >= 0:
        return rc
    msg = lib().gmfa_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if rc == GMFA_E_INVALID:
        raise ValueError(msg)
    raise RuntimeError(msg)


def ptr(t) -> int | None:
    """Device/host address of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_of(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream
