"""ctypes mirror of include/mbx.h (the C ABI of libmbx.so).

The structs below must match mbx.h byte for byte; ``check_layout()`` compares
their sizes with what the library reports (mbx_abi_sizes).
"""
import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libmbx.so")

MAXBIN_PERCOL, MAXCOLS, MAXNINSTR, MAXKEYS, MAXBOUNDS = 4, 4, 6, 3, 8

# status codes
OK, ERR_INVALID, ERR_OOM, ERR_HIP, ERR_ALIGN, ERR_CAPACITY, ERR_NODEVICE, \
    ERR_ADMISSION, ERR_UNSUPPORTED = range(9)
# opcodes (vprg::ops, primitive.cuh:169-173)
END, AND, OR, NOT, ANDM, ORM = range(6)
KEY_IDENT, KEY_DIV, KEY_BUCKET = range(3)
MEAS_COUNT, MEAS_COL, MEAS_MUL, MEAS_SUB = range(4)
PLAN_FUSED, PLAN_MATERIALISE, PLAN_SCAN = 0, 1, 2
PLAN_MASK = 3
ACCUMULATE = 16
PLAN_EAGER = 32
PLAN_DIRECT = 64
PLAN_BASEFUSE = 128
PLAN_DEDICATED = 256


class ColRef(C.Structure):
    _fields_ = [("data", C.c_void_p), ("fk", C.c_void_p),
                ("dimsz", C.c_uint32), ("elem_bytes", C.c_uint32)]


class Wah(C.Structure):
    _fields_ = [("words", C.c_void_p), ("excl", C.c_void_p), ("skip", C.c_void_p),
                ("n", C.c_size_t), ("nskip", C.c_size_t)]


class Col(C.Structure):
    _fields_ = [("middle", C.c_void_p * MAXBIN_PERCOL),
                ("leftmost", C.c_void_p), ("rightmost", C.c_void_p)]


class Instr(C.Structure):
    _fields_ = [("opcode", C.c_uint8), ("dstreg", C.c_uint8),
                ("src1", C.c_uint8), ("src2", C.c_uint8)]


class Vprg(C.Structure):
    _fields_ = [("cols", Col * MAXCOLS), ("factsz", C.c_uint32),
                ("instrs", Instr * MAXNINSTR)]


class Pred(C.Structure):
    _fields_ = [("src", ColRef), ("lo", C.c_uint64), ("hi", C.c_uint64)]


class KeyPart(C.Structure):
    _fields_ = [("src", ColRef), ("kind", C.c_uint32), ("div", C.c_uint32),
                ("base", C.c_int64), ("stride", C.c_int64),
                ("nbounds", C.c_uint32), ("pad_", C.c_uint32),
                ("bounds", C.c_uint64 * MAXBOUNDS)]


class RowOp(C.Structure):
    _fields_ = [("preds", Pred * MAXCOLS), ("npreds", C.c_uint32),
                ("nkeys", C.c_uint32), ("keys", KeyPart * MAXKEYS),
                ("measure", C.c_uint32), ("pad_", C.c_uint32),
                ("ma", ColRef), ("mb", ColRef)]


class BinSet(C.Structure):
    _fields_ = [("data", C.c_void_p), ("elem_bytes", C.c_uint32),
                ("nbins", C.c_uint32), ("mins", C.POINTER(C.c_uint64)),
                ("maxs", C.POINTER(C.c_uint64)), ("slab", C.c_void_p)]


class Query(C.Structure):
    _fields_ = [("prg", Vprg), ("op", RowOp), ("nr_grp", C.c_uint32),
                ("flags", C.c_uint32), ("d_out", C.c_void_p)]


class Tune(C.Structure):
    _fields_ = [("tile_words", C.c_uint32), ("wg_per_cu", C.c_uint32),
                ("lds_groups", C.c_uint32), ("pad_", C.c_uint32)]


class Plan(C.Structure):
    _fields_ = [("order", C.c_uint8 * MAXCOLS), ("norder", C.c_uint8),
                ("reorderable", C.c_uint8), ("pad_", C.c_uint16),
                ("flags", C.c_uint32)]


class ServeCfg(C.Structure):
    _fields_ = [("nstreams", C.c_uint32), ("max_batch", C.c_uint32),
                ("mem_budget", C.c_uint64), ("plain", C.c_uint32),
                ("fused_rows", C.c_uint32), ("tune", Tune)]


class ServeStats(C.Structure):
    _fields_ = [("completed", C.c_uint32), ("rejected", C.c_uint32),
                ("wall_s", C.c_double), ("p50_s", C.c_double),
                ("p99_s", C.c_double), ("mean_s", C.c_double),
                ("max_s", C.c_double), ("batches", C.c_uint32),
                ("admission_waits", C.c_uint32), ("mean_batch", C.c_double),
                ("peak_mem", C.c_uint64), ("cancelled", C.c_uint32),
                ("failed", C.c_uint32), ("fixed_mem", C.c_uint64)]


# request states of the server (enum mbx_req_state)
REQ_QUEUED, REQ_ADMITTED, REQ_RUNNING, REQ_DONE, REQ_REJECTED, REQ_CANCELLED, \
    REQ_FAILED = range(7)


NFEAT = 16
ABI_STRUCTS = [ColRef, Col, Instr, Vprg, Pred, KeyPart, RowOp, Query, Tune, Plan,
               ServeCfg, ServeStats, Wah]

_vp, _sz, _u32, _u64, _i32 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_uint64, C.c_int
_P = C.POINTER

# name -> (restype, argtypes)
PROTOTYPES = {
    "mbx_init": (_i32, [_i32, _sz, _P(_vp)]),
    "mbx_destroy": (_i32, [_vp]),
    "mbx_strerror": (C.c_char_p, [_i32]),
    "mbx_last_error": (C.c_char_p, []),
    "mbx_device_count": (_i32, [_P(_i32)]),
    "mbx_sync": (_i32, [_vp]),
    "mbx_stream": (_vp, [_vp]),
    "mbx_set_stream": (_i32, [_vp, _vp]),
    "mbx_stream_create": (_i32, [_P(_vp)]),
    "mbx_stream_destroy": (_i32, [_vp]),
    "mbx_stream_join": (_i32, [_vp, _P(_vp), _i32]),
    "mbx_alloc": (_i32, [_vp, _sz, _P(_vp)]),
    "mbx_free": (_i32, [_vp, _vp]),
    "mbx_pool_stats": (_i32, [_vp, _P(_sz), _P(_sz), _P(_sz), _P(_sz)]),
    "mbx_memcpy_h2d": (_i32, [_vp, _vp, _vp, _sz]),
    "mbx_memcpy_d2h": (_i32, [_vp, _vp, _vp, _sz]),
    "mbx_memcpy_d2d": (_i32, [_vp, _vp, _vp, _sz]),
    "mbx_fill_u32": (_i32, [_vp, _vp, _u32, _sz]),
    "mbx_hostpool_create": (_i32, [_sz, _P(_vp)]),
    "mbx_hostpool_alloc": (_i32, [_vp, _sz, _P(_vp)]),
    "mbx_hostpool_free": (_i32, [_vp, _vp]),
    "mbx_hostpool_stats": (_i32, [_vp, _P(_sz), _P(_sz), _P(_sz), _P(_sz), _P(_sz)]),
    "mbx_hostpool_destroy": (_i32, [_vp]),
    "mbx_bitmap_words": (_sz, [_u32]),
    "mbx_bitmap_stride_words": (_sz, [_u32]),
    "mbx_create_bin": (_i32, [_vp, _P(ColRef), _u32, _P(_u64), _P(_u64), _sz,
                              _P(_vp), _P(_sz)]),
    "mbx_create_bin_into": (_i32, [_vp, _P(ColRef), _u32, _P(_u64), _P(_u64),
                                   _sz, _vp, _sz]),
    "mbx_create_bin_multi": (_i32, [_vp, _vp, _u32, _u32, _P(BinSet), _sz,
                                    _sz]),
    "mbx_create_select": (_i32, [_vp, _vp, _u32, _u32, _u32, _u32, _P(_vp)]),
    "mbx_create_join": (_i32, [_vp, _vp, _vp, _u32, _u32, _u32, _u32, _u32,
                               _P(_vp)]),
    "mbx_wah_groups_of": (_sz, [_u32]),
    "mbx_wah_select": (_i32, [_vp, _P(ColRef), _u32, _u64, _u64, _vp]),
    "mbx_wah_groups": (_i32, [_vp, _vp, _u32, _vp]),
    "mbx_wah_compress": (_i32, [_vp, _vp, _sz, _vp, _P(_sz)]),
    "mbx_wah_scan": (_i32, [_vp, _vp, _sz, _vp]),
    "mbx_wah_and_no1": (_i32, [_vp, _vp, _vp, _sz, _vp, _vp, _sz, _vp]),
    "mbx_wah_decompress": (_i32, [_vp, _vp, _vp, _sz, _vp, _sz]),
    "mbx_wah_longest_1fill": (_i32, [_vp, _vp, _sz, _P(_u64)]),
    "mbx_wah_skip_len": (_sz, [_sz]),
    "mbx_wah_scan_skip": (_i32, [_vp, _vp, _sz, _vp, _vp, _sz]),
    "mbx_wah_and_no1_skip": (_i32, [_vp, _P(Wah), _P(Wah), _vp]),
    "mbx_wah_decompress_skip": (_i32, [_vp, _P(Wah), _vp, _sz]),
    "mbx_prog_grpby": (_i32, [_vp, _P(Vprg), _P(RowOp), _u32, _vp, _u32]),
    "mbx_prog_grpby_noalloc": (_i32, [_vp, _P(Vprg), _P(RowOp), _u32, _vp,
                                      _u32]),
    "mbx_prog_eval": (_i32, [_vp, _P(Vprg), _vp, _vp]),
    "mbx_prog_eval_ex": (_i32, [_vp, _P(Vprg), _vp, _vp, _u32]),
    "mbx_prog_shape": (_i32, [_P(Vprg), _P(_u32), _P(C.c_int)]),
    "mbx_compact": (_i32, [_vp, _vp, _vp, _u32, _vp, _vp]),
    "mbx_rows_grpby": (_i32, [_vp, _vp, _vp, _P(Vprg), _P(RowOp), _u32, _vp,
                              _u32]),
    "mbx_alg_bytes": (_i32, [_vp, _P(Vprg), _P(RowOp), _u32, _P(_u64),
                             _P(_u64)]),
    "mbx_alg_bytes_batch": (_i32, [_vp, _P(Query), _u32, _P(_u64), _P(_u64)]),
    "mbx_batch_create": (_i32, [_vp, _P(Query), _u32, _P(Tune), _P(_vp)]),
    "mbx_batch_run": (_i32, [_vp, _vp]),
    "mbx_batch_run_eager": (_i32, [_vp, _vp]),
    "mbx_batch_destroy": (_i32, [_vp, _vp]),
    "mbx_autotune": (_i32, [_vp, _P(Query), _u32, _P(Tune), _P(C.c_float)]),
    "mbx_autotune_report": (_i32, [_vp, _P(Query), _u32, _P(Tune),
                                   _P(C.c_float), _P(Tune), _P(C.c_float),
                                   _i32, _P(_i32)]),
    "mbx_default_tune": (_i32, [_P(Tune)]),
    "mbx_serve_replay": (_i32, [_vp, _P(Query), _u32, _P(C.c_double),
                                _P(ServeCfg), _vp, _P(C.c_double),
                                _P(ServeStats)]),
    "mbx_serve_fixed_mem": (_i32, [_P(ServeCfg), _P(_u64)]),
    "mbx_serve_open": (_i32, [_vp, _P(ServeCfg), _P(_vp)]),
    "mbx_serve_submit": (_i32, [_vp, _P(Query), _vp, _P(_u64)]),
    "mbx_serve_poll": (_i32, [_vp, _u64, _P(_i32), _P(C.c_double)]),
    "mbx_serve_wait": (_i32, [_vp, _u64, C.c_double, _P(_i32),
                              _P(C.c_double)]),
    "mbx_serve_cancel": (_i32, [_vp, _u64]),
    "mbx_serve_close": (_i32, [_vp, _P(ServeStats)]),
    "mbx_timer_create": (_i32, [_vp, _P(_vp)]),
    "mbx_timer_start": (_i32, [_vp, _vp]),
    "mbx_timer_stop": (_i32, [_vp, _vp, _P(C.c_float)]),
    "mbx_timer_destroy": (_i32, [_vp]),
    "mbx_timer_record_stop": (_i32, [_vp, _vp]),
    "mbx_timer_elapsed": (_i32, [_vp, _P(C.c_float)]),
    "mbx_copy_kernel": (_i32, [_vp, _vp, _vp, _sz]),
    "mbx_abi_sizes": (_i32, [_P(_sz), _i32]),
    # plans and the cost-model MLP
    "mbx_col_density": (_i32, [_vp, _P(Vprg), _P(C.c_double)]),
    "mbx_col_stats": (_i32, [_vp, _P(Vprg), _P(C.c_double), _P(C.c_double)]),
    "mbx_plan_features_ex": (_i32, [_P(Vprg), _P(C.c_double), _P(C.c_double),
                                    _u32, _P(Plan), _P(C.c_float)]),
    "mbx_plan_candidates": (_i32, [_P(Vprg), _P(C.c_double), _P(Plan), _i32,
                                   _P(_i32)]),
    "mbx_plan_choose_ex": (_i32, [_vp, _vp, _P(Vprg), _P(C.c_double),
                                  _P(C.c_double), _u32, C.c_float, _P(Plan),
                                  _P(C.c_float)]),
    "mbx_plan_describe": (_i32, [_P(Vprg), _P(Plan)]),
    "mbx_plan_apply": (_i32, [_P(Vprg), _P(Plan), _P(Vprg)]),
    "mbx_plan_features": (_i32, [_P(Vprg), _P(C.c_double), _u32, _P(Plan),
                                 _P(C.c_float)]),
    "mbx_plan_enumerate": (_i32, [_P(Vprg), _P(Plan), _i32, _P(_i32)]),
    "mbx_plan_heuristic": (_i32, [_P(Vprg), _P(C.c_double), _P(Plan)]),
    "mbx_plan_choose": (_i32, [_vp, _vp, _P(Vprg), _P(C.c_double), _u32,
                               _P(Plan), _P(C.c_float)]),
    "mbx_prog_grpby_plan": (_i32, [_vp, _P(Vprg), _P(RowOp), _u32, _vp,
                                   _P(Plan)]),
    "mbx_mlp_create": (_i32, [_vp, _i32, _i32, _i32, _u64, _P(_vp)]),
    "mbx_mlp_destroy": (_i32, [_vp, _vp]),
    "mbx_mlp_nparams": (_i32, [_vp]),
    "mbx_mlp_set_params": (_i32, [_vp, _vp, _vp]),
    "mbx_mlp_get_params": (_i32, [_vp, _vp, _vp]),
    "mbx_mlp_forward": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "mbx_mlp_loss_grad": (_i32, [_vp, _vp, _vp, _vp, _i32, _P(C.c_double),
                                 _vp]),
    "mbx_mlp_train": (_i32, [_vp, _vp, _vp, _vp, _i32, _i32, C.c_float,
                             _P(C.c_double)]),
    # host-only planning helpers (include/mbx_bmpidx.hpp)
    "mbx_col_from_bin": (_i32, [_sz, _P(_u64), _P(_u64), _u64, _u64, _P(_u64)]),
    "mbx_col_from_bin2": (_i32, [_sz, _P(_u64), _P(_u64), _sz, _P(_u64),
                                 _P(_u64), _u64, _u64, _P(_u64)]),
    "mbx_recipe_plan": (_i32, [_i32, _i32, _u32, _u32, _P(_u32), _P(_u32),
                               _P(_i32), _P(_i32)]),
    # generators and column files
    "mbx_rand_at": (_u64, [_u64, _u64]),
    "mbx_zipf_table": (_i32, [_u32, C.c_double, _P(_u64)]),
    "mbx_gen_uniform_u32_host": (_i32, [_vp, _sz, _u64, _u32, _u32]),
    "mbx_gen_zipf_u32_host": (_i32, [_vp, _sz, _u64, _u32, C.c_double]),
    "mbx_gen_uniform_u32": (_i32, [_vp, _vp, _sz, _u64, _u32, _u32]),
    "mbx_gen_zipf_u32": (_i32, [_vp, _vp, _sz, _u64, _u32, C.c_double]),
    "mbx_gen_map_u32": (_i32, [_vp, _vp, _vp, _sz, _u32, _u32]),
    "mbx_colfile_write": (_i32, [C.c_char_p, _vp, _sz, _u32, _i32]),
    "mbx_colfile_read": (_i32, [C.c_char_p, _P(_vp), _P(_sz), _P(_u32)]),
    "mbx_colfile_free": (None, [_vp]),
    "mbx_colfile_load": (_i32, [_vp, C.c_char_p, _P(_vp), _P(_sz), _P(_u32)]),
}


def bind(lib):
    """Attach prototypes; returns the names that the library does not export."""
    missing = []
    for name, (res, args) in PROTOTYPES.items():
        try:
            f = getattr(lib, name)
        except AttributeError:
            missing.append(name)
            continue
        f.restype = res
        f.argtypes = args
    return missing


def check_layout(lib):
    n = len(ABI_STRUCTS)
    out = (C.c_size_t * n)()
    m = lib.mbx_abi_sizes(out, n)
    if m != n:
        raise RuntimeError(f"ABI struct count mismatch: lib {m}, python {n}")
    bad = [(s.__name__, C.sizeof(s), out[i]) for i, s in enumerate(ABI_STRUCTS)
           if C.sizeof(s) != out[i]]
    if bad:
        raise RuntimeError(f"ABI layout mismatch (name, python, C): {bad}")
