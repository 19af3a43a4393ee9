"""ctypes binding of libes_sieve.so (include/es_sieve.h).

The library is the only compute path: if it is missing or cannot be loaded this module raises,
and nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libes_sieve.so")

ES_ABI_VERSION = 3
ES_OK, ES_ERR_INVALID, ES_ERR_DEVICE, ES_ERR_OVERFLOW, ES_ERR_LIMIT = 0, -1, -2, -3, -4

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)


class PlanDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("k", C.c_uint32), ("n_rings", C.c_uint32),
        ("ring_mod", u32p), ("ring_radix", u32p), ("term_lo", u64p), ("term_hi", u64p),
        ("M_lo", C.c_uint64), ("M_hi", C.c_uint64),
        ("n_filters", C.c_uint32), ("filt_q", u32p), ("filt_off", u32p), ("filt_bits", u32p),
        ("n_filt_words", C.c_uint32), ("n_slow", C.c_uint32), ("slow_order", u32p),
        ("n_masks", C.c_uint32), ("mask_mod", u32p), ("mask_src", u32p), ("mask_off", u32p),
        ("mask_bits", u32p), ("n_mask_words", C.c_uint32),
        ("strip_target", C.c_uint32), ("threads", C.c_uint32), ("out_cap", C.c_uint64),
        ("lds_budget", C.c_uint32), ("global_masks", C.c_uint32), ("launch_log2", C.c_uint32),
        ("inner_mod", C.c_uint32), ("lead_radix", C.c_uint32), ("layout", C.c_uint32),
    ]


class EngineInfo(C.Structure):
    _fields_ = [
        ("R_lo", C.c_uint64), ("R_hi", C.c_uint64), ("R_outer", C.c_uint64),
        ("strips", C.c_uint64), ("r0", C.c_uint32), ("q0", C.c_uint32),
        ("strip_len", C.c_uint32), ("n_masks_used", C.c_uint32), ("lds_bytes", C.c_uint32),
        ("words", C.c_uint32), ("threads", C.c_uint32), ("device", C.c_uint32),
        ("global_masks", C.c_uint32), ("layout", C.c_uint32), ("steps", C.c_uint32),
        ("groups", C.c_uint32),
    ]


_lib = None


def lib() -> C.CDLL:
    """Load libes_sieve.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with build_native.py "
                           "(the MI355X engine has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    eng = C.c_void_p
    sig = {
        "es_last_error": (C.c_char_p, []),
        "es_device_query": (C.c_int, [C.POINTER(C.c_int), u64p, u64p, C.c_char_p, C.c_uint32]),
        "es_engine_create": (C.c_int, [C.POINTER(PlanDesc), C.POINTER(eng)]),
        "es_engine_destroy": (C.c_int, [eng]),
        "es_engine_info_get": (C.c_int, [eng, C.POINTER(EngineInfo)]),
        "es_engine_set_tconst": (C.c_int, [eng, u32p, C.c_uint32]),
        "es_engine_set_block": (C.c_int, [eng, C.c_uint64]),
        "es_engine_reset": (C.c_int, [eng]),
        "es_engine_scan": (C.c_int, [eng, C.c_uint64, C.c_uint64]),
        "es_engine_scan_window": (C.c_int, [eng, C.c_uint64, C.c_uint64]),
        "es_engine_scan_windows": (C.c_int, [eng, u64p, C.c_uint32, u64p, C.c_uint64, u64p,
                                             u64p, C.POINTER(C.c_uint8)]),
        "es_engine_count": (C.c_int, [eng, u64p]),
        "es_engine_read": (C.c_int, [eng, C.c_uint64, C.c_uint64, u64p]),
        "es_engine_collect": (C.c_int, [eng, u64p, C.c_uint64, u64p, u64p, C.POINTER(C.c_int)]),
        "es_engine_sync": (C.c_int, [eng]),
        "es_engine_kernel_time": (C.c_int, [eng, C.POINTER(C.c_double), u64p, u64p]),
        "es_engine_scan_window_static": (C.c_int, [eng, C.c_uint64, C.c_uint64, C.c_void_p,
                                                   C.c_uint64, C.c_void_p, C.c_void_p]),
        "es_batch_create": (C.c_int, [C.c_uint64, C.POINTER(eng)]),
        "es_batch_destroy": (C.c_int, [eng]),
        "es_batch_add": (C.c_int, [eng, eng, C.c_uint64, C.c_uint64, C.c_uint64, u32p]),
        "es_batch_run": (C.c_int, [eng, C.POINTER(C.c_double), u32p]),
        "es_batch_collect": (C.c_int, [eng, C.c_uint32, u64p, C.c_uint64, u64p]),
        "es_batch_capture": (C.c_int, [eng, u32p]),
        "es_batch_set_block": (C.c_int, [eng, C.c_uint32, C.c_uint64]),
        "es_batch_set_blocks": (C.c_int, [eng, u64p, C.c_uint32]),
        "es_batch_replay": (C.c_int, [eng, C.POINTER(C.c_double)]),
        "es_fpass_create": (C.c_int, [C.c_uint32, u32p, u32p, u64p, C.c_uint32, C.c_uint32,
                                      C.c_uint64, C.POINTER(eng)]),
        "es_fpass_destroy": (C.c_int, [eng]),
        "es_fpass_upload": (C.c_int, [eng, u64p, C.c_uint64, u32p, C.c_uint64]),
        "es_fpass_generate": (C.c_int, [eng, C.c_uint64, C.c_uint64, C.c_uint32, u32p, u32p,
                                        u32p, C.c_uint64]),
        "es_fpass_run": (C.c_int, [eng, u64p]),
        "es_fpass_read": (C.c_int, [eng, u32p, C.c_uint64]),
        "es_fpass_read_stream": (C.c_int, [eng, C.c_uint64, C.c_uint64, u64p, u32p]),
        "es_fpass_last_ms": (C.c_int, [eng, C.POINTER(C.c_double)]),
        "es_fpass_set_shape": (C.c_int, [eng, C.c_uint32, C.c_uint32]),
        "es_mitm_create": (C.c_int, [C.c_void_p, C.POINTER(eng)]),
        "es_mitm_destroy": (C.c_int, [eng]),
        "es_mitm_sizes": (C.c_int, [eng, u64p, u64p, u32p]),
        "es_mitm_scan": (C.c_int, [eng, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                   C.c_uint64, u64p]),
        "es_mitm_read": (C.c_int, [eng, u64p, C.c_uint64]),
        "es_mitm_last": (C.c_int, [eng, C.POINTER(C.c_double), u64p]),
        "es_calib_valu": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(C.c_double), u64p]),
        "es_calib_lds": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(C.c_double),
                                   u64p]),
        "es_calib_hbm_read": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.POINTER(C.c_double)]),
        "es_mlp_create": (C.c_int, [C.c_uint32, C.POINTER(eng)]),
        "es_mlp_destroy": (C.c_int, [eng]),
        "es_mlp_set_params": (C.c_int, [eng, f32p, C.c_uint32]),
        "es_mlp_get_params": (C.c_int, [eng, f32p, C.c_uint32]),
        "es_mlp_forward": (C.c_int, [eng, f32p, C.c_uint64, f32p]),
        "es_mlp_gradients": (C.c_int, [eng, f32p, f32p, C.c_uint32, f32p, f32p, f32p]),
        "es_mlp_adam_step": (C.c_int, [eng, f32p, f32p, C.c_uint32, f32p, C.c_float, C.c_float,
                                       C.c_float, C.c_float, f32p]),
        "es_mlp_upload_dataset": (C.c_int, [eng, f32p, f32p, C.c_uint64]),
        "es_mlp_train": (C.c_int, [eng, C.c_uint32, C.c_uint32, C.c_uint64, f32p, C.c_float,
                                   C.c_float, C.c_float, C.c_float, f32p,
                                   C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class EngineError(RuntimeError):
    pass


def check(rc: int) -> None:
    """Map an es_status to the reference's exception kinds."""
    if rc == ES_OK:
        return
    msg = lib().es_last_error().decode(errors="replace")
    if rc in (ES_ERR_INVALID, ES_ERR_LIMIT):
        raise AssertionError(msg)
    if rc == ES_ERR_OVERFLOW:
        raise RuntimeError(f"survivor buffer overflow: {msg}")
    raise EngineError(msg)


def ptr32(a):
    return a.ctypes.data_as(u32p)


def ptr64(a):
    return a.ctypes.data_as(u64p)


def device_query() -> dict:
    n = C.c_int(0)
    fr, tot = C.c_uint64(0), C.c_uint64(0)
    name = C.create_string_buffer(256)
    check(lib().es_device_query(C.byref(n), C.byref(fr), C.byref(tot), name, 256))
    text = name.value.decode(errors="replace")
    cus = None
    if text.endswith(" CUs)"):            # "<name> (<arch>, <N> CUs)"
        try:
            cus = int(text.rsplit(",", 1)[1].split()[0])
        except (IndexError, ValueError):
            cus = None
    return {"devices": n.value, "free_mib": fr.value, "total_mib": tot.value, "name": text,
            "cus": cus}
