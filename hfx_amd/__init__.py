"""hfx_amd -- MI355X-native SEQUEST search core (drop-in for HFX's GPU path).

The compute path is libhfx_hip.so (hand-written HIP kernels for gfx950 behind
a C ABI, include/hfx_hip.h).  This package binds that ABI with ctypes and
provides the host-side glue that mirrors HFX's Search.hs orchestration.
Importing it never falls back to a CPU implementation: `lib()` raises when the
shared library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhfx_hip.so")

HFX_OK = 0

_lib = None


class HfxError(RuntimeError):
    pass


class HfxDb(C.Structure):
    """struct hfx_db (include/hfx_hip.h)."""

    _fields_ = [
        ("d_ions", C.c_void_p),
        ("d_r", C.c_void_p),
        ("d_tc", C.c_void_p),
        ("d_tn", C.c_void_p),
        ("d_mass", C.c_void_p),
        ("d_idx_sorted", C.c_void_p),
        ("d_r_sorted", C.c_void_p),
        ("num_pep", C.c_uint32),
        ("num_ions", C.c_uint64),
    ]


class HfxSpectra(C.Structure):
    """struct hfx_spectra (include/hfx_hip.h)."""

    _fields_ = [
        ("d_x", C.c_void_p),
        ("x_stride", C.c_uint32),
        ("d_len", C.c_void_p),
        ("d_zmax", C.c_void_p),
        ("d_mass", C.c_void_p),
        ("num_spec", C.c_uint32),
    ]


class HfxModset(C.Structure):
    """struct hfx_modset (include/hfx_hip.h)."""

    _fields_ = [
        ("d_mpep_pep_idx", C.c_void_p),
        ("d_mpep_mod_idx", C.c_void_p),
        ("d_mpep_unrank_off", C.c_void_p),
        ("d_unrank", C.c_void_p),
        ("d_mod_ma_count", C.c_void_p),
        ("d_mod_delta", C.c_void_p),
        ("d_ma_mass", C.c_void_p),
        ("num_ma", C.c_uint32),
        ("num_mpep", C.c_uint64),
    ]


_P = C.c_void_p
_U32 = C.c_uint32
_U64 = C.c_uint64
_F = C.c_float
_I = C.c_int

# name -> (restype, argtypes)
_SIGS = {
    "hfx_last_error": (C.c_char_p, []),
    "hfx_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "hfx_ctx_destroy": (_I, [_P]),
    "hfx_malloc": (_I, [C.POINTER(_P), C.c_size_t]),
    "hfx_free": (_I, [_P]),
    "hfx_host_alloc": (_I, [C.POINTER(_P), C.c_size_t]),
    "hfx_host_free": (_I, [_P]),
    "hfx_memcpy_h2d": (_I, [_P, _P, C.c_size_t, _P]),
    "hfx_memcpy_d2h": (_I, [_P, _P, C.c_size_t, _P]),
    "hfx_memcpy_d2d": (_I, [_P, _P, C.c_size_t, _P]),
    "hfx_memset": (_I, [_P, _I, C.c_size_t, _P]),
    "hfx_stream_create": (_I, [C.POINTER(_P)]),
    "hfx_stream_destroy": (_I, [_P]),
    "hfx_stream_sync": (_I, [_P]),
    "hfx_device_sync": (_I, []),
    "hfx_mem_info": (_I, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "hfx_device_count": (_I, [C.POINTER(_I)]),
    "hfx_event_create": (_I, [C.POINTER(_P)]),
    "hfx_event_destroy": (_I, [_P]),
    "hfx_event_record": (_I, [_P, _P]),
    "hfx_event_elapsed_ms": (_I, [C.POINTER(_F), _P, _P]),
    "hfx_blas_create": (_I, [C.POINTER(_P)]),
    "hfx_blas_destroy": (_I, [_P]),
    "hfx_blas_set_stream": (_I, [_P, _P]),
    # reference ABI (algorithms.h)
    "sort_idx_f": (None, [_P, _P, _U32, _U32]),
    "sort_val_f": (None, [_P, _P, _U32]),
    "sort_b40c_f": (None, [_P, _P, _U32]),
    "sort_rf": (None, [_P, _P, _U32]),
    "sort_idx_rf": (None, [_P, _P, _U32, _U32]),
    "findIndicesInRange_f": (_U32, [_P, _P, _U32, _F, _F]),
    "mvm_ff": (None, [_P, _P, _P, _P, _U32, _U32]),
    "prepare_ions": (None, [_P, _U32, _P, _U32]),
    "findSpecCandsByMass": (_U32, [_P, _P, _P, _P, _P, _U32, _P, _U32, _F]),
    "findBeginEnd_f": (_U32, [_P, _P, _P, _P, _P, _U32, _P, _U32, _F, _F]),
    "findModablePeptides": (
        _U32,
        [_P, _P, _P, _P, _P, _U32, _U32, _P, _P, _P, _P, _P, _P, _P, _P, _U32, _P, _U32],
    ),
    "calcTotalModCands": (_U32, [_P, _P, _P, _P, _P, _P, _U32, _U32, _U32]),
    "prepareGenMod": (_U32, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _U32, _U32]),
    "genModCands": (None, [_P, _P, _P, _P, _P, _P, _P, _P, _U32, _U32]),
    "addModIons": (
        None,
        [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _U32, _U32, _P, _P, _P, _P, _U32, _U32, _U32, _P],
    ),
    "addIons": (None, [_P, _P, _P, _P, _P, _P, _P, _U32, _U32, _U32]),
    "prepare_scoring": (None, [_P, _P, _P, _P, _U32]),
    # host helpers
    "hfx_mass_table": (_I, [_P, _I, _P]),
    "hfx_digestion_rule": (_I, [_I, _P]),
    "hfx_digest": (
        _I,
        [_P, _P, _U32, _P, _P, _U32, _F, _F, _I, _U32, _U32, _P, _P, _P, _U64,
         C.POINTER(_U64)],
    ),
    "hfx_mod_info": (_I, [_P, _U32, _P, _P, C.POINTER(_U32), _P, _P, _P, _U32, C.POINTER(_U32)]),
    "hfx_ms2_mass": (_F, [_F, _F]),
    "hfx_dta_precursor": (_F, [_F, _F]),
    # new entry points
    "hfx_mvm_rows": (_I, [_P, _P, _P, _U32, _P, _U32, _U32, _P]),
    "hfx_preprocess": (
        _I,
        [_P, _P, _U32, _P, _P, _P, _P, _P, _P, _P, _U32, _I, _I, _I, _P],
    ),
    "hfx_window": (_I, [_P, _P, _P, C.POINTER(HfxDb), _P, _U32, _F, _P]),
    "hfx_scan_offsets": (_I, [_P, _P, _P, _U32, _P]),
    "hfx_xcorr_sparse": (
        _I,
        [_P, _P, _P, _P, _P, _U64, C.POINTER(HfxDb), C.POINTER(HfxSpectra), _P],
    ),
    "hfx_xcorr_sparse_noalloc": (
        _I,
        [_P, _P, _P, _P, _U64, C.POINTER(HfxDb), C.POINTER(HfxSpectra), _P],
    ),
    "hfx_topn": (
        _I,
        [_P, _P, _P, _P, _P, _P, _U32, _U32, _P, _P, _U64, _P, _P, C.POINTER(HfxDb),
         C.POINTER(HfxSpectra), _P],
    ),
    "hfx_topn_exact": (
        _I,
        [_P, _P, _P, _P, _P, _U32, _P, _U32, _P, _P, _U64, _P, _P, C.POINTER(HfxDb),
         C.POINTER(HfxSpectra), _P],
    ),
    "hfx_topn_rel": (
        _I,
        [_P, _P, _P, _P, _P, _P, _U32, _U32, _F, _F, _P, _P, _U64, _P, _P, C.POINTER(HfxDb),
         C.POINTER(HfxSpectra), _P],
    ),
    "hfx_topn_rel_q": (
        _I,
        [_P, _P, _P, _P, _P, _P, _U32, _U32, _F, _F, _F, _P, _P, _U64, _P, _P,
         C.POINTER(HfxDb), C.POINTER(HfxSpectra), _P],
    ),
    "hfx_xcorr_dense_bf16x2": (
        _I,
        [_P, _P, _P, _P, _P, _P, C.POINTER(HfxDb), C.POINTER(HfxSpectra), _P],
    ),
    "hfx_xcorr_dense_i8x3": (
        _I,
        [_P, _P, _P, _P, _P, _P, C.POINTER(HfxDb), C.POINTER(HfxSpectra), _P],
    ),
    "hfx_mpep_offsets": (_I, [_P, _P, _P, _U32, _P, _U32, _P]),
    "hfx_mod_windows": (
        _I, [_P, _P, _P, _P, _P, C.POINTER(HfxDb), _P, _U32, _P, _U32, _F, _P]),
    "hfx_xcorr_sparse_mod": (
        _I,
        [_P, _P, _P, _U64, C.POINTER(HfxModset), C.POINTER(HfxDb), C.POINTER(HfxSpectra), _P],
    ),
    "hfx_topn_mod": (
        _I,
        [_P, _P, _P, _P, _P, _P, _U32, _U32, _P, _P, C.POINTER(HfxModset), C.POINTER(HfxDb),
         C.POINTER(HfxSpectra), _P],
    ),
    "hfx_topn_exact_mod": (
        _I,
        [_P, _P, _P, _P, _P, _U32, _P, _U32, _P, _P, C.POINTER(HfxModset), C.POINTER(HfxDb),
         C.POINTER(HfxSpectra), _P],
    ),
    "hfx_evalue_fit": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _U32, _U32, _F, _P]),
    "hfx_evalues": (_I, [_P, _P, _P, _P, _P, _U32, _U32, _P]),
    "hfx_sp_peak_words": (_U32, [_U32]),
    "hfx_sp_peaks": (_I, [_P, _P, _U32, C.POINTER(HfxSpectra), _P]),
    "hfx_sp_ions": (
        _I,
        [_P, _P, _P, _P, _P, _P, _P, _U64, _P, _U32, C.POINTER(HfxDb), C.POINTER(HfxSpectra),
         _I, _P],
    ),
    "hfx_sp_ions_topn": (
        _I,
        [_P, _P, _P, _P, _P, _P, _P, _U32, C.POINTER(HfxDb), C.POINTER(HfxSpectra), _I, _P],
    ),
    "hfx_sp_ions_topn_mod": (
        _I,
        [_P, _P, _P, _P, _P, _P, _P, _U32, C.POINTER(HfxModset), C.POINTER(HfxDb),
         C.POINTER(HfxSpectra), _I, _P],
    ),
    "hfx_seqdb_from_fasta": (_I, [C.c_char_p, C.c_char_p, _U64, _P, _I, _U32, _F, _F,
                                  C.POINTER(_P)]),
    "hfx_seqdb_free": (_I, [_P]),
    "hfx_seqdb_info": (_I, [_P, _P]),
    "hfx_seqdb_copy": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "hfx_seqdb_write_index": (_I, [_P, C.c_char_p, C.c_char_p]),
    "hfx_seqdb_read_index": (_I, [C.c_char_p, C.POINTER(_P)]),
    "hfx_seqdb_lookup": (_I, [_P, _U32, _P, _U32, C.POINTER(_U32), C.POINTER(_U32),
                              C.POINTER(_F)]),
    "hfx_digestion_rule_name": (C.c_char_p, [_I]),
    "hfx_dta_parse": (_I, [C.c_char_p, _U64, C.POINTER(_F), C.POINTER(_F), _P, _P, _U32,
                           C.POINTER(_U32)]),
    "hfx_mgf_parse": (_I, [C.c_char_p, _U64, C.POINTER(_P)]),
    "hfx_mgf_info": (_I, [_P, _P]),
    "hfx_mgf_copy": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "hfx_mgf_free": (_I, [_P]),
    "hfx_format_number": (_I, [C.c_double, _I, _I, _I, _P, _U32]),
    "hfx_format_table": (_I, [_P, _U32, _U32, _I, _P, _U64, C.POINTER(_U64)]),
    "hfx_format_detail": (_I, [_P, _U32, _P, _U64, C.POINTER(_U64)]),
    "hfx_format_scan": (_I, [_P, _P, _U32, _U32, _U32, _P, _U64, C.POINTER(_U64)]),
    "hfx_pepix_build": (_I, [_P, C.POINTER(HfxDb), _U32, _U32, _U32, _P, C.POINTER(_P)]),
    "hfx_pepix_destroy": (_I, [_P]),
    "hfx_pepix_info": (_I, [_P, _P]),
    "hfx_batch_create": (_I, [_P, C.POINTER(_P)]),
    "hfx_batch_destroy": (_I, [_P]),
    "hfx_batch_plan": (_I, [_P, C.POINTER(HfxDb), C.POINTER(HfxSpectra), _F, _U32, _U32, _P]),
    "hfx_batch_plan_ix": (_I, [_P, _P, C.POINTER(HfxSpectra), _F, _U32, _U32, _P]),
    "hfx_batch_use_index": (_I, [_P, _P]),
    "hfx_batch_required_bytes": (_I, [_U32, _U64, _U32, _U32, _U32, C.POINTER(_U64)]),
    "hfx_batch_split": (_I, [_P, _U32, _U32, _U32, _U32, _U64, _P, _U32, C.POINTER(_U32)]),
    "hfx_batch_run": (_I, [_P, _P]),
    "hfx_batch_capture": (_I, [_P, _P]),
    "hfx_batch_replay": (_I, [_P, _P]),
    "hfx_batch_autotune": (_I, [_P, _P, _P]),
    "hfx_batch_set_config": (_I, [_P, _U32, _U32, _U32]),
    "hfx_batch_set_precision": (_I, [_P, _U32]),
    "hfx_batch_time": (_I, [_P, _P, _P]),
    "hfx_batch_info": (_I, [_P, _P]),
    "hfx_batch_dense_blocks": (_I, [_P, _P, _U32, _P]),
    "hfx_batch_order": (_I, [_P, _P]),
    "hfx_batch_outputs": (_I, [_P, _P]),
}

# symbols that may be absent from a partial build (none at present)
_OPTIONAL: set[str] = set()


def lib() -> C.CDLL:
    """The loaded libhfx_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HfxError(
            f"{LIB_PATH} is missing: build it with `python -m hfx_amd.build` "
            "(there is no CPU fallback)"
        )
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        if not hasattr(L, name):
            if name in _OPTIONAL:
                continue
            raise HfxError(f"{LIB_PATH} does not export {name}")
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


def last_error() -> str:
    return (lib().hfx_last_error() or b"").decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    if rc != HFX_OK:
        raise HfxError(f"{what} failed with status {rc}: {last_error()}")
