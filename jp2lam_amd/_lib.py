"""ctypes binding of ``libjp2lam_amd.so`` (include/jp2lam_amd.h).

The library is loaded from this package directory only.  There is no CPU
fallback: if the HIP library is missing or no device is present, the codec
calls raise.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libjp2lam_amd.so"

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)


class EncodeOptionsC(C.Structure):
    _fields_ = [("quality", C.c_uint8), ("format", C.c_uint8)]


class ComponentC(C.Structure):
    _fields_ = [("data", i32p), ("width", C.c_uint32), ("height", C.c_uint32), ("precision", C.c_uint32),
                ("is_signed", C.c_uint32), ("dx", C.c_uint32), ("dy", C.c_uint32)]


class ImageViewC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("colorspace", C.c_uint32),
                ("num_components", C.c_uint32), ("components", C.POINTER(ComponentC))]


class ImageC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("colorspace", C.c_uint32),
                ("num_components", C.c_uint32), ("precision", C.c_uint32), ("planes", i32p)]


class MetadataC(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "width", "height", "colorspace", "num_components", "bits_per_component", "has_ipr_metadata",
        "decomposition_levels", "code_block_width", "code_block_height", "layers", "transform", "use_mct",
        "guard_bits", "num_quant_steps")] + [("tile_part_count", C.c_uint64), ("first_tile_payload_len", C.c_uint64)]


class EncodeItemC(C.Structure):
    _fields_ = [("pixels", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32),
                ("quality", C.c_uint8), ("format", C.c_uint8), ("on_device", C.c_uint8), ("pad", C.c_uint8),
                ("out", u8p), ("out_len", C.c_size_t), ("status", C.c_int32)]


class DecodeItemC(C.Structure):
    _fields_ = [("bytes", C.c_void_p), ("len", C.c_size_t), ("pixels", C.c_void_p), ("cap", C.c_size_t),
                ("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32),
                ("on_device", C.c_uint8), ("pad", C.c_uint8 * 3), ("status", C.c_int32)]


class ComponentProfileC(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("width", "height", "precision", "is_signed", "dx", "dy")]


class BatchProfileC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("colorspace", C.c_uint32),
                ("num_components", C.c_uint32), ("quality", C.c_uint8), ("format", C.c_uint8),
                ("pad", C.c_uint8 * 2)]


class MetricsC(C.Structure):
    _fields_ = [("psnr_db", C.c_double), ("ssim", C.c_double)]


class ContentStatsC(C.Structure):
    _fields_ = [("bright_share", C.c_float), ("entropy", C.c_float), ("colorfulness", C.c_float),
                ("content_class", C.c_uint32), ("quality", C.c_uint32)]


class ServedC(C.Structure):
    _fields_ = [("status", C.c_int32), ("data", u8p), ("len", C.c_size_t), ("width", C.c_uint32),
                ("height", C.c_uint32), ("channels", C.c_uint32), ("batch_size", C.c_uint32),
                ("queue_ms", C.c_double), ("run_ms", C.c_double), ("total_ms", C.c_double)]


# (name, restype, argtypes) for every symbol the header declares
SIGNATURES = [
    ("jp2lam_encode", C.c_int, [C.POINTER(ImageViewC), C.POINTER(EncodeOptionsC), C.POINTER(u8p), C.POINTER(C.c_size_t)]),
    ("jp2lam_analyze_content", C.c_int, [C.POINTER(ImageViewC), C.POINTER(ContentStatsC)]),
    ("jp2lam_analyze_content_u8", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int,
                                            C.POINTER(ContentStatsC)]),
    ("jp2lam_analyze_content_u8_noalloc", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                                    C.POINTER(ContentStatsC)]),
    ("jp2lam_encode_with_psnr", C.c_int, [C.POINTER(ImageViewC), C.POINTER(EncodeOptionsC), C.POINTER(u8p),
                                          C.POINTER(C.c_size_t), C.POINTER(MetricsC)]),
    ("jp2lam_encode_u8", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(EncodeOptionsC),
                                   C.POINTER(u8p), C.POINTER(C.c_size_t)]),
    ("jp2lam_decode_jp2", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(ImageC)]),
    ("jp2lam_decode_jp2_u8", C.c_int, [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("jp2lam_inspect_jp2", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(MetadataC)]),
    ("jp2lam_free", None, [C.c_void_p]),
    ("jp2lam_image_free", None, [C.POINTER(ImageC)]),
    ("jp2lam_last_error", C.c_char_p, []),
    ("jp2lam_batch_encoder_new", C.c_void_p, [C.POINTER(EncodeOptionsC)]),
    ("jp2lam_batch_encoder_encode_one", C.c_int, [C.c_void_p, C.POINTER(ImageViewC), C.POINTER(u8p), C.POINTER(C.c_size_t)]),
    ("jp2lam_batch_encoder_encoded_count", C.c_size_t, [C.c_void_p]),
    ("jp2lam_batch_encoder_has_profile", C.c_int, [C.c_void_p]),
    ("jp2lam_batch_encoder_free", None, [C.c_void_p]),
    ("jp2lam_batch_encoder_profile", C.c_int, [C.c_void_p, C.POINTER(BatchProfileC), C.POINTER(ComponentProfileC),
                                              C.c_size_t]),
    ("jp2lam_batch_decoder_new", C.c_void_p, []),
    ("jp2lam_batch_decoder_decode_one", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(ImageC)]),
    ("jp2lam_batch_decoder_decoded_count", C.c_size_t, [C.c_void_p]),
    ("jp2lam_batch_decoder_free", None, [C.c_void_p]),
    ("jp2lam_batch_decoder_profile", C.c_int, [C.c_void_p, C.POINTER(BatchProfileC), C.POINTER(ComponentProfileC),
                                              C.c_size_t]),
    ("jp2lam_encode_many", C.c_int, [C.POINTER(EncodeItemC), C.c_size_t]),
    ("jp2lam_decode_many", C.c_int, [C.POINTER(DecodeItemC), C.c_size_t]),
    ("jp2g_forward_dwt", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    ("jp2g_inverse_dwt", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]),
    ("jp2g_t1_encode_block", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_size_t]),
    ("jp2g_t1_decode_block", C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.c_void_p, C.c_size_t, C.c_void_p]),
    ("jp2lam_device_alloc", C.c_void_p, [C.c_size_t]),
    ("jp2lam_device_copy", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    ("jp2lam_device_free", None, [C.c_void_p]),
    ("jp2lam_server_new", C.c_void_p, [C.c_int, C.c_size_t, C.c_int]),
    ("jp2lam_server_submit_encode", C.c_uint64, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                                 C.c_uint8, C.c_uint8, C.c_uint8]),
    ("jp2lam_server_submit_decode", C.c_uint64, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("jp2lam_server_wait", C.c_int, [C.c_void_p, C.c_uint64, C.c_int, C.POINTER(ServedC)]),
    ("jp2lam_server_stats", None, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("jp2lam_server_new_simulated", C.c_void_p, [C.c_int, C.c_size_t, C.c_int, C.c_uint32]),
    ("jp2lam_server_cancel", C.c_int, [C.c_void_p, C.c_uint64]),
    ("jp2lam_server_faults", None, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("jp2lam_server_free", None, [C.c_void_p]),
    ("jp2g_set_debug", None, [C.c_int]),
    ("jp2g_set_t1_mode", None, [C.c_int, C.c_uint32]),
    ("jp2g_last_t1_lanes", C.c_int, []),
    ("jp2g_last_t1_coder", C.c_int, []),
    ("jp2g_set_hybrid_frac", None, [C.c_float]),
    ("jp2g_set_queue", None, [C.c_uint32, C.c_int]),
    ("jp2g_get_queue", None, [C.POINTER(C.c_uint32), C.POINTER(C.c_int)]),
    ("jp2g_set_graphs", None, [C.c_int]),
    ("jp2g_graph_stats", None, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("jp2g_set_counters", None, [C.c_int]),
    ("jp2g_counters", None, [C.c_void_p]),
    ("jp2g_reset_counters", None, []),
    ("jp2g_last_block_symbols", C.c_uint32, [C.c_int, C.c_int]),
    ("jp2g_last_block_count", C.c_int, []),
    ("jp2g_last_block_info", None, [C.c_int, C.c_void_p]),
    ("jp2g_last_block_passes", None, [C.c_int, C.c_void_p]),
    ("jp2g_last_block_bytes", None, [C.c_int, C.c_void_p]),
    ("jp2g_last_lambda", C.c_double, [C.c_int]),
    ("jp2g_last_timings", None, [C.c_void_p]),
    ("jp2g_tables", None, [C.c_void_p] * 7),
    ("jp2g_device_available", C.c_int, []),
]

_lib = None


def load() -> C.CDLL:
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m jp2lam_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
