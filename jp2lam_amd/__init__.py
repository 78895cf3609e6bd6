"""jp2lam_amd: MI355X-native JPEG 2000 Part 1 codec core.

A Python mirror of the jp2lam crate API (src/lib.rs:19-29) over the C ABI
in ``include/jp2lam_amd.h``: :func:`encode`, :func:`decode_jp2`,
:func:`inspect_jp2`, :class:`BatchEncoder` / :class:`BatchDecoder`,
:func:`encode_with_psnr`, :class:`EncodeOptions` with quality 0..=100,
:class:`Image` with planar i32 components and the :class:`Jp2LamError`
family.  Every codec call runs the
HIP kernels in ``libjp2lam_amd.so``; there is no CPU fallback.

Extensions beyond the crate API: :func:`encode_u8` / :func:`decode_u8` for
interleaved byte images, and :func:`encode_many` / :func:`decode_many`
for mixed-size batches in one device launch sequence.
"""
from __future__ import annotations

import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib

__all__ = [
    "ColorSpace", "OutputFormat", "Preset", "EncodeOptions", "Component", "Image", "DecodeMetadata",
    "Jp2LamError", "InvalidInput", "EncodeFailed", "DecodeFailed", "UnsupportedFeature", "IoError",
    "encode", "encode_to_writer", "encode_with_psnr", "EncodeMetrics", "decode_jp2", "decode_from_reader", "inspect_jp2", "BatchEncoder",
    "BatchDecoder", "BatchProfile", "BatchComponentProfile", "encode_batch", "decode_batch", "encode_u8", "decode_u8", "encode_many", "decode_many",
    "Server", "Served", "ContentClass", "ContentStats", "analyze_content", "auto_quality",
    "auto_class_and_quality", "analyze_content_u8", "PHOTO_QUALITY", "GRAPHICS_QUALITY",
]


# ---- model (src/model.rs) ---------------------------------------------------------
class ColorSpace(enum.IntEnum):
    """ColorSpace in the declaration order of src/model.rs:1-12."""
    GRAY = 0
    RGB = 1
    SRGB = 2
    YUV = 3
    YCBCR = 4

    def component_count(self) -> int:
        return 1 if self is ColorSpace.GRAY else 3


class OutputFormat(enum.IntEnum):
    JP2 = 0
    J2K = 1


class Preset(enum.Enum):
    """Named presets (src/model.rs:40-62)."""
    DOCUMENT_LOW = 30
    DOCUMENT_HIGH = 85
    WEB_LOW = 42
    WEB_HIGH = 62

    @property
    def quality(self) -> int:
        return self.value


@dataclass
class EncodeOptions:
    """Quality 0..=100 (100 = lossless 5/3) and container format; the
    default is q85 JP2 (src/model.rs:155-177)."""
    quality: int = Preset.DOCUMENT_HIGH.value
    format: OutputFormat = OutputFormat.JP2

    @classmethod
    def from_preset(cls, preset: Preset, format: OutputFormat = OutputFormat.JP2) -> "EncodeOptions":
        return cls(preset.quality, format)


@dataclass
class Component:
    data: np.ndarray
    width: int
    height: int
    precision: int = 8
    signed: bool = False
    dx: int = 1
    dy: int = 1


@dataclass
class Image:
    width: int
    height: int
    components: List[Component]
    colorspace: ColorSpace = ColorSpace.GRAY

    @classmethod
    def from_gray_bytes(cls, width: int, height: int, data) -> "Image":
        buf = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data.reshape(-1)
        if buf.size != width * height:
            raise InvalidInput(f"Gray buffer length {buf.size} does not match {width}×{height}={width * height}")
        return cls(width, height, [Component(buf.astype(np.int32), width, height)], ColorSpace.GRAY)

    @classmethod
    def from_rgb_bytes(cls, width: int, height: int, data) -> "Image":
        buf = np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else data.reshape(-1)
        if buf.size != width * height * 3:
            raise InvalidInput(
                f"RGB buffer length {buf.size} does not match {width}×{height}×3={width * height * 3}")
        px = buf.reshape(-1, 3).astype(np.int32)
        comps = [Component(np.ascontiguousarray(px[:, c]), width, height) for c in range(3)]
        return cls(width, height, comps, ColorSpace.SRGB)

    def to_numpy(self) -> np.ndarray:
        """Samples as uint8, (H, W) or (H, W, C)."""
        planes = [np.asarray(c.data, dtype=np.int32).reshape(self.height, self.width) for c in self.components]
        a = planes[0] if len(planes) == 1 else np.stack(planes, axis=-1)
        return a.astype(np.uint8)


@dataclass
class DecodeMetadata:
    width: int
    height: int
    colorspace: ColorSpace
    num_components: int
    bits_per_component: int
    has_ipr_metadata: bool
    decomposition_levels: int
    code_block_width: int
    code_block_height: int
    layers: int
    transform: int
    use_mct: bool
    guard_bits: int
    num_quant_steps: int
    tile_part_count: int
    first_tile_payload_len: int


# ---- errors (src/error.rs) -------------------------------------------------------------
class Jp2LamError(Exception):
    code = 0
    prefix = "error"

    def __init__(self, message: str):
        super().__init__(message)
        self.message = message

    def __str__(self) -> str:
        return f"{self.prefix}: {self.message}"

    def is_decode_failure(self) -> bool:
        return isinstance(self, (DecodeFailed, UnsupportedFeature))

    def is_unsupported_feature(self) -> bool:
        return isinstance(self, UnsupportedFeature)


class InvalidInput(Jp2LamError):
    code, prefix = 1, "invalid input"


class EncodeFailed(Jp2LamError):
    code, prefix = 2, "encode failed"


class DecodeFailed(Jp2LamError):
    code, prefix = 3, "decode failed"


class UnsupportedFeature(Jp2LamError):
    code, prefix = 4, "unsupported feature"


class IoError(Jp2LamError):
    code, prefix = 5, "I/O error"


_ERRORS = {1: InvalidInput, 2: EncodeFailed, 3: DecodeFailed, 4: UnsupportedFeature, 5: IoError}


def _check(status: int) -> None:
    if status != 0:
        msg = _lib.load().jp2lam_last_error().decode("utf-8", "replace")
        raise _ERRORS.get(status, IoError)(msg)


# ---- marshalling ------------------------------------------------------------------------
class _View:
    """Keeps the i32 planes alive while the C view is in use."""

    def __init__(self, image: Image):
        self.arrays = [np.ascontiguousarray(np.asarray(c.data, dtype=np.int32)) for c in image.components]
        comps = (_lib.ComponentC * max(len(image.components), 1))()
        for i, (c, a) in enumerate(zip(image.components, self.arrays)):
            comps[i] = _lib.ComponentC(a.ctypes.data_as(_lib.i32p), c.width, c.height, c.precision,
                                       1 if c.signed else 0, c.dx, c.dy)
            if a.size != c.width * c.height:
                raise InvalidInput(f"component {i} has {a.size} samples, expected {c.width * c.height}")
        self.comps = comps
        self.view = _lib.ImageViewC(image.width, image.height, int(image.colorspace), len(image.components), comps)


def _opts(options: Optional[EncodeOptions]) -> _lib.EncodeOptionsC:
    o = options or EncodeOptions()
    return _lib.EncodeOptionsC(min(int(o.quality), 255), int(o.format))


def _take_bytes(ptr, n: int) -> bytes:
    lib = _lib.load()
    try:
        return C.string_at(ptr, n)
    finally:
        lib.jp2lam_free(ptr)


def _image_from_c(img: _lib.ImageC) -> Image:
    lib = _lib.load()
    try:
        n = img.width * img.height
        flat = np.ctypeslib.as_array(img.planes, shape=(img.num_components * n,)).copy()
    finally:
        lib.jp2lam_image_free(C.byref(img))
    comps = [Component(flat[c * n:(c + 1) * n], img.width, img.height) for c in range(img.num_components)]
    return Image(img.width, img.height, comps, ColorSpace(img.colorspace))


def _as_buffer(data) -> bytes:
    return data if isinstance(data, (bytes, bytearray)) else bytes(data)


# ---- crate API ---------------------------------------------------------------------------
def encode(image: Image, options: Optional[EncodeOptions] = None) -> bytes:
    """encode(&Image, &EncodeOptions) (src/encode/mod.rs:84-101)."""
    lib = _lib.load()
    v = _View(image)
    out = _lib.u8p()
    n = C.c_size_t()
    o = _opts(options)
    _check(lib.jp2lam_encode(C.byref(v.view), C.byref(o), C.byref(out), C.byref(n)))
    return _take_bytes(out, n.value)


@dataclass(frozen=True)
class EncodeMetrics:
    """EncodeMetrics (src/encode/mod.rs:114-122): PSNR in dB (inf for
    lossless) and mean SSIM over 8x8 luma blocks (1.0 for lossless)."""
    psnr_db: float
    ssim: float


def encode_with_psnr(image: Image, options: Optional[EncodeOptions] = None) -> Tuple[bytes, EncodeMetrics]:
    """encode_with_psnr (src/encode/mod.rs:124-135): the codestream and the
    reference's internal reconstruction metrics (backend.rs:297-433),
    computed on the device from the encode's own coefficients."""
    lib = _lib.load()
    v = _View(image)
    out = _lib.u8p()
    n = C.c_size_t()
    o = _opts(options)
    m = _lib.MetricsC()
    _check(lib.jp2lam_encode_with_psnr(C.byref(v.view), C.byref(o), C.byref(out), C.byref(n), C.byref(m)))
    return _take_bytes(out, n.value), EncodeMetrics(m.psnr_db, m.ssim)


# ---- content-adaptive quality (src/content.rs) --------------------------------------
PHOTO_QUALITY = 78
GRAPHICS_QUALITY = 55


class ContentClass(enum.Enum):
    """ContentClass (content.rs:55-78)."""
    PHOTO = 0
    GRAPHICS = 1

    def quality(self) -> int:
        return GRAPHICS_QUALITY if self is ContentClass.GRAPHICS else PHOTO_QUALITY

    def label(self) -> str:
        return "graphics" if self is ContentClass.GRAPHICS else "photo"


@dataclass(frozen=True)
class ContentStats:
    """ContentStats (content.rs:80-106): f32 statistics of one histogram pass."""
    bright_share: float
    entropy: float
    colorfulness: float

    def classify(self) -> ContentClass:
        # content.rs:96-105, compared in f32 like the reference
        e, b = np.float32(self.entropy), np.float32(self.bright_share)
        graphics = e < np.float32(3.0) or (b > np.float32(0.45) and e < np.float32(5.0))
        return ContentClass.GRAPHICS if graphics else ContentClass.PHOTO


def _content(st: "_lib.ContentStatsC") -> ContentStats:
    return ContentStats(st.bright_share, st.entropy, st.colorfulness)


def analyze_content(image: Image) -> ContentStats:
    """analyze (content.rs:119-166): luma histogram and chroma spread of
    ``image``, one pass on the device."""
    lib = _lib.load()
    v = _View(image)
    st = _lib.ContentStatsC()
    _check(lib.jp2lam_analyze_content(C.byref(v.view), C.byref(st)))
    return _content(st)


def analyze_content_u8(pixels: np.ndarray) -> ContentStats:
    """analyze on interleaved u8 pixels, (H, W) gray or (H, W, 3) RGB."""
    lib = _lib.load()
    px = np.ascontiguousarray(pixels, dtype=np.uint8)
    ch = 1 if px.ndim == 2 else px.shape[2]
    st = _lib.ContentStatsC()
    _check(lib.jp2lam_analyze_content_u8(px.ctypes.data, px.shape[1], px.shape[0], ch, 0, C.byref(st)))
    return _content(st)


def auto_quality(image: Image) -> int:
    """auto_quality (content.rs:168-171)."""
    return analyze_content(image).classify().quality()


def auto_class_and_quality(image: Image) -> Tuple[ContentClass, int]:
    """auto_class_and_quality (content.rs:173-177)."""
    cls = analyze_content(image).classify()
    return cls, cls.quality()


def encode_to_writer(image: Image, options: Optional[EncodeOptions], writer) -> None:
    data = encode(image, options)
    try:
        writer.write(data)
    except Exception as e:  # encode/mod.rs:103-112 maps writer errors to EncodeFailed
        raise EncodeFailed(str(e)) from e


def decode_jp2(data) -> Image:
    """decode_jp2(&[u8]) (src/decode/mod.rs:56-61)."""
    lib = _lib.load()
    buf = _as_buffer(data)
    img = _lib.ImageC()
    _check(lib.jp2lam_decode_jp2(buf, len(buf), C.byref(img)))
    return _image_from_c(img)


def decode_from_reader(reader) -> Image:
    try:
        data = reader.read()
    except Exception as e:  # decode/mod.rs:131-137
        raise IoError(f"failed to read JP2 input: {e}") from e
    return decode_jp2(data)


def inspect_jp2(data) -> DecodeMetadata:
    """inspect_jp2(&[u8]) (src/decode/mod.rs:38-49)."""
    lib = _lib.load()
    buf = _as_buffer(data)
    m = _lib.MetadataC()
    _check(lib.jp2lam_inspect_jp2(buf, len(buf), C.byref(m)))
    return DecodeMetadata(m.width, m.height, ColorSpace(m.colorspace), m.num_components, m.bits_per_component,
                          bool(m.has_ipr_metadata), m.decomposition_levels, m.code_block_width,
                          m.code_block_height, m.layers, m.transform, bool(m.use_mct), m.guard_bits,
                          m.num_quant_steps, m.tile_part_count, m.first_tile_payload_len)


@dataclass(frozen=True)
class BatchComponentProfile:
    """BatchComponentProfile (src/batch.rs:15-23)."""
    width: int
    height: int
    precision: int
    signed: bool
    dx: int
    dy: int


@dataclass(frozen=True)
class BatchProfile:
    """BatchProfile (src/batch.rs:5-13)."""
    width: int
    height: int
    colorspace: ColorSpace
    components: tuple
    quality: int
    format: OutputFormat


def _read_profile(getter, handle) -> Optional[BatchProfile]:
    p = _lib.BatchProfileC()
    comps = (_lib.ComponentProfileC * 8)()
    rc = getter(handle, C.byref(p), comps, len(comps))
    if rc < 0:
        raise InvalidInput("null batch handle")
    if rc == 0:
        return None
    if p.num_components > len(comps):
        comps = (_lib.ComponentProfileC * p.num_components)()
        getter(handle, C.byref(p), comps, len(comps))
    cs = tuple(BatchComponentProfile(k.width, k.height, k.precision, bool(k.is_signed), k.dx, k.dy)
               for k in comps[:p.num_components])
    return BatchProfile(p.width, p.height, ColorSpace(p.colorspace), cs, p.quality, OutputFormat(p.format))


class BatchEncoder:
    """BatchEncoder (src/batch.rs:32-84): the first image fixes the profile."""

    def __init__(self, options: Optional[EncodeOptions] = None):
        self._lib = _lib.load()
        self.options = options or EncodeOptions()
        o = _opts(self.options)
        self._h = self._lib.jp2lam_batch_encoder_new(C.byref(o))
        if not self._h:
            raise IoError("could not allocate a batch encoder")

    def encode_one(self, image: Image) -> bytes:
        v = _View(image)
        out = _lib.u8p()
        n = C.c_size_t()
        _check(self._lib.jp2lam_batch_encoder_encode_one(self._h, C.byref(v.view), C.byref(out), C.byref(n)))
        return _take_bytes(out, n.value)

    def encode_all(self, images: Iterable[Image]) -> List[bytes]:
        return [self.encode_one(im) for im in images]

    @property
    def encoded_count(self) -> int:
        return int(self._lib.jp2lam_batch_encoder_encoded_count(self._h))

    @property
    def has_profile(self) -> bool:
        return bool(self._lib.jp2lam_batch_encoder_has_profile(self._h))

    def profile(self) -> Optional["BatchProfile"]:
        """BatchEncoder::profile (src/batch.rs:47-49)."""
        return _read_profile(self._lib.jp2lam_batch_encoder_profile, self._h)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.jp2lam_batch_encoder_free(h)


class BatchDecoder:
    """BatchDecoder (src/batch.rs:86-130): decodes, then checks the profile."""

    def __init__(self):
        self._lib = _lib.load()
        self._h = self._lib.jp2lam_batch_decoder_new()
        if not self._h:
            raise IoError("could not allocate a batch decoder")

    def decode_one(self, data) -> Image:
        buf = _as_buffer(data)
        img = _lib.ImageC()
        _check(self._lib.jp2lam_batch_decoder_decode_one(self._h, buf, len(buf), C.byref(img)))
        return _image_from_c(img)

    def decode_all(self, streams: Iterable) -> List[Image]:
        return [self.decode_one(s) for s in streams]

    @property
    def decoded_count(self) -> int:
        return int(self._lib.jp2lam_batch_decoder_decoded_count(self._h))

    def profile(self) -> Optional["BatchProfile"]:
        """BatchDecoder::profile (src/batch.rs:91-93)."""
        return _read_profile(self._lib.jp2lam_batch_decoder_profile, self._h)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.jp2lam_batch_decoder_free(h)


def encode_batch(images: Iterable[Image], options: Optional[EncodeOptions] = None) -> List[bytes]:
    return BatchEncoder(options).encode_all(images)


def decode_batch(streams: Iterable) -> List[Image]:
    return BatchDecoder().decode_all(streams)


# ---- byte-image fast paths ------------------------------------------------------------
def encode_u8(pixels: np.ndarray, quality: int = 85, format: OutputFormat = OutputFormat.JP2) -> bytes:
    """Encode a uint8 (H, W) gray or (H, W, 3) sRGB array."""
    lib = _lib.load()
    px = np.ascontiguousarray(pixels, dtype=np.uint8)
    ch = 1 if px.ndim == 2 else px.shape[2]
    o = _lib.EncodeOptionsC(min(int(quality), 255), int(format))
    out = _lib.u8p()
    n = C.c_size_t()
    _check(lib.jp2lam_encode_u8(px.ctypes.data, px.shape[1], px.shape[0], ch, C.byref(o), C.byref(out), C.byref(n)))
    return _take_bytes(out, n.value)


def decode_u8(data) -> np.ndarray:
    """Decode a JP2 into a uint8 (H, W) or (H, W, 3) array."""
    lib = _lib.load()
    buf = _as_buffer(data)
    meta = inspect_jp2(buf)
    out = np.empty(meta.width * meta.height * meta.num_components, dtype=np.uint8)
    w, h, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
    _check(lib.jp2lam_decode_jp2_u8(buf, len(buf), out.ctypes.data, out.size, C.byref(w), C.byref(h), C.byref(c)))
    return out.reshape(h.value, w.value) if c.value == 1 else out.reshape(h.value, w.value, c.value)


def encode_many(pages: Sequence, format: OutputFormat = OutputFormat.JP2, device: bool = False):
    """Encode ``[(pixels, quality), ...]`` of mixed sizes in one device batch.

    With ``device=True`` each ``pixels`` is ``(device_pointer, H, W, C)``.
    Returns a list of ``bytes`` or :class:`Jp2LamError` per page."""
    lib = _lib.load()
    items = (_lib.EncodeItemC * len(pages))()
    keep = []
    for i, (px, q) in enumerate(pages):
        if device:
            ptr, h, w, ch = px
        else:
            a = np.ascontiguousarray(px, dtype=np.uint8)
            keep.append(a)
            ptr, h, w, ch = a.ctypes.data, a.shape[0], a.shape[1], 1 if a.ndim == 2 else a.shape[2]
        items[i] = _lib.EncodeItemC(ptr, w, h, ch, min(int(q), 255), int(format), 1 if device else 0, 0, None, 0, 0)
    lib.jp2lam_encode_many(items, len(pages))
    out = []
    for it in items:
        if it.status == 0:
            out.append(_take_bytes(it.out, it.out_len))
        else:
            out.append(_ERRORS.get(it.status, IoError)("page failed"))
    return out


def decode_many(streams: Sequence, device_outputs: Optional[Sequence] = None):
    """Decode several JP2 streams in one device batch into uint8 arrays (or
    into ``(device_pointer, capacity)`` targets)."""
    lib = _lib.load()
    bufs = [_as_buffer(s) for s in streams]
    items = (_lib.DecodeItemC * len(bufs))()
    outs = []
    for i, b in enumerate(bufs):
        if device_outputs is not None:
            ptr, cap = device_outputs[i]
            outs.append(None)
            items[i] = _lib.DecodeItemC(C.cast(C.c_char_p(b), C.c_void_p), len(b), ptr, cap, 0, 0, 0, 1)
        else:
            m = inspect_jp2(b)
            a = np.empty(m.width * m.height * m.num_components, dtype=np.uint8)
            outs.append(a)
            items[i] = _lib.DecodeItemC(C.cast(C.c_char_p(b), C.c_void_p), len(b), a.ctypes.data, a.size, 0, 0, 0, 0)
    lib.jp2lam_decode_many(items, len(bufs))
    res = []
    for i, it in enumerate(items):
        if it.status != 0:
            res.append(_ERRORS.get(it.status, IoError)("stream failed"))
        elif outs[i] is None:
            res.append((it.height, it.width, it.channels))
        else:
            a = outs[i]
            res.append(a.reshape(it.height, it.width) if it.channels == 1 else a.reshape(it.height, it.width, it.channels))
    return res


# ---- serving: continuous batching (no counterpart in the reference) ----------
class Served:
    """A finished request: ``value`` is the codestream (encode) or an
    ``(H, W[, C])`` uint8 array (decode), or the :class:`Jp2LamError`.
    Times are milliseconds: queue (submit->start), run (start->done),
    total (submit->done); ``batch_size`` is the batch it ran in."""

    __slots__ = ("value", "queue_ms", "run_ms", "total_ms", "batch_size")

    def __init__(self, value, queue_ms, run_ms, total_ms, batch_size):
        self.value, self.queue_ms, self.run_ms = value, queue_ms, run_ms
        self.total_ms, self.batch_size = total_ms, batch_size


class Server:
    """Request loop with continuous batching over ``slots`` device engines.

    ``submit_encode``/``submit_decode`` return an id at once; ``wait(id)``
    returns a :class:`Served`.  Requests that arrive while batches run are
    admitted as a slot frees, up to ``mem_budget_mib`` of device memory per
    server and ``max_batch`` requests per batch."""

    def __init__(self, slots: int = 2, mem_budget_mib: int = 16384, max_batch: int = 16,
                 simulated_delay_us: Optional[int] = None):
        """``simulated_delay_us``: run the request loop on the simulated
        executor (no device work; for tests of admission, cancellation and
        fault handling) with that delay per batch."""
        self._lib = _lib.load()
        if simulated_delay_us is None:
            self._h = self._lib.jp2lam_server_new(slots, mem_budget_mib, max_batch)
        else:
            self._h = self._lib.jp2lam_server_new_simulated(slots, mem_budget_mib, max_batch, int(simulated_delay_us))
        self._kinds = {}
        if not self._h:
            raise IoError(self._lib.jp2lam_last_error().decode())

    def submit_encode(self, pixels, quality: int = 85, format: OutputFormat = OutputFormat.JP2,
                      device=None) -> int:
        """``pixels``: (H, W) or (H, W, 3) uint8 array, or with ``device``
        = ``(device_pointer, H, W, C)`` for pixels already in HBM."""
        if device is not None:
            ptr, h, w, ch = device
            rid = self._lib.jp2lam_server_submit_encode(self._h, ptr, w, h, ch, min(int(quality), 255), int(format), 1)
        else:
            a = np.ascontiguousarray(pixels, dtype=np.uint8)
            ch = 1 if a.ndim == 2 else a.shape[2]
            rid = self._lib.jp2lam_server_submit_encode(self._h, a.ctypes.data, a.shape[1], a.shape[0], ch,
                                                        min(int(quality), 255), int(format), 0)
        if not rid:
            raise InvalidInput(self._lib.jp2lam_last_error().decode())
        self._kinds[rid] = "encode"
        return rid

    def submit_decode(self, data) -> int:
        b = _as_buffer(data)
        rid = self._lib.jp2lam_server_submit_decode(self._h, C.cast(C.c_char_p(b), C.c_void_p), len(b))
        if not rid:
            raise InvalidInput(self._lib.jp2lam_last_error().decode())
        self._kinds[rid] = "decode"
        return rid

    def wait(self, rid: int, timeout_ms: int = -1) -> Optional[Served]:
        out = _lib.ServedC()
        rc = self._lib.jp2lam_server_wait(self._h, rid, timeout_ms, C.byref(out))
        if rc < 0:
            raise InvalidInput(self._lib.jp2lam_last_error().decode())
        if rc == 0:
            return None
        kind = self._kinds.pop(rid, "encode")
        if out.status != 0:
            value = _ERRORS.get(out.status, IoError)(self._lib.jp2lam_last_error().decode())
        elif kind == "encode":
            value = _take_bytes(out.data, out.len) if out.len else b""
        else:
            arr = np.frombuffer(_take_bytes(out.data, out.len), dtype=np.uint8) if out.len else np.zeros(0, np.uint8)
            value = arr.reshape(out.height, out.width) if out.channels == 1 else \
                arr.reshape(out.height, out.width, out.channels)
        return Served(value, out.queue_ms, out.run_ms, out.total_ms, out.batch_size)

    def cancel(self, rid: int) -> bool:
        """True if the request was still queued and is now cancelled."""
        rc = self._lib.jp2lam_server_cancel(self._h, rid)
        if rc < 0:
            raise InvalidInput(self._lib.jp2lam_last_error().decode())
        return rc == 1

    def faults(self) -> dict:
        v = (C.c_uint64 * 3)()
        self._lib.jp2lam_server_faults(self._h, v)
        return {"rejected": int(v[0]), "cancelled": int(v[1]), "device_faults": int(v[2])}

    def stats(self) -> dict:
        v = (C.c_uint64 * 4)()
        self._lib.jp2lam_server_stats(self._h, v)
        return {"submitted": v[0], "completed": v[1], "batches": v[2], "largest_batch": v[3]}

    def close(self):
        if self._h:
            self._lib.jp2lam_server_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
