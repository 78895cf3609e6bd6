"""ctypes binding of the oracle (oracle/build/liboracle.so).

Test infrastructure only: the oracle is the CPU restatement of the
reference codec that the parity tests and the CPU baseline compare against.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "build" / "liboracle.so"

vp = C.c_void_p


class RdPoint(C.Structure):
    """ora_rd_point: one point of a PCRD curve (pcrd.rs PcrdPoint)."""
    _fields_ = [("passes", C.c_int), ("bytes", C.c_uint32), ("dist", C.c_double), ("slope", C.c_double)]


class RawRecord(C.Structure):
    """ora_raw_record: RawPassRecord (pcrd.rs:33-50)."""
    _fields_ = [("pass_index", C.c_int), ("bytes", C.c_uint32), ("cumulative_bytes", C.c_uint32),
                ("distortion_delta", C.c_double)]


def build_oracle() -> Path:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)
    return ORACLE_LIB


class Oracle:
    def __init__(self, path: Path = ORACLE_LIB):
        if not path.exists():
            build_oracle()
        L = C.CDLL(str(path))
        sig = {
            "ora_encode_u8": (vp, [vp, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "ora_encode_planes": (vp, [vp, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)]),
            "ora_enc_release": (None, [vp]),
            "ora_enc_error": (C.c_char_p, [vp]),
            "ora_enc_file": (C.c_size_t, [vp, C.POINTER(vp)]),
            "ora_enc_codestream": (C.c_size_t, [vp, C.POINTER(vp)]),
            "ora_enc_nblocks": (C.c_int, [vp]),
            "ora_enc_lambda": (C.c_double, [vp, C.c_int]),
            "ora_enc_plan": (None, [vp, vp]),
            "ora_enc_block_info": (None, [vp, C.c_int, vp]),
            "ora_enc_block_passes": (None, [vp, C.c_int, vp]),
            "ora_enc_block_bytes": (None, [vp, C.c_int, vp]),
            "ora_coefficients_u8": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_int, C.c_int, vp]),
            "ora_decode_u8": (C.c_int, [vp, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_int), vp, C.c_size_t, C.c_char_p]),
            "ora_t1_encode_flat": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_int]),
            "ora_t1_decode_block": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                                              C.c_size_t, vp, C.c_char_p]),
            "ora_mq_encode_stream": (C.c_int, [vp, vp, C.c_int, vp, C.c_int]),
            "ora_mq_decode_stream": (None, [vp, C.c_int, vp, C.c_int, vp]),
            "ora_tables": (None, [vp] * 7),
            "ora_fwd53": (None, [vp, C.c_size_t, C.c_size_t, C.c_int]),
            "ora_inv53": (None, [vp, C.c_size_t, C.c_size_t, C.c_int]),
            "ora_fwd97": (None, [vp, C.c_size_t, C.c_size_t, C.c_int]),
            "ora_inv97": (None, [vp, C.c_size_t, C.c_size_t, C.c_int]),
            "ora_fwd97_f64": (None, [vp, C.c_size_t, C.c_size_t, C.c_int]),
            "ora_quantize": (C.c_int32, [C.c_float, C.c_float]),
            "ora_dequantize": (C.c_float, [C.c_int32, C.c_float]),
            "ora_inverse_ict": (None, [vp, vp, vp, C.c_size_t]),
            "ora_inverse_rct": (None, [vp, vp, vp, C.c_size_t]),
            "ora_forward_ict": (None, [vp, vp, vp, C.c_size_t, C.c_int, vp]),
            "ora_forward_rct": (None, [vp, vp, vp, C.c_size_t, C.c_int, vp]),
            "ora_encode_stepsize": (None, [C.c_int32, C.c_uint32, C.POINTER(C.c_uint8), C.POINTER(C.c_uint16)]),
            "ora_irreversible_expounded_quant": (None, [C.c_uint32, C.c_uint8, C.c_uint8, C.c_int,
                                                        C.POINTER(C.c_uint8), C.POINTER(C.c_uint16)]),
            "ora_get_norm_97": (C.c_double, [C.c_uint32, C.c_int]),
            "ora_max_decompositions": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "ora_encode_metrics_u8": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_int, C.c_int, vp]),
            "ora_metric_recon_u8": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_int, C.c_int, vp]),
            "ora_metrics_flat": (None, [vp, vp, C.c_uint32, C.c_uint32, C.c_int, vp]),
            "ora_decode_recon_f32": (C.c_int, [vp, C.c_size_t, vp, C.c_size_t]),
            "ora_analyze_content_flat": (None, [vp, C.c_size_t, C.c_int, C.c_int, vp]),
            "ora_classify_content": (C.c_int, [vp]),
            "ora_pcrd_raw_curve": (C.c_int, [vp, C.c_int, vp]),
            "ora_pcrd_prune_hull": (C.c_int, [vp, C.c_int, vp, C.POINTER(C.c_int)]),
            "ora_pcrd_choose_point": (C.c_int, [vp, C.c_int, C.c_double, C.POINTER(C.c_int)]),
            "ora_pass_delta_pass_kind_aware": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                                            C.c_double, C.c_int, C.c_int, C.c_int]),
            "ora_pass_delta_baseline": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]),
            "ora_quality_to_lambda": (C.c_double, [C.c_int, C.c_uint32]),
            "ora_calibrate_lambda": (C.c_double, [C.c_double, C.c_double, C.c_int, C.c_uint32]),
            "ora_quality_step_scaler": (C.c_double, [C.c_int]),
            "ora_scale_one_quant": (None, [C.c_uint8, C.c_uint16, C.c_double, C.POINTER(C.c_uint8),
                                           C.POINTER(C.c_uint16)]),
            "ora_enc_block_rd": (C.c_int, [vp, C.c_int, vp, vp, vp, C.POINTER(C.c_int32)]),
            "ora_enc_max_slope": (C.c_double, [vp, C.c_int]),
            "ora_enc_counters": (None, [vp, vp]),
            "ora_set_faithful": (None, [C.c_int]),
            "ora_ycbcr_to_rgb": (None, [vp, C.c_size_t, vp]),
            "ora_set_threads": (None, [C.c_int]),
            "ora_enc_block_nsym": (C.c_uint32, [vp, C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L

    # ---- whole codec ----
    def encode(self, px: np.ndarray, quality: int, fmt: int = 0, keep: bool = False):
        px = np.ascontiguousarray(px, dtype=np.uint8)
        h, w = px.shape[:2]
        ch = 1 if px.ndim == 2 else px.shape[2]
        st = C.c_int()
        e = self.L.ora_encode_u8(px.ctypes.data, w, h, ch, quality, fmt, C.byref(st))
        if st.value:
            msg = self.L.ora_enc_error(e).decode()
            self.L.ora_enc_release(e)
            raise RuntimeError(f"oracle encode failed ({st.value}): {msg}")
        p = vp()
        n = self.L.ora_enc_file(e, C.byref(p))
        data = C.string_at(p, n)
        if keep:
            return data, Encoding(self, e)
        self.L.ora_enc_release(e)
        return data

    def decode(self, data: bytes) -> np.ndarray:
        w, h, nc = C.c_uint32(), C.c_uint32(), C.c_int()
        err = C.create_string_buffer(256)
        cap = 1 << 20
        while True:
            out = np.empty(cap, dtype=np.uint8)
            rc = self.L.ora_decode_u8(data, len(data), C.byref(w), C.byref(h), C.byref(nc), out.ctypes.data, cap, err)
            if rc == 5 and b"too small" in err.value:
                cap = w.value * h.value * max(nc.value, 1) + 16
                continue
            if rc:
                raise OracleError(rc, err.value.decode())
            n = w.value * h.value * nc.value
            a = out[:n].copy()
            return a.reshape(h.value, w.value) if nc.value == 1 else a.reshape(h.value, w.value, nc.value)

    # ---- encode_with_psnr metrics ----
    def encode_metrics(self, px: np.ndarray, quality: int):
        """(psnr_db, ssim) of encode_with_psnr on u8 pixels."""
        px = np.ascontiguousarray(px, dtype=np.uint8)
        h, w = px.shape[:2]
        ch = 1 if px.ndim == 2 else px.shape[2]
        out = np.zeros(2, dtype=np.float64)
        rc = self.L.ora_encode_metrics_u8(px.ctypes.data, w, h, ch, quality, out.ctypes.data)
        if rc:
            raise OracleError(rc, "metrics")
        return float(out[0]), float(out[1])

    def metric_recon(self, px: np.ndarray, quality: int) -> np.ndarray:
        """The truncated f32 reconstruction the metrics compare (ch, h, w)."""
        px = np.ascontiguousarray(px, dtype=np.uint8)
        h, w = px.shape[:2]
        ch = 1 if px.ndim == 2 else px.shape[2]
        out = np.zeros((ch, h, w), dtype=np.float32)
        assert self.L.ora_metric_recon_u8(px.ctypes.data, w, h, ch, quality, out.ctypes.data) == 0
        return out

    def metrics_of(self, orig: np.ndarray, recon: np.ndarray):
        """(psnr_db, ssim) of recon (ch, h, w, level-shifted f32) against
        orig (ch, h, w) samples."""
        orig = np.ascontiguousarray(orig, dtype=np.int32)
        recon = np.ascontiguousarray(recon, dtype=np.float32)
        ch, h, w = orig.shape
        out = np.zeros(2, dtype=np.float64)
        self.L.ora_metrics_flat(orig.ctypes.data, recon.ctypes.data, w, h, ch, out.ctypes.data)
        return float(out[0]), float(out[1])

    def decode_recon(self, data: bytes, shape) -> np.ndarray:
        """The decoder's f32 planes after the inverse 9/7 (ch, h, w)."""
        out = np.zeros(shape, dtype=np.float32)
        rc = self.L.ora_decode_recon_f32(data, len(data), out.ctypes.data, out.size)
        if rc:
            raise OracleError(rc, "decode_recon")
        return out

    # ---- content analysis ----
    def analyze_content(self, planes: np.ndarray, colorspace: int):
        """(bright_share, entropy, colorfulness) as f32 and the class
        (1 graphics, 0 photo) of planes (ncomp, h, w) i32."""
        planes = np.ascontiguousarray(planes, dtype=np.int32)
        nc = planes.shape[0]
        out = np.zeros(3, dtype=np.float32)
        self.L.ora_analyze_content_flat(planes.ctypes.data, planes[0].size, nc, colorspace, out.ctypes.data)
        return out, int(self.L.ora_classify_content(out.ctypes.data))

    def coefficients(self, px: np.ndarray, quality: int) -> np.ndarray:
        px = np.ascontiguousarray(px, dtype=np.uint8)
        h, w = px.shape[:2]
        ch = 1 if px.ndim == 2 else px.shape[2]
        out = np.zeros((ch, h, w), dtype=np.int32)
        assert self.L.ora_coefficients_u8(px.ctypes.data, w, h, ch, quality, out.ctypes.data) == 0
        return out

    # ---- stages ----
    def fwd97(self, plane: np.ndarray, levels: int) -> np.ndarray:
        a = np.ascontiguousarray(plane, dtype=np.float32).copy()
        self.L.ora_fwd97(a.ctypes.data, a.shape[1], a.shape[0], levels)
        return a

    def fwd97_f64(self, plane: np.ndarray, levels: int) -> np.ndarray:
        a = np.ascontiguousarray(plane, dtype=np.float64).copy()
        self.L.ora_fwd97_f64(a.ctypes.data, a.shape[1], a.shape[0], levels)
        return a

    def inv97(self, plane: np.ndarray, levels: int) -> np.ndarray:
        a = np.ascontiguousarray(plane, dtype=np.float32).copy()
        self.L.ora_inv97(a.ctypes.data, a.shape[1], a.shape[0], levels)
        return a

    def fwd53(self, plane: np.ndarray, levels: int) -> np.ndarray:
        a = np.ascontiguousarray(plane, dtype=np.int32).copy()
        self.L.ora_fwd53(a.ctypes.data, a.shape[1], a.shape[0], levels)
        return a

    def inv53(self, plane: np.ndarray, levels: int) -> np.ndarray:
        a = np.ascontiguousarray(plane, dtype=np.int32).copy()
        self.L.ora_inv53(a.ctypes.data, a.shape[1], a.shape[0], levels)
        return a

    def t1_encode(self, coeffs: np.ndarray, band: int, mb: int):
        c = np.ascontiguousarray(coeffs, dtype=np.int32)
        h, w = c.shape
        info = np.zeros(6, dtype=np.int32)  # mbp zbp cls npasses nbytes mq_symbols
        passes = np.zeros(5 * 96, dtype=np.int32)
        cap = w * h * 4 + 1024
        out = np.zeros(cap, dtype=np.uint8)
        rc = self.L.ora_t1_encode_flat(c.ctypes.data, w, h, band, mb, info.ctypes.data, passes.ctypes.data,
                                       out.ctypes.data, cap)
        assert rc == 0
        n = int(info[3])
        return info, passes[:5 * n].reshape(n, 5), bytes(out[:info[4]])

    def t1_decode(self, w, h, band, mb, zbp, npasses, data: bytes) -> np.ndarray:
        out = np.zeros((h, w), dtype=np.int32)
        err = C.create_string_buffer(256)
        rc = self.L.ora_t1_decode_block(w, h, band, mb, zbp, npasses, data, len(data), out.ctypes.data, err)
        if rc:
            raise OracleError(rc, err.value.decode())
        return out

    def mq_encode(self, ctx, bits) -> bytes:
        c = np.ascontiguousarray(ctx, dtype=np.uint8)
        b = np.ascontiguousarray(bits, dtype=np.uint8)
        cap = 2 * len(c) + 64
        out = np.zeros(cap, dtype=np.uint8)
        n = self.L.ora_mq_encode_stream(c.ctypes.data, b.ctypes.data, len(c), out.ctypes.data, cap)
        assert n >= 0
        return bytes(out[:n])

    def mq_decode(self, data: bytes, ctx) -> np.ndarray:
        c = np.ascontiguousarray(ctx, dtype=np.uint8)
        out = np.zeros(len(c), dtype=np.uint8)
        self.L.ora_mq_decode_stream(data, len(data), c.ctypes.data, len(c), out.ctypes.data)
        return out

    # ---- PCRD (dwt/pcrd.rs) ------------------------------------------------
    def raw_curve(self, records):
        """build_raw_curve: records are (pass_index, bytes, cumulative, dD).
        Returns (status, [RdPoint])."""
        recs = (RawRecord * max(len(records), 1))(*[RawRecord(*r) for r in records])
        out = (RdPoint * (len(records) + 1))()
        rc = self.L.ora_pcrd_raw_curve(recs, len(records), out)
        return rc, list(out) if rc == 0 else None

    def hull(self, points):
        arr = (RdPoint * len(points))(*points)
        out = (RdPoint * len(points))()
        m = C.c_int()
        rc = self.L.ora_pcrd_prune_hull(arr, len(points), out, C.byref(m))
        assert rc == 0
        return list(out)[:m.value]

    def choose(self, points, lam):
        """choose_block_point: (status, chosen RdPoint)."""
        arr = (RdPoint * len(points))(*points)
        k = C.c_int(-1)
        rc = self.L.ora_pcrd_choose_point(arr, len(points), lam, C.byref(k))
        return rc, (points[k.value] if rc == 0 else None)

    def hull_curve(self, records):
        rc, raw = self.raw_curve(records)
        assert rc == 0, rc
        return self.hull(raw)

    def tables(self):
        qe = np.zeros(94, np.uint16)
        mps, nmps, nlps = (np.zeros(94, np.uint8) for _ in range(3))
        zc = np.zeros(2048, np.uint8)
        sc, spb = np.zeros(256, np.uint8), np.zeros(256, np.uint8)
        self.L.ora_tables(*(a.ctypes.data for a in (qe, mps, nmps, nlps, zc, sc, spb)))
        return qe, mps, nmps, nlps, zc, sc, spb


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(f"{code}: {msg}")
        self.code = code
        self.msg = msg


class Encoding:
    """Per-block view of an oracle encode."""

    def __init__(self, ora: Oracle, handle):
        self.o = ora
        self.h = handle

    def __del__(self):
        if self.h:
            self.o.L.ora_enc_release(self.h)
            self.h = None

    @property
    def nblocks(self) -> int:
        return self.o.L.ora_enc_nblocks(self.h)

    def info(self, i) -> np.ndarray:
        a = np.zeros(16, np.int32)
        self.o.L.ora_enc_block_info(self.h, i, a.ctypes.data)
        return a

    def passes(self, i) -> np.ndarray:
        n = int(self.info(i)[11])
        a = np.zeros(5 * max(n, 1), np.int32)
        self.o.L.ora_enc_block_passes(self.h, i, a.ctypes.data)
        return a[:5 * n].reshape(n, 5)

    def bytes(self, i) -> bytes:
        n = int(self.info(i)[12])
        a = np.zeros(max(n, 1), np.uint8)
        self.o.L.ora_enc_block_bytes(self.h, i, a.ctypes.data)
        return bytes(a[:n])

    def lam(self, c) -> float:
        return self.o.L.ora_enc_lambda(self.h, c)

    def max_slope(self, c) -> float:
        return self.o.L.ora_enc_max_slope(self.h, c)

    def rd(self, i):
        """(raw bytes, raw cumulative distortion, hull pass counts) of block
        i's PCRD curve, origin first."""
        rb = np.zeros(97, np.uint32)
        rdist = np.zeros(97, np.float64)
        hp = np.zeros(97, np.int32)
        hn = C.c_int32()
        n = self.o.L.ora_enc_block_rd(self.h, i, rb.ctypes.data, rdist.ctypes.data, hp.ctypes.data, C.byref(hn))
        return rb[:n], rdist[:n], hp[:hn.value]

    def counters(self):
        """encode/counters.rs totals of this encode: blocks, empty blocks, MQ
        symbols, cleanup / SP / MR passes, pass bytes."""
        a = np.zeros(7, np.uint64)
        self.o.L.ora_enc_counters(self.h, a.ctypes.data)
        return [int(v) for v in a]

    def nsym(self, i) -> int:
        return int(self.o.L.ora_enc_block_nsym(self.h, i))

