"""Deterministic synthetic pages for the benchmark workloads.

Each generator returns a C-contiguous ``uint8`` array, ``(H, W)`` for
grayscale or ``(H, W, 3)`` for sRGB, from a seeded ``numpy`` PCG64 stream.

* :func:`gradient_page` -- 512x512 gray: horizontal gradient 40..220 plus
  N(0, 8) noise and 40 dark rectangles (value 0..20, 4-30 px sides).
* :func:`document_page` -- gray document: background 245 + N(0, 3), dark
  glyph rectangles (10..60, 3-20 px tall, 2-14 px wide) on text lines every
  40 px until 12 % of the area is covered.
* :func:`photo_page` -- sRGB: per channel a sum of 8 oriented sinusoids
  (period 16-800 px, amplitude 10-40) plus 1/f noise (sigma 12) plus an
  offset 60..190.
* :func:`mixed_pages` -- the mixed-size queue: 70 % documents, 30 % photos,
  width 1200..5100, height 1600..6600 (even), quality from {50,75,85,100}.
"""
from __future__ import annotations

import numpy as np

__all__ = ["gradient_page", "document_page", "photo_page", "mixed_pages", "mixed_pages_parallel", "page_spec"]


def _clamp_u8(a: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(a), 0, 255).astype(np.uint8)


def gradient_page(width: int = 512, height: int = 512, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = np.arange(width, dtype=np.float64)
    base = 40.0 + 180.0 * x / max(width - 1, 1)
    img = base[None, :] + rng.normal(0.0, 8.0, size=(height, width))
    img = _clamp_u8(img)
    for _ in range(40):
        rw = int(rng.integers(4, 31))
        rh = int(rng.integers(4, 31))
        x0 = int(rng.integers(0, max(width - rw, 1)))
        y0 = int(rng.integers(0, max(height - rh, 1)))
        img[y0:y0 + rh, x0:x0 + rw] = int(rng.integers(0, 21))
    return np.ascontiguousarray(img)


def document_page(width: int = 2550, height: int = 3300, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    img = _clamp_u8(245.0 + rng.normal(0.0, 3.0, size=(height, width)))
    target = 0.12 * width * height
    covered = 0
    lines = list(range(20, max(height - 20, 21), 40))
    if not lines:
        lines = [0]
    # glyphs are placed left to right along each line; lines are revisited
    # until the coverage target is met
    cursor = {y: int(rng.integers(0, 40)) for y in lines}
    li = 0
    guard = 0
    while covered < target and guard < 10_000_000:
        guard += 1
        y = lines[li % len(lines)]
        li += 1
        gh = int(rng.integers(3, 21))
        gw = int(rng.integers(2, 15))
        x0 = cursor[y]
        if x0 + gw >= width:
            cursor[y] = int(rng.integers(0, 8))
            continue
        y0 = min(max(y - gh, 0), height - gh)
        val = int(rng.integers(10, 61))
        img[y0:y0 + gh, x0:x0 + gw] = val
        covered += gh * gw
        cursor[y] = x0 + gw + int(rng.integers(1, 6))
    return np.ascontiguousarray(img)


def _pink_noise(rng: np.random.Generator, n: int, h: int, w: int, sigma: float) -> np.ndarray:
    """n independent (h, w) fields of 1/f noise with standard deviation sigma."""
    try:  # multithreaded FFTs when scipy is present; numpy's otherwise
        from scipy import fft as _fft

        kw = {"workers": -1}
    except ImportError:  # pragma: no cover
        _fft = np.fft
        kw = {}
    white = rng.standard_normal(size=(n, h, w), dtype=np.float32)
    f = _fft.rfft2(white, axes=(1, 2), **kw)
    fy = np.fft.fftfreq(h).astype(np.float32)[:, None]
    fx = np.fft.rfftfreq(w).astype(np.float32)[None, :]
    r = np.sqrt(fx * fx + fy * fy)
    r[0, 0] = 1.0
    f *= (1.0 / r)[None]
    f[:, 0, 0] = 0.0
    out = _fft.irfft2(f, s=(h, w), axes=(1, 2), **kw)
    for c in range(n):
        std = float(out[c].std())
        if std > 0:
            out[c] *= np.float32(sigma / std)
    return out


def photo_page(width: int = 4000, height: int = 3000, seed: int = 3) -> np.ndarray:
    rng = np.random.default_rng(seed)
    x = np.arange(width, dtype=np.float64)
    y = np.arange(height, dtype=np.float64)
    noise = _pink_noise(rng, 3, height, width, 12.0)
    out = np.empty((height, width, 3), dtype=np.uint8)
    for c in range(3):
        # sum of 8 oriented sinusoids amp * sin(k (x cos t + y sin t) + phase),
        # split by angle addition into one (H x 16) @ (16 x W) product
        rows = np.empty((height, 16), dtype=np.float64)
        cols = np.empty((16, width), dtype=np.float64)
        for i in range(8):
            period = float(rng.uniform(16.0, 800.0))
            amp = float(rng.uniform(10.0, 40.0))
            theta = float(rng.uniform(0.0, np.pi))
            phase = float(rng.uniform(0.0, 2.0 * np.pi))
            k = 2.0 * np.pi / period
            ax = k * np.cos(theta) * x
            by = k * np.sin(theta) * y + phase
            rows[:, 2 * i], rows[:, 2 * i + 1] = np.cos(by), np.sin(by)
            cols[2 * i], cols[2 * i + 1] = amp * np.sin(ax), amp * np.cos(ax)
        acc = (rows @ cols).astype(np.float32)
        acc += noise[c]
        acc += np.float32(rng.uniform(60.0, 190.0))
        out[:, :, c] = _clamp_u8(acc)
    return np.ascontiguousarray(out)


def page_spec(index: int, seed: int = 4) -> dict:
    """Size, kind and quality of page ``index`` of the mixed queue."""
    rng = np.random.default_rng(seed * 1000 + index)
    w = int(rng.integers(600, 2551)) * 2
    h = int(rng.integers(800, 3301)) * 2
    kind = "gray" if rng.uniform() < 0.7 else "rgb"
    q = int(rng.choice([50, 75, 85, 100]))
    return {"width": w, "height": h, "kind": kind, "quality": q, "seed": seed * 1000 + index}


def mixed_pages(count: int, seed: int = 4, scale: float = 1.0, start: int = 0):
    """Yield ``(pixels, quality)`` for pages ``start .. start + count - 1``
    of the mixed-size page queue.  ``scale`` shrinks both dimensions (tests
    use small pages)."""
    for i in range(start, start + count):
        s = page_spec(i, seed)
        w = max(2, int(s["width"] * scale) // 2 * 2)
        h = max(2, int(s["height"] * scale) // 2 * 2)
        if s["kind"] == "gray":
            px = document_page(w, h, s["seed"])
        else:
            px = photo_page(w, h, s["seed"])
        yield px, s["quality"]


def _mixed_page(args):
    index, seed, scale = args
    return next(mixed_pages(1, seed, scale, start=index))


def mixed_pages_parallel(count: int, seed: int = 4, scale: float = 1.0, workers: int = 8):
    """``list(mixed_pages(count, seed, scale))``, generated by ``workers``
    processes (started fresh, so this is safe after the GPU is in use)."""
    if workers <= 1 or count <= 1:
        return list(mixed_pages(count, seed, scale))
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor

    with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_mixed_page, [(i, seed, scale) for i in range(count)], chunksize=1))

