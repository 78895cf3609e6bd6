"""Host parsers under AddressSanitizer/UBSan, on the CPU.

The decode path parses caller bytes on the host (JP2 boxes, codestream
markers, packet headers) before anything reaches the device.  These tests
build those parsers (jp2lam_amd/csrc/host.cpp) with g++ and the sanitizers,
then run a mutation fuzz over streams the oracle encodes, plus known
malformed inputs through the C ABI.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from jp2lam_amd import synth

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "jp2lam_amd" / "csrc"
FUZZ_SRC = ROOT / "tests" / "fuzz" / "parse_fuzz.cpp"


@pytest.fixture(scope="module")
def fuzz_bin(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("fuzz") / "parse_fuzz"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", str(FUZZ_SRC), str(CSRC / "host.cpp"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return out


def _seed_streams(ora, tmp: Path):
    rng = np.random.default_rng(7)
    paths = []
    cases = [
        (synth.document_page(96, 80, seed=11), 100),
        (synth.document_page(120, 72, seed=12), 85),
        (synth.document_page(33, 47, seed=13), 20),
        (rng.integers(0, 256, (40, 52, 3), dtype=np.uint8), 75),
        (rng.integers(0, 256, (24, 18, 3), dtype=np.uint8), 99),
        (rng.integers(0, 256, (17, 9), dtype=np.uint8), 100),
        (synth.document_page(200, 150, seed=14), 50),
    ]
    for k, (px, q) in enumerate(cases):
        p = tmp / f"seed{k}.jp2"
        p.write_bytes(ora.encode(px, q))
        paths.append(p)
    return paths


def test_parser_mutation_fuzz_asan(fuzz_bin, ora, tmp_path):
    seeds = _seed_streams(ora, tmp_path)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(fuzz_bin), "30000", "20261204", *map(str, seeds)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "fuzz ok" in r.stdout
    # the mutants must exercise both outcomes
    parsed = int(r.stdout.split(",")[1].split()[0])
    assert parsed > 0


def _jp2_around(codestream: bytes, width: int, height: int) -> bytes:
    def box(t: bytes, body: bytes) -> bytes:
        return (8 + len(body)).to_bytes(4, "big") + t + body

    ihdr = height.to_bytes(4, "big") + width.to_bytes(4, "big") + (1).to_bytes(2, "big") + bytes([7, 7, 0, 0])
    colr = bytes([1, 0, 0]) + (17).to_bytes(4, "big")
    return (box(b"jP  ", bytes([0x0D, 0x0A, 0x87, 0x0A])) + box(b"ftyp", b"jp2 " + bytes(4) + b"jp2 ")
            + box(b"jp2h", box(b"ihdr", ihdr) + box(b"colr", colr)) + box(b"jp2c", codestream))


def test_tile_part_comment_past_end_is_rejected(lib, ora):
    """A COM segment in the tile-part header whose length runs past the end
    of the codestream is an error, not a read past the buffer."""
    good = ora.encode(synth.document_page(64, 64, seed=3), 85)
    meta = _find_codestream(good)
    cs = good[meta:]
    # main header up to the first SOT
    sot = cs.index(b"\xff\x90")
    for com_len in (0xFFF0, 0x0001, 0x0000):
        bad = cs[:sot] + b"\xff\x90\x00\x0a\x00\x00\x00\x00\x00\x00\x00\x01" + b"\xff\x64" + com_len.to_bytes(2, "big")
        data = _jp2_around(bad, 64, 64)
        from jp2lam_amd import _lib

        m = _lib.MetadataC()
        buf = C.create_string_buffer(data, len(data))
        st = lib.jp2lam_inspect_jp2(buf, len(data), C.byref(m))
        assert st == 3, (com_len, st)
        msg = lib.jp2lam_last_error().decode()
        assert "marker" in msg, msg


def _find_codestream(jp2: bytes) -> int:
    """Offset of the codestream inside a JP2 (after the jp2c box header)."""
    pos = 0
    while pos + 8 <= len(jp2):
        n = int.from_bytes(jp2[pos:pos + 4], "big")
        t = jp2[pos + 4:pos + 8]
        if t == b"jp2c":
            return pos + 8
        pos += n
    raise AssertionError("no jp2c box")
