"""Launch sequences replayed as HIP graphs code the data of the call that
replays them.

A batch shape seen twice is captured and later calls of that shape launch
the graph; the descriptor tables and pixels are copied into the arena before
each launch.  These tests run one shape several times with different
content, with eager work of other shapes (and their synchronisations)
between the replays, and check every result against the oracle and against
the same call with graphs switched off (the default)."""
import ctypes as C

import numpy as np
import pytest

import jp2lam_amd as j
from jp2lam_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def graphs_on(gpu):
    gpu.jp2g_set_graphs(1)
    yield
    gpu.jp2g_set_graphs(0)


def _stats(lib):
    cap, lau = C.c_uint64(0), C.c_uint64(0)
    lib.jp2g_graph_stats(C.byref(cap), C.byref(lau))
    return cap.value, lau.value


def test_encode_replays_code_new_content(gpu, lib, ora):
    pages = [synth.photo_page(480, 352, seed=s) for s in (11, 12, 13, 14, 15)]
    other = synth.document_page(300, 410, seed=7)
    c0, l0 = _stats(lib)
    got = []
    for k, px in enumerate(pages):
        got.append(j.encode_u8(px, 75))
        # eager work of another shape between the replays
        if k % 2 == 0:
            assert j.encode_u8(other, 50) == ora.encode(other, 50)
    c1, l1 = _stats(lib)
    assert c1 - c0 >= 1 and l1 - l0 >= len(pages) - 1
    for k, (px, s) in enumerate(zip(pages, got)):
        assert s == ora.encode(px, 75), k
    lib.jp2g_set_graphs(0)
    try:
        assert [j.encode_u8(px, 75) for px in pages] == got
    finally:
        lib.jp2g_set_graphs(1)


def test_batched_encode_replays_match_oracle(gpu, lib, ora):
    """A batch of two lossy pages and a lossless one, replayed with new
    content."""
    for s in range(4):
        items = [(synth.document_page(520, 700, seed=20 + s), 85),
                 (synth.photo_page(333, 257, seed=30 + s), 50),
                 (synth.gradient_page(200, 120, seed=40 + s), 100)]
        got = j.encode_many(items)
        for k, ((px, q), g) in enumerate(zip(items, got)):
            assert g == ora.encode(px, q), (s, k)


def test_decode_replays_code_new_content(gpu, lib, ora):
    streams = [ora.encode(synth.photo_page(480, 352, seed=s), 75) for s in (51, 52, 53, 54)]
    other = ora.encode(synth.document_page(256, 320, seed=9), 85)
    c0, l0 = _stats(lib)
    got = []
    for k, s in enumerate(streams * 2):
        got.append(j.decode_u8(s))
        if k % 3 == 1:
            assert np.array_equal(j.decode_u8(other), ora.decode(other))
    c1, l1 = _stats(lib)
    assert c1 - c0 >= 1 and l1 - l0 >= 1
    for k, g in enumerate(got):
        assert np.array_equal(g, ora.decode(streams[k % len(streams)])), k
