"""Serving loop (continuous batching): every result equals the same request
run alone, whatever batch it landed in; failures stay with their request."""
import threading
import time

import numpy as np
import pytest

import jp2lam_amd as j
from jp2lam_amd import synth

pytestmark = pytest.mark.gpu


def _requests():
    reqs = []
    for i, (px, q) in enumerate(synth.mixed_pages(10, seed=11, scale=0.1)):
        reqs.append(("enc", px, q))
    for i in range(6):
        px = synth.document_page(300 + 37 * i, 200 + 23 * i, seed=40 + i) if i % 2 == 0 else \
            synth.photo_page(150 + 11 * i, 120 + 7 * i, seed=40 + i)
        q = [85, 100, 60, 75, 99, 30][i]
        reqs.append(("dec", j.encode_u8(px, q), None))
    return reqs


def _alone(kind, a, q):
    if kind == "enc":
        return j.encode_u8(a, q)
    try:
        return j.decode_u8(a)
    except j.Jp2LamError as e:
        return e


def _same(got, want):
    if isinstance(want, Exception):
        return isinstance(got, j.Jp2LamError) and type(got) is type(want)
    if isinstance(want, bytes):
        return got == want
    return isinstance(got, np.ndarray) and got.shape == want.shape and np.array_equal(got, want)


def test_mixed_stream_equals_alone(gpu):
    reqs = _requests()
    alone = [_alone(k, a, q) for k, a, q in reqs]
    with j.Server(slots=2, mem_budget_mib=4096, max_batch=8) as srv:
        ids = []
        for k, a, q in reqs:
            ids.append(srv.submit_encode(a, q) if k == "enc" else srv.submit_decode(a))
        got = [srv.wait(i, 60000) for i in ids]
        st = srv.stats()
    assert all(g is not None for g in got)
    for n, (g, want) in enumerate(zip(got, alone)):
        assert _same(g.value, want), n
    assert st["completed"] == len(reqs)
    assert st["batches"] < len(reqs), st  # requests shared batches
    assert max(g.batch_size for g in got) > 1


def test_concurrent_submitters(gpu):
    """Requests from several threads while batches run: each result is the
    request's own."""
    pages = [(synth.document_page(256 + 16 * i, 192 + 8 * i, seed=60 + i), [85, 100, 50][i % 3]) for i in range(9)]
    want = [j.encode_u8(px, q) for px, q in pages]
    results = [None] * len(pages)
    with j.Server(slots=2, mem_budget_mib=2048, max_batch=4) as srv:
        def run(idx):
            px, q = pages[idx]
            time.sleep(0.002 * idx)
            results[idx] = srv.wait(srv.submit_encode(px, q), 60000)

        th = [threading.Thread(target=run, args=(i,)) for i in range(len(pages))]
        for t in th:
            t.start()
        for t in th:
            t.join()
    for i, r in enumerate(results):
        assert r is not None and r.value == want[i], i
        assert r.total_ms >= r.run_ms >= 0


def test_bad_request_fails_alone(gpu):
    good = synth.document_page(200, 150, seed=3)
    data = j.encode_u8(good, 85)
    with j.Server(slots=1, mem_budget_mib=1024, max_batch=8) as srv:
        a = srv.submit_encode(good, 85)
        b = srv.submit_decode(b"\x00\x00\x00\x0cjP  \r\n\x87\n" + b"\x00" * 20)
        c = srv.submit_decode(data)
        ra, rb, rc = srv.wait(a), srv.wait(b), srv.wait(c)
    assert ra.value == data
    assert isinstance(rb.value, j.Jp2LamError)
    assert np.array_equal(rc.value, j.decode_u8(data))


def test_device_resident_input(gpu, lib):
    px = synth.document_page(320, 240, seed=9)
    d = lib.jp2lam_device_alloc(px.size)
    assert d
    try:
        assert lib.jp2lam_device_copy(d, px.ctypes.data, px.size, 0) == 0
        with j.Server(slots=1, mem_budget_mib=1024, max_batch=4) as srv:
            r = srv.wait(srv.submit_encode(None, 85, device=(d, 240, 320, 1)))
        assert r.value == j.encode_u8(px, 85)
        # device input through the batch API, device output through decode_many
        out = lib.jp2lam_device_alloc(px.size)
        try:
            data = j.encode_many([((d, 240, 320, 1), 85)], device=True)[0]
            assert data == r.value
            res = j.decode_many([data], device_outputs=[(out, px.size)])[0]
            assert not isinstance(res, Exception)
            back = np.empty_like(px)
            assert lib.jp2lam_device_copy(back.ctypes.data, out, px.size, 1) == 0
            assert np.array_equal(back, j.decode_u8(data))
        finally:
            lib.jp2lam_device_free(out)
    finally:
        lib.jp2lam_device_free(d)
