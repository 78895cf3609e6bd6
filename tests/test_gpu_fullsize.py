"""Full-size configurations on the device, against the oracle:

- the mixed-size page queue (pages drawn as in the 256-page workload: widths
  1200..5100, heights 1600..6600, 70 % documents / 30 % photos, q in
  {50, 75, 85, 100}, seed 4), encoded and decoded as one queue call;
- the same queue cut into many device batches on several engines;
- the 4000x3000 sRGB photo at q75, encoded and decoded;
- Poisson serving of full-size pages, 60 % encodes / 40 % decodes, each
  result equal to the request alone.

The oracle runs in host threads (ctypes releases the interpreter lock)."""
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import jp2lam_amd as j
from jp2lam_amd import synth

pytestmark = pytest.mark.gpu

N_MIXED = 16


@pytest.fixture(scope="module")
def mixed_full():
    return list(synth.mixed_pages(N_MIXED, seed=4))


def _oracle_decode(ora, data):
    try:
        return ora.decode(data)
    except Exception as e:  # the reference decoder rejects some of its own streams
        return e


@pytest.fixture(scope="module")
def mixed_ref(mixed_full, ora):
    with ThreadPoolExecutor(8) as ex:
        streams = list(ex.map(lambda pq: ora.encode(*pq), mixed_full))
        decs = list(ex.map(lambda d: _oracle_decode(ora, d), streams))
    return streams, decs


def _check_decodes(got, want):
    for k, (g, w) in enumerate(zip(got, want)):
        if isinstance(w, Exception):
            assert isinstance(g, j.DecodeFailed), k
        else:
            assert isinstance(g, np.ndarray) and np.array_equal(g, w), k


def test_mixed_queue_full_size_matches_oracle(gpu, mixed_full, mixed_ref):
    streams = j.encode_many(mixed_full)
    want_streams, want_decs = mixed_ref
    for k, (s, w) in enumerate(zip(streams, want_streams)):
        assert s == w, (k, mixed_full[k][0].shape, mixed_full[k][1])
    _check_decodes(j.decode_many(streams), want_decs)


def test_mixed_queue_in_many_batches_equals_one(gpu, lib, mixed_full, mixed_ref):
    """Small device batches on three engines at once give the same bytes
    and pixels as the one-batch queue."""
    want_streams, want_decs = mixed_ref
    lib.jp2g_set_queue(12000, 3)
    try:
        streams = j.encode_many(mixed_full)
        decs = j.decode_many(streams)
    finally:
        lib.jp2g_set_queue(0, 0)
    assert streams == want_streams
    _check_decodes(decs, want_decs)


def test_photo_4000x3000_q75_encode_decode(gpu, ora):
    px = synth.photo_page(4000, 3000, seed=3)
    ours = j.encode_u8(px, 75)
    ref = ora.encode(px, 75)
    assert ours == ref
    assert np.array_equal(j.decode_u8(ours), ora.decode(ref))


def test_full_size_serving_mix_equals_alone(gpu, mixed_full):
    """Poisson arrivals (100/s) of full-size pages: 60 % encodes of the
    queue's pages, 40 % decodes of their streams; every result equals the
    same request alone."""
    rng = np.random.default_rng(55)
    dec_ok = []
    streams = [j.encode_u8(px, q) for px, q in mixed_full]
    alone_dec = {}
    for k, s in enumerate(streams):
        try:
            alone_dec[k] = j.decode_u8(s)
            dec_ok.append(k)
        except j.Jp2LamError as e:
            alone_dec[k] = e
    n = 40
    arrivals = np.cumsum(rng.exponential(0.01, n))
    kinds = rng.uniform(size=n) < 0.6
    picks = [int(rng.integers(0, len(mixed_full))) if e else dec_ok[int(rng.integers(0, len(dec_ok)))]
             for e in kinds]
    with j.Server(slots=2, mem_budget_mib=32768, max_batch=16) as srv:
        ids = []
        t0 = time.perf_counter()
        for i in range(n):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            k = picks[i]
            ids.append(srv.submit_encode(*mixed_full[k]) if kinds[i] else srv.submit_decode(streams[k]))
        got = [srv.wait(i, 120000) for i in ids]
    for i, g in enumerate(got):
        assert g is not None, i
        k = picks[i]
        if kinds[i]:
            assert g.value == streams[k], i
        else:
            assert isinstance(g.value, np.ndarray) and np.array_equal(g.value, alone_dec[k]), i
