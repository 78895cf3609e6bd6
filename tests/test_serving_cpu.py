"""The serving loop's admission, cancellation and device-fault handling on
the CPU, through the C ABI, with the simulated executor (no device work):
the request loop, the budget check, the queue and the error mapping are the
same code the device server runs."""
import os
import struct
import subprocess
import sys
import textwrap
import time

import numpy as np
import pytest

import jp2lam_amd as j


def _digest(px, q, cs):
    h = 1469598103934665603
    for b in px.tobytes():
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return struct.pack("<4IQ", px.shape[1], px.shape[0], cs, q, h)


def test_simulated_requests_complete_in_order_of_their_own_results():
    rng = np.random.default_rng(1)
    pages = [rng.integers(0, 256, (8 + i, 10 + i), dtype=np.uint8) for i in range(12)]
    with j.Server(slots=2, mem_budget_mib=1024, max_batch=4, simulated_delay_us=2000) as srv:
        ids = [srv.submit_encode(p, 60 + i) for i, p in enumerate(pages)]
        dec = srv.submit_decode(b"abcdef")
        got = [srv.wait(i, 10000) for i in ids]
        gd = srv.wait(dec, 10000)
        st = srv.stats()
    for i, (p, g) in enumerate(zip(pages, got)):
        assert g.value == _digest(p, 60 + i, 0), i  # GRAY = 0
    assert gd.value.tobytes() == b"abcdef"
    assert st["completed"] == 13 and st["batches"] < 13  # requests shared batches


def test_admission_refuses_requests_over_the_slot_budget():
    # 1 MiB over two slots: 512 KiB per slot; an encode is estimated at 48
    # bytes per sample, so 10,923 samples is the largest page admitted
    with j.Server(slots=2, mem_budget_mib=1, max_batch=4, simulated_delay_us=0) as srv:
        small = np.zeros((100, 100), np.uint8)
        big = np.zeros((120, 100), np.uint8)
        assert srv.wait(srv.submit_encode(small, 85), 5000).value is not None
        with pytest.raises(j.InvalidInput, match="per-slot budget"):
            srv.submit_encode(big, 85)
        assert srv.faults()["rejected"] == 1
        # the server keeps serving after a refusal
        assert srv.wait(srv.submit_encode(small, 50), 5000).value == _digest(small, 50, 0)


def test_cancel_removes_a_queued_request_only():
    with j.Server(slots=1, mem_budget_mib=1024, max_batch=1, simulated_delay_us=200000) as srv:
        px = np.zeros((4, 4), np.uint8)
        first = srv.submit_encode(px, 85)
        time.sleep(0.05)  # the slot has taken the first request
        queued = [srv.submit_encode(px, 80 + i) for i in range(3)]
        assert srv.cancel(queued[1]) is True
        assert srv.cancel(first) is False  # already running
        r_first = srv.wait(first, 5000)
        r = [srv.wait(i, 5000) for i in queued]
        assert r_first.value == _digest(px, 85, 0)
        assert isinstance(r[1].value, j.IoError) and "cancelled" in str(r[1].value)
        assert r[0].value == _digest(px, 80, 0) and r[2].value == _digest(px, 82, 0)
        assert srv.faults()["cancelled"] == 1
        with pytest.raises(j.InvalidInput):
            srv.cancel(10 ** 9)


FAULT_SCRIPT = textwrap.dedent("""
    import numpy as np, sys
    import jp2lam_amd as j
    px = np.zeros((4, 4), np.uint8)
    with j.Server(slots=1, mem_budget_mib=1024, max_batch=1, simulated_delay_us=0) as srv:
        kinds = []
        for k in range(3):
            r = srv.wait(srv.submit_encode(px, 85), 5000).value
            kinds.append(type(r).__name__ + ":" + (str(r) if isinstance(r, Exception) else "ok"))
        d = srv.wait(srv.submit_decode(b"xyz"), 5000).value
        kinds.append(type(d).__name__ + ":" + (str(d) if isinstance(d, Exception) else "ok"))
        print("|".join(kinds)); print(srv.faults()["device_faults"])
""")


@pytest.mark.parametrize("plan,expect", [
    ("encode:2", ["bytes:ok", "EncodeFailed", "bytes:ok", "ndarray:ok"]),
    ("decode", ["bytes:ok", "bytes:ok", "bytes:ok", "DecodeFailed"]),
])
def test_injected_device_fault_maps_to_encode_or_decode_failed(plan, expect):
    env = dict(os.environ, JP2LAM_FAULT_INJECT=plan, PYTHONPATH=os.pathsep.join(sys.path))
    out = subprocess.run([sys.executable, "-c", FAULT_SCRIPT], capture_output=True, text=True, env=env, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    kinds, nf = out.stdout.strip().splitlines()
    got = kinds.split("|")
    for g, e in zip(got, expect):
        assert g.startswith(e), (g, e)
        if "Failed" in e:
            assert "device error" in g and "injected" in g
    assert int(nf) == 1
