"""Benchmark: encode + decode Mpixel/s of one 2550x3300 8-bit grayscale
document page at quality 85 (9/7, 5 levels, 64x64 code-blocks, JP2), with
the page resident in HBM and the decoded pixels written back to HBM.

One step = encode the page (device pixels -> JP2 bytes on the host, Tier-2
on the host) + decode those bytes (host Tier-2 parse -> device pixels).
Prints one JSON line; see --help.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# HIP maps this process's streams onto GPU_MAX_HW_QUEUES hardware queues (4
# by default), and two streams on one queue run one after the other.  The
# run holds up to eight streams: torch's, the process-wide engine's and its
# single-image decode side stream, the two serving slots and the three queue
# engines.  With four queues the two serving slots shared one and ran in
# turn: 64-page bursts 670-684 pages/s with four queues, 971-1,032 with
# eight, same box.  Raised here, before anything starts HIP.
HW_QUEUES = 8
if int(os.environ.get("GPU_MAX_HW_QUEUES") or 0) < HW_QUEUES:
    os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--quality", type=int, default=85)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-check", action="store_true", help="skip the post-run oracle check")
    ap.add_argument("--batch-pages", type=int, default=16,
                    help="pages in the batched-encode measurement (0 = skip)")
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="length of the serving leg (Poisson arrivals of page encodes); 0 = skip")
    ap.add_argument("--serve-rate", type=float, default=100.0, help="serving arrival rate, requests/s")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other workloads (512x512 q100 J2K, 4000x3000 sRGB q75, mixed-page queue)")
    ap.add_argument("--mixed-pages", type=int, default=256,
                    help="pages of the mixed-size queue leg (the full queue is 256)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB (a known-size kernel "
                         "for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


_T0 = time.perf_counter()


def note(msg):
    """Progress line on stderr (stdout carries only the result line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def enc_plain_host(j, px, q):
    return j.encode_u8(px, q)


CODERS = {1: "wave per block", 2: "lane per block", 4: "hybrid (long streams scalar, short on lanes)"}


def level_areas(w, h, levels):
    areas = []
    for _ in range(levels):
        areas.append(w * h)
        w, h = (w + 1) // 2, (h + 1) // 2
    return areas


def main():
    args = parse()
    if args.gpus != 1:
        sys.exit("bench.py measures one GPU: every box here has a single MI355X; pages shard across GPUs "
                 "without exchange (see DESIGN.md), so run it with --gpus 1")
    import numpy as np
    import torch

    import jp2lam_amd as j
    from jp2lam_amd import _lib, synth

    lib = _lib.load()
    if not torch.cuda.is_available() or not lib.jp2g_device_available():
        sys.exit("bench.py needs a HIP device")
    W, H, Q = 2550, 3300, args.quality
    px = synth.document_page(W, H, seed=2)
    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(px).to(dev)
    d_out = torch.empty_like(d_in)
    torch.cuda.synchronize()

    tim = (np.zeros(4, np.float64), np.zeros(4, np.float64))
    note(f"plain path: {W}x{H} q{Q}, {args.warmup} warm-up + {args.steps} timed steps")

    def step():
        enc = j.encode_many([((d_in.data_ptr(), H, W, 1), Q)], device=True)[0]
        if isinstance(enc, Exception):
            raise enc
        t = np.zeros(4, np.float64)
        lib.jp2g_last_timings(t.ctypes.data)
        res = j.decode_many([enc], device_outputs=[(d_out.data_ptr(), d_out.numel())])[0]
        if isinstance(res, Exception):
            raise res
        t2 = np.zeros(4, np.float64)
        lib.jp2g_last_timings(t2.ctypes.data)
        return enc, t, t2

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    enc_t = np.zeros(4)
    dec_t = np.zeros(4)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        enc, te, td = step()
        enc_t += te
        dec_t += td
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1000.0 / args.steps
    enc_t /= args.steps
    dec_t /= args.steps
    mpx = W * H / 1e6 / (ms / 1000.0)

    # ---- roofline of the dominant kernel (HIP events on the engine stream) ----
    P = W * H
    areas = level_areas(W, H, 5)
    stages = {
        # name: (avg ms, algorithmic bytes per launch, bound)
        "encode_dwt (level shift + 5 levels)": (enc_t[0], P * (1 + 4) + 8 * sum(areas)),
        "t1_encode_kernel": (enc_t[1], 4 * P + len(enc)),
        "pcrd+pack": (enc_t[2], 0),
        "t1_decode_kernel": (dec_t[1], len(enc) + 4 * P),
        "decode_idwt (5 levels + finalize)": (dec_t[0], 8 * sum(areas) + P * (4 + 1)),
    }
    dom = max(stages, key=lambda k: stages[k][0])
    dms, dbytes = stages[dom]
    achieved = dbytes / (dms / 1e3) / 1e9 if dms > 0 else 0.0
    # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    # passes (scaled by the copy-kernel calibration recorded with them)
    traffic = None
    tfile = ROOT / "profiles" / "hbm_traffic_doc2550x3300_q85.json"
    if tfile.exists() and (W, H, Q) == (2550, 3300, 85):
        kern = json.loads(tfile.read_text())["kernels"].get(dom.split(" ")[0])
        if kern:
            traffic = int(kern["bytes"])
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "algorithmic_bytes": int(dbytes), "avg_ms": round(dms, 4),
            "note": "Tier-1 is bound by the serial MQ chain of its heaviest code-block, not by HBM"}

    # Tier-1 work units, outside the timed region, from the device counters:
    # coefficient bit-planes (sum over blocks of w*h*mbp) and MQ decisions,
    # in all and on the critical path (the block with the most decisions:
    # all passes for the encoder, the kept passes for the decoder)
    lib.jp2g_set_debug(1)
    lib.jp2g_reset_counters()
    lib.jp2g_set_counters(1)
    try:
        j.encode_many([((d_in.data_ptr(), H, W, 1), Q)], device=True)
        info = np.zeros(16, np.int32)
        cbp = 0
        for bi in range(lib.jp2g_last_block_count()):
            lib.jp2g_last_block_info(bi, info.ctypes.data)
            cbp += int(info[5] - info[3]) * int(info[6] - info[4]) * int(info[7])
        nblk = lib.jp2g_last_block_count()
        sym_all = [lib.jp2g_last_block_symbols(bi, 0) for bi in range(nblk)]
        sym_kept = [lib.jp2g_last_block_symbols(bi, 1) for bi in range(nblk)]
        ctr = np.zeros(7, np.uint64)
        lib.jp2g_counters(ctr.ctypes.data)
    finally:
        lib.jp2g_set_counters(0)
        lib.jp2g_set_debug(0)
    roof["t1_encode_coeff_bitplanes_per_s"] = round(cbp / (enc_t[1] / 1e3), 1) if enc_t[1] > 0 else None
    roof["t1_decode_coeff_bitplanes_per_s"] = round(cbp / (dec_t[1] / 1e3), 1) if dec_t[1] > 0 else None
    crit_e, crit_d = max(sym_all), max(sym_kept)
    roof["t1_critical_path"] = {
        "note": "Tier-1 is a serial MQ chain per code-block: the kernel ends with its heaviest block, so its "
                "time over that block's decisions is the cost per decision on the critical path",
        "mq_decisions_per_page": int(ctr[2]), "mq_decisions_kept_per_page": int(sum(sym_kept)),
        "encode_heaviest_block_decisions": int(crit_e),
        "encode_ns_per_decision": round(enc_t[1] * 1e6 / crit_e, 1) if crit_e else None,
        "decode_heaviest_block_decisions": int(crit_d),
        "decode_ns_per_decision": round(dec_t[1] * 1e6 / crit_d, 1) if crit_d else None,
        "encode_decisions_per_s": round(int(ctr[2]) / (enc_t[1] / 1e3), 1) if enc_t[1] > 0 else None,
        "decode_decisions_per_s": round(sum(sym_kept) / (dec_t[1] / 1e3), 1) if dec_t[1] > 0 else None}
    tfile_u = ROOT / "profiles" / "utilisation_t1_dwt_doc2550x3300_q85.json"
    if tfile_u.exists() and (W, H, Q) == (2550, 3300, 85):
        u = json.loads(tfile_u.read_text()).get(f"plain:{dom.split(' ')[0]}")
        if u:
            roof["utilisation_from_profile"] = {k: u[k] for k in ("valu_busy", "salu_busy", "waves_per_cu",
                                                                  "lds_bank_conflict_pct") if k in u}
    stage_ms = {k: round(v[0], 4) for k, v in stages.items()}
    stage_ms["encode_host_tier2"] = round(enc_t[3], 4)

    # ---- plain encode only, then batched encode of distinct pages ----------
    def enc_plain():
        r = j.encode_many([((d_in.data_ptr(), H, W, 1), Q)], device=True)[0]
        if isinstance(r, Exception):
            raise r
        return r

    batched = None
    if args.batch_pages > 0:
        enc_plain()
        torch.cuda.synchronize()
        reps = max(3, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(reps):
            enc_plain()
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - t0) * 1000.0 / reps
        nb = args.batch_pages
        pages = [torch.from_numpy(synth.document_page(W, H, seed=2 + 1000 * k)).to(dev) for k in range(nb)]
        items = [((pg.data_ptr(), H, W, 1), Q) for pg in pages]
        torch.cuda.synchronize()
        res = j.encode_many(items, device=True)
        bad = [r for r in res if isinstance(r, Exception)]
        if bad:
            raise bad[0]
        breps = max(2, reps // 2)
        t0 = time.perf_counter()
        for _ in range(breps):
            j.encode_many(items, device=True)
        torch.cuda.synchronize()
        batch_ms = (time.perf_counter() - t0) * 1000.0 / breps
        tb = np.zeros(4, np.float64)
        lib.jp2g_last_timings(tb.ctypes.data)
        batched = {"pages": nb, "t1_coder": CODERS.get(lib.jp2g_last_t1_coder()),
                   "encode_ms_per_batch": round(batch_ms, 3),
                   "encode_mpx_s": round(nb * W * H / 1e6 / (batch_ms / 1e3), 2),
                   "plain_encode_ms_per_page": round(plain_ms, 3),
                   "plain_encode_mpx_s": round(W * H / 1e6 / (plain_ms / 1e3), 2),
                   "speedup_vs_plain": round(plain_ms * nb / batch_ms, 3),
                   "stage_ms": {"dwt": round(tb[0], 3), "t1_encode": round(tb[1], 3), "pcrd_pack": round(tb[2], 3),
                                "host_tier2": round(tb[3], 3)}}
        del pages

    # ---- serving: Poisson arrivals of page encodes, continuous batching ------
    serve = None
    if args.serve_seconds > 0:
        import threading

        pool = [synth.document_page(W, H, seed=2 + 1000 * k) for k in range(4)]
        want = [enc_plain_host(j, p, Q) for p in pool]
        n = max(1, int(args.serve_rate * args.serve_seconds))
        note(f"serving: {n} Poisson arrivals at {args.serve_rate:g}/s")
        rng = np.random.default_rng(5)
        arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, n))
        with j.Server(slots=2, mem_budget_mib=32768, max_batch=16) as srv:
            for p in pool:  # warm the two engines' arenas
                srv.wait(srv.submit_encode(p, Q))

            def replay():
                ids = [None] * n
                done = [None] * n

                def waiter():
                    for i in range(n):
                        while ids[i] is None:
                            time.sleep(0.0002)
                        done[i] = srv.wait(ids[i])

                th = threading.Thread(target=waiter)
                th.start()
                t0 = time.perf_counter()
                for i in range(n):
                    d = t0 + arrivals[i] - time.perf_counter()
                    if d > 0:
                        time.sleep(d)
                    ids[i] = srv.submit_encode(pool[i % len(pool)], Q)
                    if (i + 1) % 1000 == 0:
                        note(f"serving: {i + 1}/{n} submitted")
                th.join()
                return done, time.perf_counter() - t0

            # One measured pass: the coder comes from the offline table, so
            # no arrival pays for tuning.
            b0 = srv.stats()["batches"]
            note("serving: measured pass")
            done, t_span = replay()
            nbatches = srv.stats()["batches"] - b0
            # saturation: bursts of requests submitted at once, so the loop
            # batches them (capacity, not the arrival rate); one burst first
            # so the arenas have grown to the batch sizes
            nb_sat = 64

            def burst(device):
                t1 = time.perf_counter()
                if device:
                    sids = [srv.submit_encode(None, Q, device=(dpool[i % len(pool)].data_ptr(), H, W, 1))
                            for i in range(nb_sat)]
                else:
                    sids = [srv.submit_encode(pool[i % len(pool)], Q) for i in range(nb_sat)]
                res = [srv.wait(i) for i in sids]
                return res, time.perf_counter() - t1

            dpool = [torch.from_numpy(p).to(dev) for p in pool]
            torch.cuda.synchronize()
            # one burst to warm, then the median of three (bursts vary by
            # 10-20 % from one to the next on a box)
            burst(True)
            runs = [burst(True) for _ in range(3)]
            sat_each = sorted(nb_sat / t for _, t in runs)
            t_sat = nb_sat / sat_each[1]
            burst(False)  # fills the pinned pool
            runs_h = [burst(False) for _ in range(3)]
            sat_each_h = sorted(nb_sat / t for _, t in runs_h)
            t_sat_h = nb_sat / sat_each_h[1]
        lat = np.array([r.total_ms for r in done])
        ok = all(done[i].value == want[i % len(pool)] for i in range(n))
        ok_sat = all(r[0][i].value == want[i % len(pool)] for r in runs + runs_h for i in range(nb_sat))
        sat_pps = nb_sat / t_sat
        serve = {"workload": f"{W}x{H} gray q{Q} page encodes, host pixels, Poisson arrivals",
                 "rate_req_s": args.serve_rate, "requests": n, "slots": 2,
                 "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                 "max_ms": round(float(lat.max()), 2),
                 "throughput_pages_s": round(n / t_span, 1),
                 "mean_batch": round(float(np.mean([r.batch_size for r in done])), 2),
                 "batches": int(nbatches), "results_equal_alone": bool(ok),
                 "saturated": {"requests": nb_sat, "input": "pages resident in HBM, as in the plain path",
                               "pages_s": round(sat_pps, 1),
                               "bursts_pages_s": [round(v, 1) for v in sat_each],
                               "mpx_s": round(sat_pps * W * H / 1e6, 1),
                               "mean_batch": round(float(np.mean([r.batch_size for b, _ in runs for r in b])), 2),
                               "p99_ms": round(float(np.percentile([r.total_ms for b, _ in runs for r in b], 99)), 2),
                               "host_pixels_pages_s": round(nb_sat / t_sat_h, 1),
                               "host_pixels_bursts_pages_s": [round(v, 1) for v in sat_each_h],
                               "results_equal_alone": bool(ok_sat)}}
        if batched:
            serve["saturated"]["speedup_vs_plain_encode"] = round(sat_pps * batched["plain_encode_ms_per_page"] / 1e3, 3)

    # ---- the other workloads: 512x512 q100 J2K, 4000x3000 sRGB q75, mixed queue ----
    configs = None
    if not args.no_configs:
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle_ctypes import Oracle

        ora = Oracle()
        configs = {}

        def timed(fn, reps):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                r = fn()
            torch.cuda.synchronize()
            return r, (time.perf_counter() - t0) * 1000.0 / reps

        # gradient page, lossless J2K: encode only (the reference decodes JP2 only)
        g = synth.gradient_page(512, 512, seed=1)
        dg = torch.from_numpy(g).to(dev)
        enc1, ms1 = timed(lambda: j.encode_many([((dg.data_ptr(), 512, 512, 1), 100)], format=j.OutputFormat.J2K,
                                                device=True)[0], max(5, args.steps))
        configs["gradient512_gray_q100_j2k_encode"] = {
            "ms": round(ms1, 3), "mpx_s": round(512 * 512 / 1e3 / ms1, 2), "codestream_bytes": len(enc1),
            "parity_vs_oracle": bool(enc1 == ora.encode(g, 100, fmt=1))}

        # photo page, sRGB q75: encode then decode, pixels resident in HBM
        W3, H3 = 4000, 3000
        ph = synth.photo_page(W3, H3, seed=3)
        dph = torch.from_numpy(ph).to(dev)
        dph_out = torch.empty_like(dph)
        enc3, e3 = timed(lambda: j.encode_many([((dph.data_ptr(), H3, W3, 3), 75)], device=True)[0], max(3, args.steps // 2))
        te3 = np.zeros(4, np.float64)
        lib.jp2g_last_timings(te3.ctypes.data)
        coder3 = CODERS.get(lib.jp2g_last_t1_coder())
        _, d3 = timed(lambda: j.decode_many([enc3], device_outputs=[(dph_out.data_ptr(), dph_out.numel())])[0],
                      max(3, args.steps // 2))
        td3 = np.zeros(4, np.float64)
        lib.jp2g_last_timings(td3.ctypes.data)
        ref3 = ora.encode(ph, 75)
        configs["photo4000x3000_srgb_q75"] = {
            "encode_ms": round(e3, 3), "decode_ms": round(d3, 3),
            "mpx_s": round(W3 * H3 / 1e3 / (e3 + d3), 2), "codestream_bytes": len(enc3), "t1_coder": coder3,
            "stage_ms": {"dwt": round(te3[0], 3), "t1_encode": round(te3[1], 3), "pcrd_pack": round(te3[2], 3),
                         "t1_decode": round(td3[1], 3), "idwt_finalize": round(td3[0], 3)},
            "parity_vs_oracle": bool(enc3 == ref3 and np.array_equal(dph_out.cpu().numpy(), ora.decode(ref3)))}
        del dph, dph_out

        # mixed-size queue: 70 % documents / 30 % photos, q in {50, 75, 85, 100},
        # pages resident in HBM, encoded as ONE jp2lam_encode_many call and
        # decoded as one jp2lam_decode_many call (the queue cuts itself into
        # tuned device batches on several engines)
        if args.mixed_pages > 0:
            workers = max(1, min(16, os.cpu_count() or 1))
            note(f"mixed queue: generating {args.mixed_pages} pages on {workers} processes")
            host = synth.mixed_pages_parallel(args.mixed_pages, seed=4, workers=workers)
            note("mixed queue: one encode_many call, one decode_many call")
            devp = [torch.from_numpy(p).to(dev) for p, _ in host]
            items = [((d.data_ptr(), p.shape[0], p.shape[1], 1 if p.ndim == 2 else 3), q)
                     for d, (p, q) in zip(devp, host)]
            outs = [torch.empty_like(d) for d in devp]
            torch.cuda.synchronize()

            def enc_all():
                res = j.encode_many(items, device=True)
                bad = [r for r in res if isinstance(r, Exception)]
                if bad:
                    raise bad[0]
                return res

            streams = enc_all()  # warm-up: grows every queue engine's arenas
            # streams the reference decoder rejects (sRGB lossless, its chroma
            # Mb quirk; zero-length first contributions, decode/t2.rs:243) are
            # rejected by ours too: they stay out of the decode half and the
            # serving leg's decode requests, and are listed with the reason
            rejected = {}
            for k in range(len(host)):
                try:
                    j.inspect_jp2(streams[k])
                    if host[k][0].ndim == 3 and host[k][1] >= 100:
                        raise j.DecodeFailed("sRGB lossless: chroma Mb (backend.rs:179-189)")
                except j.Jp2LamError as e:
                    rejected[k] = str(e)
            dec_ids = [k for k in range(len(host)) if k not in rejected]

            def dec_all():
                res = j.decode_many([streams[k] for k in dec_ids],
                                    device_outputs=[(outs[k].data_ptr(), outs[k].numel()) for k in dec_ids])
                for k, r in zip(dec_ids, res):
                    if isinstance(r, Exception):
                        try:  # the single call carries the reason
                            j.decode_u8(streams[k])
                        except j.Jp2LamError as e:
                            r = e
                        rejected[k] = str(r)
                return res

            dec_all()
            dec_ids = [k for k in dec_ids if k not in rejected]
            em = dm = float("inf")
            for _ in range(2):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                again = enc_all()
                torch.cuda.synchronize()
                em = min(em, (time.perf_counter() - t0) * 1e3)
                t0 = time.perf_counter()
                dec_all()
                torch.cuda.synchronize()
                dm = min(dm, (time.perf_counter() - t0) * 1e3)
            same = again == streams
            qmpx = sum(p.shape[0] * p.shape[1] for p, _ in host) / 1e6
            dmpx = sum(host[k][0].shape[0] * host[k][0].shape[1] for k in dec_ids) / 1e6
            # every page: the queue's bytes and pixels equal the page alone
            note("mixed queue: every page against its single call")
            same_enc = all(streams[k] == j.encode_u8(host[k][0], host[k][1]) for k in range(len(host)))
            same_dec = all(np.array_equal(outs[k].cpu().numpy(), j.decode_u8(streams[k])) for k in dec_ids)
            # the oracle on the smallest gray and the smallest sRGB page
            ocheck = {}
            for kind in (2, 3):
                ks = [k for k in range(len(host)) if host[k][0].ndim == kind]
                if ks:
                    k = min(ks, key=lambda i: host[i][0].size)
                    ref = ora.encode(host[k][0], host[k][1])
                    ok = streams[k] == ref
                    if ok and k in dec_ids:
                        ok = np.array_equal(outs[k].cpu().numpy(), ora.decode(ref))
                    ocheck[f"page{k}_{'gray' if kind == 2 else 'srgb'}_q{host[k][1]}"] = bool(ok)
            import ctypes

            qc, qe = ctypes.c_uint32(), ctypes.c_int()
            lib.jp2g_get_queue(ctypes.byref(qc), ctypes.byref(qe))
            configs["mixed_queue"] = {
                "pages": len(host), "megapixels": round(qmpx, 1), "calls": "one encode_many, one decode_many",
                "queue_chunk_blocks": int(qc.value), "queue_engines": int(qe.value),
                "gray_pages": sum(1 for p, _ in host if p.ndim == 2),
                "decoded_pages": len(dec_ids), "decode_rejected_as_by_reference": {str(k): v for k, v in rejected.items()},
                "encode_s": round(em / 1e3, 3), "decode_s": round(dm / 1e3, 3),
                "encode_mpx_s": round(qmpx / (em / 1e3), 1), "decode_mpx_s": round(dmpx / (dm / 1e3), 1),
                "repeat_identical": bool(same), "every_page_equals_single_call": bool(same_enc and same_dec),
                "parity_vs_oracle": ocheck}
            del devp, outs

            # mixed serving: Poisson arrivals of 60 % encodes of the queue's
            # pages and 40 % decodes of their codestreams, host memory; each
            # result is checked against the same request run alone
            if args.serve_seconds > 0:
                import threading

                want_enc = {}
                want_dec = {k: j.decode_u8(streams[k]) for k in dec_ids}
                nreq = max(1, int(args.serve_rate * args.serve_seconds))
                rng = np.random.default_rng(55)
                arr = np.cumsum(rng.exponential(1.0 / args.serve_rate, nreq))
                is_enc = rng.uniform(size=nreq) < 0.6
                pick = [int(rng.integers(0, len(host))) if e else dec_ids[int(rng.integers(0, len(dec_ids)))]
                        for e in is_enc]
                for k in set(p for p, e in zip(pick, is_enc) if e):
                    want_enc[k] = j.encode_u8(host[k][0], host[k][1])
                host_sizes_desc = sorted(range(len(host)), key=lambda k: -host[k][0].size)
                with j.Server(slots=2, mem_budget_mib=32768, max_batch=16) as srv:
                    def stream_once():
                        rids = [None] * nreq
                        rdone = [None] * nreq

                        def waiter2():
                            for i in range(nreq):
                                while rids[i] is None:
                                    time.sleep(0.0002)
                                rdone[i] = srv.wait(rids[i])

                        th = threading.Thread(target=waiter2)
                        th.start()
                        t0m = time.perf_counter()
                        for i in range(nreq):
                            dly = t0m + arr[i] - time.perf_counter()
                            if dly > 0:
                                time.sleep(dly)
                            k = pick[i]
                            rids[i] = (srv.submit_encode(host[k][0], host[k][1]) if is_enc[i]
                                       else srv.submit_decode(streams[k]))
                            if (i + 1) % 1000 == 0:
                                note(f"mixed serving: {i + 1}/{nreq} submitted")
                        th.join()
                        return rdone, t0m, time.perf_counter()

                    # a few requests of each kind first grow the engines'
                    # arenas and the pinned pool; then one measured pass
                    for k in sorted(host_sizes_desc[:4]):
                        srv.wait(srv.submit_encode(host[k][0], host[k][1]))
                    for k in dec_ids[:2]:
                        srv.wait(srv.submit_decode(streams[k]))
                    note(f"mixed serving: {nreq} arrivals, measured pass")
                    rdone, t0m, t_endm = stream_once()
                okm = all((rdone[i].value == want_enc[pick[i]]) if is_enc[i]
                          else np.array_equal(rdone[i].value, want_dec[pick[i]]) for i in range(nreq))
                lat = np.array([r.total_ms for r in rdone])
                le = lat[is_enc]
                ld = lat[~is_enc]
                served_mpx = sum(host[pick[i]][0].shape[0] * host[pick[i]][0].shape[1] for i in range(nreq)) / 1e6
                configs["mixed_serving"] = {
                    "workload": "Poisson arrivals: 60 % encodes of the mixed-queue pages (host pixels), "
                                "40 % decodes of their codestreams", "rate_req_s": args.serve_rate,
                    "requests": nreq, "encodes": int(is_enc.sum()), "slots": 2,
                    "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
                    "encode_p99_ms": round(float(np.percentile(le, 99)), 2) if le.size else None,
                    "decode_p99_ms": round(float(np.percentile(ld, 99)), 2) if ld.size else None,
                    "max_ms": round(float(lat.max()), 2), "throughput_req_s": round(nreq / (t_endm - t0m), 1),
                    "served_mpx_s": round(served_mpx / (t_endm - t0m), 1),
                    "mean_batch": round(float(np.mean([r.batch_size for r in rdone])), 2),
                    "results_equal_alone": bool(okm)}

    if args.calib_copy_mib > 0:
        n = args.calib_copy_mib << 20
        a = torch.ones(n, dtype=torch.uint8, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    # ---- parity check of the measured output (outside the timed region) ----
    check = None
    if not args.no_check:
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle_ctypes import Oracle

        ora = Oracle()
        ref = ora.encode(px, Q)
        out = d_out.cpu().numpy()
        check = bool(enc == ref and np.array_equal(out, ora.decode(ref)))

    # ---- CPU baseline: the oracle (C restatement, OpenMP) on the same page ----
    # Three figures on the host cores: the faithful CPU path (the work the
    # reference does, including the contrast-mask DCT and plan_coding_passes
    # whose results it discards), the lean path (same bytes without that
    # work), both on every core, and the lean path on one thread.  `value` is
    # the faithful figure: that is the reference's CPU path.
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle_ctypes import Oracle

        ora = Oracle()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        note(f"cpu baseline: oracle on {threads} threads (lean, faithful) and 1 thread (lean)")

        def cpu_run(faithful, nthreads, runs):
            ora.L.ora_set_faithful(1 if faithful else 0)
            ora.L.ora_set_threads(nthreads)
            try:
                te, td = [], []
                for _ in range(runs):
                    a = time.perf_counter()
                    d = ora.encode(px, Q)
                    b_ = time.perf_counter()
                    ora.decode(d)
                    te.append(b_ - a)
                    td.append(time.perf_counter() - b_)
                    assert d == enc
                return statistics.median(te), statistics.median(td)
            finally:
                ora.L.ora_set_faithful(0)
                ora.L.ora_set_threads(0)

        cpu_run(False, threads, 1)  # warm-up
        le, ld = cpu_run(False, threads, max(1, args.cpu_runs))
        fe, fd = cpu_run(True, threads, 1)
        se, sd = cpu_run(False, 1, 1)
        mp = P / 1e6
        cpu = {"value": round(mp / (fe + fd), 3), "unit": "Mpixel/s", "cores": threads, "kind": "port",
               "sample": f"one full 2550x3300 q{Q} page encode+decode per figure (oracle C restatement, "
                         f"same bytes as the GPU): faithful = the reference's CPU work including its discarded "
                         f"contrast-mask DCT and plan_coding_passes, 1 run; lean = without that work, median of "
                         f"{max(1, args.cpu_runs)} after 1 warm-up; T1 encode and large 9/7 row passes over "
                         f"OpenMP threads, T1 decode serial like the reference",
               "faithful": {"threads": threads, "encode_mpx_s": round(mp / fe, 3), "decode_mpx_s": round(mp / fd, 3),
                            "encode_decode_mpx_s": round(mp / (fe + fd), 3)},
               "lean": {"threads": threads, "encode_mpx_s": round(mp / le, 3), "decode_mpx_s": round(mp / ld, 3),
                        "encode_decode_mpx_s": round(mp / (le + ld), 3)},
               "lean_1_thread": {"threads": 1, "encode_mpx_s": round(mp / se, 3), "decode_mpx_s": round(mp / sd, 3),
                                 "encode_decode_mpx_s": round(mp / (se + sd), 3)}}
        if batched:
            cpu["batched_gpu_encode_vs_faithful_cpu_encode"] = round(batched["encode_mpx_s"] / (mp / fe), 1)
            cpu["batched_gpu_encode_vs_lean_cpu_encode"] = round(batched["encode_mpx_s"] / (mp / le), 1)

    line = {
        "metric": "encode+decode Mpixel/s on 1 GPU vs plain GPU & CPU; p50/p99 page latency",
        "value": round(mpx, 3),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded document page: background 245+N(0,3), 12% dark glyph rectangles)",
        "config": {"workload": "plain per-image GPU path: 2550x3300 gray document page, q85 9/7 5 levels, "
                               "64x64 code-blocks, JP2; encode then decode, pixels resident in HBM",
                   "width": W, "height": H, "quality": Q, "codestream_bytes": len(enc),
                   "gpu_max_hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"])},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stage_ms": stage_ms,
        "parity_vs_oracle": check,
        "batched_encode": batched,
        "serving": serve,
        "other_workloads": configs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
