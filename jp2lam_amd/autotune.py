"""Offline autotuner for the batched Tier-1 coder choice.

Run on an otherwise idle MI355X::

    python -m jp2lam_amd.autotune [--out-json jp2lam_amd/t1_tuning.json]
                                  [--out-inc jp2lam_amd/csrc/t1_tuning.inc]

The reference codes every code-block the same way (encode/backend/native/
t1.rs:287-290); this build has three Tier-1 coders whose speed depends on
the batch (one wavefront per block, one lane per block, and the hybrid that
sends formed streams longer than ``frac`` x the batch's longest to scalar
coders and the rest to lanes).  The sweep encodes batches of document pages
and of the mixed-size queue (70 % documents, 30 % photos, q in {50, 75, 85,
100}) and a 4000x3000 sRGB photo, with every coder forced in turn, and
records the Tier-1 stage time (HIP events on the engine stream).  For each
size class (bit length of the batch's code-block count) it keeps the
(coder, frac) with the lowest geometric mean of time / best time over the
class's workloads, and writes that table as a C++ include that the engine
compiles in.  Output bytes are the same under every coder; the sweep checks
that too.  Nothing is probed while requests are served.

A second sweep tunes the queue form of jp2lam_encode_many /
jp2lam_decode_many: a long mixed-page list is encoded and decoded whole
with each (code-blocks per device batch, engines at once) pair, and the
pair with the lowest encode + decode wall time is written to
``t1_tuning_queue.inc``.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
if str(PKG.parent) not in sys.path:
    sys.path.insert(0, str(PKG.parent))

# The queue sweep runs up to three engines (a stream each, plus the
# single-image decode's side stream): with HIP's default of four hardware
# queues, streams share queues and run in turn, which would skew the sweep.
# Raised before anything starts HIP, as bench.py does.
if int(os.environ.get("GPU_MAX_HW_QUEUES") or 0) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

import jp2lam_amd as j  # noqa: E402
from jp2lam_amd import _lib, synth  # noqa: E402

# (name, t1 mode, hybrid split fraction)
CONFIGS = [("wave", 1, 0.0), ("lanes", 2, 0.0)] + [(f"hybrid{f:.2f}", 4, f) for f in (0.55, 0.6, 0.65, 0.7, 0.75, 0.85)]
CODER_ID = {1: "kT1Wave", 2: "kT1Lanes", 4: "kT1Hybrid"}


def log(msg: str) -> None:
    print(f"[autotune {time.monotonic() - T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


T0 = time.monotonic()


def workloads(doc_counts, mixed_counts, mixed_seed):
    docs = []
    for k in range(max(doc_counts)):
        docs.append((synth.document_page(2550, 3300, seed=100 + k), 85))
    log(f"{len(docs)} document pages generated")
    mixed = list(synth.mixed_pages(max(mixed_counts), seed=mixed_seed))
    log(f"{len(mixed)} mixed pages generated")
    out = [(f"doc{k}", docs[:k]) for k in doc_counts]
    out += [(f"mixed{k}", mixed[:k]) for k in mixed_counts]
    out.append(("photo4000x3000", [(synth.photo_page(4000, 3000, seed=3), 75)]))
    return out


def time_config(lib, pages, mode, frac, reps):
    lib.jp2g_set_t1_mode(mode, 0)
    if mode == 4:
        lib.jp2g_set_hybrid_frac(frac)
    tim = (C.c_double * 4)()
    best, streams = math.inf, None
    for _ in range(reps):
        out = j.encode_many(pages)
        lib.jp2g_last_timings(tim)
        best = min(best, tim[1])
        streams = out
    return best, streams


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out-json", default=str(PKG / "t1_tuning.json"))
    ap.add_argument("--out-inc", default=str(PKG / "csrc" / "t1_tuning.inc"))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--doc-counts", default="1,2,3,4,5,6,8,12,16,24,32")
    ap.add_argument("--mixed-counts", default="1,2,3,4,6,8,12,16,24")
    ap.add_argument("--mixed-seed", type=int, default=4)
    ap.add_argument("--out-queue-inc", default=str(PKG / "csrc" / "t1_tuning_queue.inc"))
    ap.add_argument("--queue-pages", type=int, default=64)
    ap.add_argument("--queue-chunks", default="32768,65536,131072,262144")
    ap.add_argument("--queue-engines", default="1,2,3")
    ap.add_argument("--skip-coder", action="store_true", help="run the queue sweep only")
    args = ap.parse_args()
    lib = _lib.load()
    if not lib.jp2g_device_available():
        raise SystemExit("autotune needs a HIP device")
    doc_counts = [int(x) for x in args.doc_counts.split(",")]
    mixed_counts = [int(x) for x in args.mixed_counts.split(",")]
    report = {}
    if not args.skip_coder:
        report.update(coder_sweep(lib, args, doc_counts, mixed_counts))
    report["queue"] = queue_sweep(lib, args)
    Path(args.out_json).write_text(json.dumps(report, indent=1) + "\n")
    log(f"wrote {args.out_json}")


def coder_sweep(lib, args, doc_counts, mixed_counts):
    # every workload runs as one device batch on the process-wide engine
    lib.jp2g_set_queue(1 << 18, 1)
    rows = []
    for name, pages in workloads(doc_counts, mixed_counts, args.mixed_seed):
        lib.jp2g_set_t1_mode(1, 0)
        j.encode_many(pages)  # warm-up, sizes the arenas
        times, ref = {}, None
        for cname, mode, frac in CONFIGS:
            t, streams = time_config(lib, pages, mode, frac, args.reps)
            if ref is None:
                ref = streams
            elif streams != ref:
                raise SystemExit(f"{name}: {cname} output differs from the wave coder's")
            times[cname] = t
        nb = sum(_blocks_of(px, q) for px, q in pages)
        rows.append({"workload": name, "pages": len(pages), "blocks": nb, "bucket": nb.bit_length(),
                     "t1_ms": {k: round(v, 4) for k, v in times.items()}})
        log(f"{name}: {nb} blocks, " + ", ".join(f"{k} {v:.2f}" for k, v in times.items()))
    lib.jp2g_set_t1_mode(0, 0)
    lib.jp2g_set_queue(0, 0)
    table = choose(rows)
    write_inc(Path(args.out_inc), table)
    log(f"wrote {args.out_inc}")
    return {"configs": [c[0] for c in CONFIGS], "rows": rows, "table": table}


def queue_sweep(lib, args):
    """Whole-list encode + decode wall time of a mixed-page queue for each
    (chunk code-blocks, engines) pair; the fastest pair is written out."""
    pages = list(synth.mixed_pages(args.queue_pages, seed=args.mixed_seed + 1))
    mpx = sum(p.shape[0] * p.shape[1] for p, _ in pages) / 1e6
    log(f"queue sweep: {len(pages)} mixed pages, {mpx:.0f} Mpx")
    lib.jp2g_set_t1_mode(0, 0)
    lib.jp2g_set_queue(0, 0)
    ref = j.encode_many(pages)
    streams = [r for r in ref if not isinstance(r, Exception)]
    j.decode_many(streams)  # grows every engine the sweep uses
    rows = []
    for eng in [int(x) for x in args.queue_engines.split(",")]:
        for chunk in [int(x) for x in args.queue_chunks.split(",")]:
            lib.jp2g_set_queue(chunk, eng)
            best_e = best_d = math.inf
            for _ in range(2):
                t0 = time.perf_counter()
                out = j.encode_many(pages)
                t1 = time.perf_counter()
                j.decode_many(streams)
                t2 = time.perf_counter()
                if out != ref:
                    raise SystemExit(f"queue chunk {chunk} engines {eng}: output differs")
                best_e, best_d = min(best_e, t1 - t0), min(best_d, t2 - t1)
            rows.append({"chunk_blocks": chunk, "engines": eng, "encode_s": round(best_e, 4),
                         "decode_s": round(best_d, 4), "encode_mpx_s": round(mpx / best_e, 1)})
            log(f"queue chunk {chunk} engines {eng}: encode {best_e:.3f} s, decode {best_d:.3f} s")
    lib.jp2g_set_queue(0, 0)
    best = min(rows, key=lambda r: r["encode_s"] + r["decode_s"])
    Path(args.out_queue_inc).write_text(
        "// Generated by jp2lam_amd/autotune.py on an MI355X; do not edit by hand.\n"
        "// Queue chunking for jp2lam_encode_many / jp2lam_decode_many: code-blocks\n"
        f"// per device batch and engines running batches at once ({len(pages)} mixed\n"
        f"// pages, {mpx:.0f} Mpx: encode {best['encode_s']:.3f} s, decode {best['decode_s']:.3f} s).\n"
        f"constexpr uint32_t kQueueChunkBlocks = {best['chunk_blocks']};\n"
        f"constexpr int kQueueEngines = {best['engines']};\n")
    log(f"wrote {args.out_queue_inc}")
    return {"pages": len(pages), "megapixels": round(mpx, 1), "rows": rows, "chosen": best}


def _blocks_of(px, q) -> int:
    """Code-block count of one page (layout.rs:37-173 geometry)."""
    h, w = px.shape[:2]
    nc = 1 if px.ndim == 2 else 3
    levels = _levels(w, h, nc)
    cb = 32 if q < 30 else 64
    rw, rh = [w], [h]
    for _ in range(levels):
        rw.insert(0, (rw[0] + 1) // 2)
        rh.insert(0, (rh[0] + 1) // 2)
    n = math.ceil(rw[0] / cb) * math.ceil(rh[0] / cb)
    for r in range(1, levels + 1):
        lw, lh, fw, fh = rw[r - 1], rh[r - 1], rw[r], rh[r]
        for bw, bh in ((fw - lw, lh), (lw, fh - lh), (fw - lw, fh - lh)):
            if bw > 0 and bh > 0:
                n += math.ceil(bw / cb) * math.ceil(bh / cb)
    return n * nc


def _levels(w, h, nc) -> int:
    m = min(w, h)
    if m <= 1:
        lv = 0
    elif m < 32:
        lv = m.bit_length() - 1
    else:
        lv = min(1 + (m // 32).bit_length() - 1, 6)
    return min(lv, 5 if nc == 1 else 6)


def choose(rows):
    """Per size class, the config with the lowest geometric-mean slowdown
    against each workload's best."""
    buckets = sorted({r["bucket"] for r in rows})
    table = []
    for b in buckets:
        rs = [r for r in rows if r["bucket"] == b]
        score = {}
        for cname, _, _ in CONFIGS:
            logs = [math.log(r["t1_ms"][cname] / min(r["t1_ms"].values())) for r in rs]
            score[cname] = math.exp(sum(logs) / len(logs))
        best = min(score, key=score.get)
        mode, frac = next((m, f) for c, m, f in CONFIGS if c == best)
        table.append({"bucket": b, "config": best, "coder": mode, "frac": frac,
                      "slowdown": round(score[best], 4), "workloads": [r["workload"] for r in rs]})
    return table


def write_inc(path: Path, table) -> None:
    lines = ["// Generated by jp2lam_amd/autotune.py on an MI355X; do not edit by hand.",
             "// Tier-1 coder per size class (bit length of the batch's code-block",
             "// count): {bucket, coder, hybrid split fraction}.  A batch uses the entry",
             "// with the largest bucket not above its own; below the first entry the",
             "// wave coder runs.  Workloads behind each entry are listed beside it."]
    lines.append("static const T1TableEntry kT1Table[] = {")
    for e in table:
        lines.append(f"    {{{e['bucket']}, {CODER_ID[e['coder']]}, {e['frac']:.2f}f}},  "
                     f"// {e['config']}: {', '.join(e['workloads'])}")
    lines.append("};")
    path.write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
