"""CPU model of the Tier-1 wave schedulers (tests/host/claim_model.cpp).

Every interleaving of the waves' atomic steps is explored, with in-order
dispatch into a bounded number of resident slots:

- the shipped hand-out (t1_pick_block: per-SIMD arrival counter, heavy
  positions from the front for a SIMD's first wave unless it decides for a
  light one after its bounded wait, light ones from the back for every other
  wave) codes every block exactly once and never blocks, for every resident
  count, including fewer slots than waves, and for either decision;
- a persistent claim loop without waits does the same;
- the exclusive-SIMD form, in which a later wave on a SIMD waits for the
  first wave's heavy/light publication, hangs in some interleavings (a first
  wave that finds the queue empty exits without publishing).
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent
SRC = ROOT / "host" / "claim_model.cpp"


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("claim") / "claim_model"
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-o", str(exe), str(SRC)], check=True)
    return exe


def _run(exe, proto, waves, blocks, simds, resident):
    r = subprocess.run([str(exe), proto, str(waves), str(blocks), str(simds), str(resident)],
                       capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout.strip()


@pytest.mark.parametrize("cfg", [(4, 4, 2, 4), (4, 4, 2, 2), (4, 4, 2, 1), (6, 6, 2, 3), (6, 6, 3, 2), (8, 8, 4, 3),
                                 (6, 6, 2, 6), (7, 7, 2, 7), (7, 7, 3, 5)])
def test_shipped_hand_out_codes_every_block_once(model, cfg):
    rc, out = _run(model, "pick", *cfg)
    assert rc == 0 and out.startswith("ok"), out


@pytest.mark.parametrize("cfg", [(4, 4, 2, 4), (4, 4, 2, 2), (6, 6, 2, 3), (6, 6, 3, 2), (5, 3, 2, 2), (3, 7, 2, 1)])
def test_wait_free_claim_loop_never_hangs(model, cfg):
    rc, out = _run(model, "claim", *cfg)
    assert rc == 0 and out.startswith("ok"), out


@pytest.mark.parametrize("cfg", [(4, 4, 2, 4), (4, 4, 2, 2), (6, 6, 3, 2)])
def test_publication_wait_can_hang(model, cfg):
    rc, out = _run(model, "mark", *cfg)
    assert rc == 1 and out.startswith("hang"), out


def test_model_reports_double_coding(model):
    """The checker itself: the hand-out with its counter increments split
    into a read and a write (a planted mistake) codes one position twice and
    leaves another uncoded in some interleaving, and the explorer says so."""
    rc, out = _run(model, "split", 4, 4, 2, 4)
    assert rc == 2 and out.startswith("wrong") and ("coded 0x" in out or "coded 2x" in out), out
    rc, _ = _run(model, "pick", 5, 3, 2, 2)
    assert rc == 3  # the hand-out launches one wave per block
