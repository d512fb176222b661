"""The ziggurat's fast test on the rare draws where it ties on the high word.

The GPU samplers decide the fast test (normal.rs:26-29, equivalently
(u >> 11) < T[l]) on the draw's high 32 bits against T[l] >> 21; when those
are equal -- one attempt in 2^32 -- they fall back to the exact 53-bit test.
tests/golden/zig_ties.json lists stream positions of such ties (found by
tests/golden/find_zig_ties.c with the oracle's stream and ziggurat): seed 42,
block 4096, three accepted and two rejected by the exact test.  Random draws
never reach this code, so these positions pin it: the loop-form sampler and
the steppers' queue schedule (lane-mask state machine, rings, deferred slow
block; the tie path among 63 ordinary lanes) must both give the oracle's
normals through and past the tie.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

R_ZIG = 3.654152885361008796
_FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "zig_ties.json")))
TIES = _FIX["ties"]
SEED, BLOCK = _FIX["seed"], _FIX["block"]


def _ulps(a, b):
    return np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))


def _check(dev, ref, prec, what):
    if prec == 64:
        tail = np.abs(ref) > R_ZIG
        np.testing.assert_array_equal(dev[~tail], ref[~tail], err_msg=what)
        assert np.all(_ulps(dev[tail], ref[tail]) <= 2), what
    else:
        np.testing.assert_allclose(dev, ref, rtol=2.4e-7, atol=1e-9, err_msg=what)


def test_fixture_holds_both_outcomes():
    assert {t["accepted"] for t in TIES} == {True, False}


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("tie", TIES, ids=lambda t: f"path{t['path']}")
def test_loop_sampler_through_a_tie(gpu, orc, tie, prec):
    n = tie["normal"] + 3
    ref, used = orc.normals(SEED, BLOCK, tie["path"], n)
    dev, dev_used = gpu.stream_normals(SEED, BLOCK, tie["path"], 1, n, precision=prec)
    assert int(dev_used[0]) == used
    _check(dev[0], ref, prec, f"path {tie['path']}")


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("tie", TIES, ids=lambda t: f"path{t['path']}")
def test_queue_schedule_through_a_tie(gpu, orc, tie, prec):
    # one wavefront of 64 paths with the tie's path as lane 17
    n = tie["normal"] + 3
    p0 = tie["path"] - 17
    dev = gpu.stream_normals_queue(SEED, BLOCK, p0, 64, n, precision=prec)
    for lane in range(64):
        ref, _ = orc.normals(SEED, BLOCK, p0 + lane, n)
        _check(dev[lane], ref, prec, f"path {p0 + lane}")


@pytest.mark.parametrize("prec", [64, 32])
def test_queue_schedule_equals_loop(gpu, prec):
    # ordinary positions: 300 paths (a ragged last wavefront) x 500 normals
    loop, _ = gpu.stream_normals(7, 2552, 1000, 300, 500, precision=prec)
    queue = gpu.stream_normals_queue(7, 2552, 1000, 300, 500, precision=prec)
    np.testing.assert_array_equal(queue, loop)
