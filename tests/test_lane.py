"""CPU checks of the dispatch-lane boundary: the sub-batch sizing rule
(ported from apps/indicatrix-worker/src/stream_emit/tests/batch_sizing.rs),
ixg_render_chunk's argument checks, and the test-side claim point and merge
against the reference's own SampleCursor cases (sample_cursor/tests.rs).
No GPU calls."""
import ctypes as C

import numpy as np

from dispatch_mirror import Merger, SampleCursor
from indicatrix_amd import _lib, abi, tracer, workloads

MAX_SUBBATCH = 2048


def test_next_subbatch_grows_when_well_under_budget():
    assert tracer.next_subbatch(4, 0.010) > 4


def test_next_subbatch_shrinks_when_over_budget():
    nxt = tracer.next_subbatch(100, 0.400)
    assert 1 <= nxt < 100
    assert nxt == 25  # 100 x (100 ms / 400 ms)


def test_next_subbatch_never_returns_zero():
    assert tracer.next_subbatch(0, 0.001) >= 1
    assert tracer.next_subbatch(1, 10.0) >= 1


def test_next_subbatch_never_exceeds_the_cap_even_under_adversarial_timing():
    batch = 1
    for _ in range(40):
        batch = tracer.next_subbatch(batch, 1e-9)
        assert batch <= MAX_SUBBATCH
    assert batch == MAX_SUBBATCH
    assert tracer.next_subbatch(7, 0.0) == 28  # an unmeasurable time grows x4
    assert tracer.next_subbatch(50, 0.100) == 50  # on target: unchanged


def test_render_chunk_rejects_bad_arguments():
    sc = workloads.film_c1_scene(8, 8)
    out = np.zeros((64, 3), np.float32)
    done = C.c_uint32(7)
    L = _lib.lib()
    assert L.ixg_render_chunk(None, sc.ref(), 0, 4, 0, out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(done),
                              None, None, None) == abi.IXG_INVALID
    assert L.ixg_render_chunk(C.c_void_p(1), sc.ref(), 2 ** 32 - 2, 4, 0,
                              out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(done), None, None,
                              None) == abi.IXG_INVALID
    assert L.ixg_render_chunk(C.c_void_p(1), sc.ref(), 0, 4, 0, None, C.byref(done), None, None,
                              None) == abi.IXG_INVALID


def test_cursor_mirror_reference_cases():
    c = SampleCursor(0, 100)
    assert c.claim(1000) == (0, 100) and c.claim(1) is None
    c = SampleCursor(0, 100)
    c.claim(90)
    assert c.claim(30) == (90, 10) and c.claim(30) is None
    c = SampleCursor(500, 600)
    assert c.claim(50) == (500, 50) and c.claim(50) == (550, 50) and c.claim(50) is None
    c = SampleCursor(0, 100)
    assert c.claim(0) is None and c.claim(100) == (0, 100)
    c = SampleCursor(0, 20)
    c.requeue(0, 8)
    c.requeue(3, 0)
    c.claim(0)
    assert c.claim_any(8) == (0, 8)
    c = SampleCursor(0, 20)
    c.claim(5)
    c.requeue(0, 5)
    assert c.claim_any(2) == (0, 2) and c.claim_any(10) == (2, 3) and c.claim_any(10) == (5, 10)
    assert c.claim_any(10) == (15, 5) and c.fully_claimed() and c.claim_any(10) is None


def test_merger_mirror_orders_and_zeroes():
    m = Merger(2)
    m.add(8, 4, np.array([[1, 2, 3], [np.nan, 1, 1]], np.float32))
    m.add(0, 8, np.array([[1, 1, 1], [2, 2, 2]], np.float32))
    merged, total = m.into_parts()
    assert total == 12 and m.ranges() == [(0, 8), (8, 4)]
    assert np.array_equal(merged, np.array([[2, 3, 4], [2, 2, 2]], np.float32))
