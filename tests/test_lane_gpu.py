"""The device as a dispatch lane (WorkerLane::render_chunk,
indicatrix-dispatch/src/lane.rs:113-142) driven the way the reference's
LanePool drives its lanes: disjoint chunks claimed from a cursor, chunk sums
merged in chunk-start order, a failed chunk's untraced tail requeued for
another lane.  The merged image must equal one render of the whole range;
every partial result must be exactly the prefix it reports."""
import ctypes as C
import threading
import time

import numpy as np
import pytest

from dispatch_mirror import Merger, SampleCursor
from indicatrix_amd import tracer, workloads

pytestmark = pytest.mark.gpu


def _scene():
    return workloads.film_c1_scene(32, 24, 12)


def _render(backend, sc, first, count):
    acc = np.zeros((sc.width * sc.height, 3), np.float64)
    assert backend.try_accumulate(sc, first, count, acc)
    return acc


def _close(a, b, rtol=2e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert np.all(np.abs(a - b) <= rtol * np.abs(b) + 1e-6), float(np.max(np.abs(a - b)))


def test_whole_chunk_equals_one_render(backend):
    sc = _scene()
    lane = tracer.GpuLane(backend)
    r = lane.render_chunk(sc, 100, 300)
    assert r.complete and r.done == 300 and r.sum.dtype == np.float32
    _close(r.sum, _render(backend, sc, 100, 300))
    assert r.rate is not None and r.rate > 0.0
    # the converged sub-batch seeds the next chunk of the scene
    r2 = lane.render_chunk(sc, 400, 50)
    assert r2.complete and r2.done == 50
    _close(r2.sum, _render(backend, sc, 400, 50))


def test_two_lanes_merge_to_one_render(backend):
    sc = _scene()
    b2 = tracer.GpuBackend()
    try:
        lanes = [tracer.GpuLane(backend, "gpu lane 0"), tracer.GpuLane(b2, "gpu lane 1")]
        cursor = SampleCursor(0, 600)
        merger = Merger(sc.width * sc.height)

        def run(lane, want):
            while True:
                claim = cursor.claim_any(want)
                if claim is None:
                    return
                first, count = claim
                r = lane.render_chunk(sc, first, count)
                assert r.done <= count
                if r.done:
                    merger.add(first, r.done, r.sum)
                cursor.requeue(first + r.done, count - r.done)

        ts = [threading.Thread(target=run, args=(lane, want)) for lane, want in zip(lanes, (48, 80))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        merged, total = merger.into_parts()
        assert total == 600 and cursor.fully_claimed()
        ranges = merger.ranges()
        assert ranges[0][0] == 0 and all(a + n == b for (a, n), (b, _) in zip(ranges, ranges[1:]))
        _close(merged, _render(backend, sc, 0, 600), rtol=1e-5)
    finally:
        b2.close()


def test_failed_chunk_keeps_its_prefix_and_the_tail_is_requeued(backend):
    sc = _scene()
    lane = tracer.GpuLane(backend)
    backend.inject_fault(3)  # the fourth sub-batch reports a device loss
    r = lane.render_chunk(sc, 0, 400)
    assert not r.complete and 0 < r.done < 400
    assert backend.is_lost()
    assert backend.try_recover()
    _close(r.sum, _render(backend, sc, 0, r.done))
    # the pool requeues [done, 400) and another chunk completes it
    cursor = SampleCursor(0, 0)
    cursor.requeue(r.done, 400 - r.done)
    first, count = cursor.claim_any(1000)
    r2 = lane.render_chunk(sc, first, count)
    assert r2.complete and r2.done == count
    merger = Merger(sc.width * sc.height)
    merger.add(0, r.done, r.sum)
    merger.add(first, r2.done, r2.sum)
    merged, total = merger.into_parts()
    assert total == 400
    _close(merged, _render(backend, sc, 0, 400), rtol=1e-5)


def test_cancel_before_and_during_a_chunk(backend):
    sc = _scene()
    lane = tracer.GpuLane(backend)
    flag = (C.c_uint8 * 1)(1)
    r = lane.render_chunk(sc, 0, 64, cancel=flag)
    assert r.error == "cancelled" and r.done == 0 and not np.any(r.sum)
    big = workloads.film_c1_scene(128, 128, 16)
    flag[0] = 0
    timer = threading.Timer(0.3, lambda: flag.__setitem__(0, 1))
    timer.start()
    t0 = time.perf_counter()
    r = lane.render_chunk(big, 0, 100000, cancel=flag)
    timer.cancel()
    assert time.perf_counter() - t0 < 20.0
    assert r.error == "cancelled" and 0 < r.done < 100000
    _close(r.sum, _render(backend, big, 0, r.done))
