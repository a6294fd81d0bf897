"""A small restatement of the coordinator side of the dispatch contract, for
the lane tests: the claim point that hands out disjoint absolute sample
ranges (SampleCursor, indicatrix-dispatch/src/sample_cursor/mod.rs: claim,
claim_any over requeued tails, requeue) and the merge of chunk sums in
chunk-start order (Merger, indicatrix-dispatch/src/merge/mod.rs:134-275:
add parks a chunk by its first sample after zeroing non-finite pixels,
into_parts folds the parked chunks in ascending first-sample order).  Test
infrastructure: a GPU lane under test is driven through it exactly as the
reference's LanePool drives a WorkerLane."""
import threading
from collections import deque

import numpy as np


class SampleCursor:
    def __init__(self, start, end):
        self._next = start
        self.end = end
        self._requeued = deque()
        self._lock = threading.Lock()

    def claim(self, want):
        """sample_cursor/mod.rs claim: (start, count) of fresh work or None."""
        if want == 0:
            return None
        with self._lock:
            if self._next >= self.end:
                return None
            start = self._next
            self._next = min(self._next + want, self.end)
            return start, self._next - start

    def claim_any(self, want):
        """claim_any: a requeued range first (split to want), else fresh work."""
        if want == 0:
            return None
        with self._lock:
            if self._requeued:
                start, count = self._requeued.popleft()
                if count > want:
                    self._requeued.appendleft((start + want, count - want))
                    return start, want
                return start, count
        return self.claim(want)

    def requeue(self, start, count):
        if count:
            with self._lock:
                self._requeued.append((start, count))

    def fully_claimed(self):
        with self._lock:
            return self._next >= self.end and not self._requeued


class Merger:
    def __init__(self, pixels):
        self.pixels = pixels
        self._parked = {}
        self._lock = threading.Lock()

    def add(self, first_sample, done, chunk_sum):
        """merge/mod.rs:134-152: park a chunk's sum by its first sample."""
        s = np.array(chunk_sum, np.float32).reshape(self.pixels, 3)
        bad = ~np.all(np.isfinite(s), axis=1)
        s[bad] = 0.0
        with self._lock:
            assert first_sample not in self._parked, "chunks must be disjoint"
            self._parked[first_sample] = (done, s)

    def into_parts(self):
        """merge/mod.rs:197-213: the f32 sum of every parked chunk, folded in
        ascending first-sample order, and the total sample count."""
        merged = np.zeros((self.pixels, 3), np.float32)
        total = 0
        for first in sorted(self._parked):
            done, s = self._parked[first]
            merged += s
            total += done
        return merged, total

    def ranges(self):
        return sorted((f, d) for f, (d, _) in self._parked.items())
