"""Batched sieving of many (plan, block, window) jobs in shared launches.

The reference runs one plan per process and one kernel launch per chunk
(sieve/gpu_sieve_v9.py:500-506, harness/grind.py); here many k values -- whose wheels differ in
geometry and table sizes -- and many blocks share launches: the jobs' strips are consecutive
workgroup ranges of one grid per kernel shape (es_batch_run in libes_sieve.so).  Survivors of a
job are those es_engine_scan_window returns for the same engine, block and window.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import es_native as nat


class Batch:
    def __init__(self, out_cap_per_job: int = 1 << 16):
        h = C.c_void_p()
        nat.check(nat.lib().es_batch_create(out_cap_per_job, C.byref(h)))
        self._h = h
        self._engines = []          # keep engines alive while the batch exists
        self.jobs = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                nat.lib().es_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def add(self, sieve, t: int, idx_lo: int, idx_hi: int) -> int:
        """Queue block t, positions [idx_lo, idx_hi) of a GpuSieve; returns the job id."""
        jid = C.c_uint32(0)
        nat.check(nat.lib().es_batch_add(self._h, sieve._h, int(t), int(idx_lo), int(idx_hi),
                                         C.byref(jid)))
        self._engines.append(sieve)
        self.jobs.append((sieve, t, idx_lo, idx_hi))
        return jid.value

    def run(self):
        """Sieve every job; returns (device ms, launches)."""
        ms = C.c_double(0)
        la = C.c_uint32(0)
        nat.check(nat.lib().es_batch_run(self._h, C.byref(ms), C.byref(la)))
        return ms.value, la.value

    def capture(self) -> int:
        """Record the counter reset and every launch of the current jobs into a hipGraph;
        returns the number of kernel launches captured."""
        la = C.c_uint32(0)
        nat.check(nat.lib().es_batch_capture(self._h, C.byref(la)))
        return la.value

    def set_block(self, job: int, t: int) -> None:
        """Move a job to block t (its window of positions is kept)."""
        nat.check(nat.lib().es_batch_set_block(self._h, job, int(t)))
        s, _, lo, hi = self.jobs[job]
        self.jobs[job] = (s, t, lo, hi)

    def set_blocks(self, ts) -> None:
        """Move every job i to block ts[i] in one upload."""
        arr = np.ascontiguousarray(np.asarray(ts, dtype=np.uint64))
        assert arr.size == len(self.jobs)
        nat.check(nat.lib().es_batch_set_blocks(self._h, nat.ptr64(arr), int(arr.size)))
        self.jobs = [(s, int(t), lo, hi) for (s, _, lo, hi), t in zip(self.jobs, arr)]

    def replay(self) -> float:
        """Run the captured graph; device ms."""
        ms = C.c_double(0)
        nat.check(nat.lib().es_batch_replay(self._h, C.byref(ms)))
        return ms.value

    def sweep(self, blocks) -> list[list[list[int]]]:
        """Survivors of every job for each block t in `blocks` (every job moved to t, then one
        graph replay per block): result[i][job] for blocks[i]."""
        out = []
        for t in blocks:
            self.set_blocks([t] * len(self.jobs))
            self.replay()
            out.append([self.collect(j) for j in range(len(self.jobs))])
        return out

    def collect(self, job: int) -> list[int]:
        n = C.c_uint64(0)
        nat.check(nat.lib().es_batch_collect(self._h, job, None, 0, C.byref(n)))
        buf = np.zeros(max(1, n.value), dtype=np.uint64)
        nat.check(nat.lib().es_batch_collect(self._h, job, nat.ptr64(buf), n.value, C.byref(n)))
        return sorted(int(x) for x in buf[:n.value])
