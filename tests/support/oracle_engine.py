"""Test double of the engine module contract, backed by the CPU oracle (tests only).

Lets the block driver's checkpoint / resume / pause logic run on a machine without a GPU:
each _launch scans its outer-index range with the oracle's C scan."""
from __future__ import annotations

import numpy as np

import es_oracle as O
from es_plan import TUNABLES, Plan, plan_wheel  # noqa: F401

LAUNCH_LOG = []


def free_vram_mib():
    return None


def load_known(*paths):
    return {}


class _Arr:
    def __init__(self, data):
        self.data = data

    def fill(self, v):
        self.data[:] = v

    def get(self):
        return np.array(self.data)

    def set(self, arr):
        self.data[:len(arr)] = np.asarray(arr, dtype=np.uint32)

    def __getitem__(self, sl):
        return _Arr(self.data[sl])


class GpuSieve:
    def __init__(self, plan):
        self.plan = plan
        self.rest_q = [plan.info[p][0] for p in plan.rest]
        self.rest_Mmodq = [plan.M % q for q in self.rest_q]
        self.tconst_h = np.zeros(max(1, len(self.rest_q)), dtype=np.uint32)
        self.tconst = _Arr(np.zeros(max(1, len(self.rest_q)), dtype=np.uint32))
        self.out_cap = int(plan.tun["max_survivors"])
        self.out = _Arr(np.zeros(self.out_cap, dtype=np.uint64))
        self.out_n = _Arr(np.zeros(1, dtype=np.uint64))
        self.r0 = plan.radix[0] if plan.radix else 1
        self.R_outer = plan.R // self.r0
        self.ochunk = max(1, int(plan.tun.get("test_ochunk", 64)))
        self.launches = 0
        self.residues_checked = 0

    def _t(self):
        # recover t from tconst of the first filter with q coprime to M is not possible in
        # general; the driver always sets tconst from t, so remember the block via the table
        for t in range(0, 1 << 16):
            if all(((t % q) * mq) % q == int(c) for q, mq, c in
                   zip(self.rest_q, self.rest_Mmodq, self.tconst.data)):
                return t
        raise AssertionError("unknown block")

    def _launch(self, ostart, ocount):
        t = self._t()
        lo, hi = ostart * self.r0, (ostart + ocount) * self.r0
        n, ids = O.c_scan(self.plan.k, self.plan.wheel, self.plan.rest, t, lo, hi, threads=1,
                          cap=self.out_cap)
        c = int(self.out_n.data[0])
        take = max(0, min(n, self.out_cap - c))
        self.out.data[c:c + take] = ids[:take]
        self.out_n.data[0] = c + n
        self.launches += 1
        LAUNCH_LOG.append((t, ostart, ocount))
