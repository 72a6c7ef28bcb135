"""NumPy FP64 restatement of the reference's observation-pair lists and run-end global BA.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this module to
check or time the product path; the product package (xclean_amd/) never imports it.

Functions and the reference code they restate (paths relative to the reference's src/xclean/):
  build_pairs        backend/ba.py:902-927
  (the global-BA functions follow below with their own citations)

Pins (tests/test_gba_oracle.py): build_pairs against a brute-force loop over every landmark's observation pairs.
"""
from __future__ import annotations

import numpy as np


def build_pairs(olm_sorted, nl):
    """backend/ba.py:902-927: (offs int32 (nl + 1), pa int32, pb int32) - every ordered observation pair (a, b) inside
    each landmark's run of the landmark-sorted observation list, landmark-major, a-major then b."""
    olm = np.asarray(olm_sorted, np.int64).ravel()
    offs = np.searchsorted(olm, np.arange(nl + 1), side="left").astype(np.int64)
    c = np.diff(offs)
    c2 = c * c
    total = int(c2.sum())
    if total == 0:
        return offs.astype(np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32)
    owner = np.repeat(np.arange(nl), c2)
    first = np.concatenate([[0], np.cumsum(c2)[:-1]])
    t = np.arange(total) - first[owner]
    pa = offs[owner] + t // c[owner]
    pb = offs[owner] + t % c[owner]
    return offs.astype(np.int32), pa.astype(np.int32), pb.astype(np.int32)
