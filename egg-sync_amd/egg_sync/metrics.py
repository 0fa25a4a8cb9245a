"""Partition comparison used to judge a learned solve against the plain one."""
from __future__ import annotations

import numpy as np


def same_partition(a, b) -> bool:
    """True when two labelings induce the same partition (labels may be renamed;
    the noise label -1 must match exactly)."""
    a = np.asarray(a).astype(np.int64)
    b = np.asarray(b).astype(np.int64)
    if a.shape != b.shape or not np.array_equal(a < 0, b < 0):
        return False
    m = a >= 0
    if not m.any():
        return True
    pairs = np.unique(np.stack([a[m], b[m]], 1), axis=0)
    return len(np.unique(pairs[:, 0])) == len(pairs) and len(np.unique(pairs[:, 1])) == len(pairs)
