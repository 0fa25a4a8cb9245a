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


def disagreement(a, b) -> dict:
    """How far two labelings are apart: points whose noise status differs, and
    points outside the best one-to-one matching of clusters (each cluster of
    `a` matched to the cluster of `b` it shares most points with)."""
    a = np.asarray(a).astype(np.int64)
    b = np.asarray(b).astype(np.int64)
    noise = int(np.sum((a < 0) != (b < 0)))
    m = (a >= 0) & (b >= 0)
    pairs, counts = np.unique(np.stack([a[m], b[m]], 1), axis=0, return_counts=True)
    order = np.argsort(-counts)
    used_a, used_b, matched = set(), set(), 0
    for k in order:
        pa, pb = int(pairs[k, 0]), int(pairs[k, 1])
        if pa not in used_a and pb not in used_b:
            used_a.add(pa)
            used_b.add(pb)
            matched += int(counts[k])
    return {"noise_differs": noise, "unmatched": int(m.sum()) - matched,
            "clusters_a": int(len(np.unique(a[a >= 0]))),
            "clusters_b": int(len(np.unique(b[b >= 0])))}
