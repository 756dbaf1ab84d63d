"""Pins of the oracle's row order by tap mask (CPU only).

The order is fixed by its definition: a permutation whose rows, window by window, have
non-increasing tap masks with ties in row order.  Brute force builds it with Python's
``sorted`` on an explicit key; a second check plants a wrong tie rule in a copy of the
oracle's result and sees it rejected.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle


def _random_map(rng, n, K, density):
    m = np.where(rng.random((n, K)) < density, rng.integers(0, 1000, (n, K)), -1).astype(np.int32)
    return m


def _brute(kmap, window, bits):
    n = kmap.shape[0]
    masks = [sum(1 << k for k in range(bits) if kmap[r, k] >= 0) for r in range(n)]
    key = (lambda r: (r // window, -masks[r], r)) if window > 0 else (lambda r: (-masks[r], r))
    return np.array(sorted(range(n), key=key), dtype=np.int32)


@pytest.mark.parametrize("K,density", [(27, 0.3), (8, 0.125), (8, 0.6), (1, 0.5), (33, 0.4)])
@pytest.mark.parametrize("window", [0, 4, 16, 1000])
def test_row_order_brute_force(K, density, window):
    rng = np.random.default_rng(K * 100 + window)
    kmap = _random_map(rng, 300, K, density)
    got = oracle.row_order(kmap, window)
    bits = min(K, 27 if window > 0 else 31)
    assert np.array_equal(got, _brute(kmap, window, bits))
    assert np.array_equal(np.sort(got), np.arange(300))


def test_row_order_one_tap_rows_group_by_tap():
    """Rows with exactly one tap (the fine-to-coarse map of a 2x2x2 stride-2 kernel) come
    out grouped by tap, lowest tap last, each group in row order."""
    rng = np.random.default_rng(5)
    tap = rng.integers(0, 8, 500)
    kmap = np.full((500, 8), -1, np.int32)
    kmap[np.arange(500), tap] = 7
    got = oracle.row_order(kmap)
    expect = np.concatenate([np.nonzero(tap == t)[0] for t in range(7, -1, -1)])
    assert np.array_equal(got, expect)


def test_row_order_rejects_a_wrong_tie_rule():
    rng = np.random.default_rng(9)
    kmap = _random_map(rng, 200, 8, 0.2)  # many equal masks
    good = oracle.row_order(kmap, 0)
    masks = ((kmap >= 0).astype(np.int64) << np.arange(8)).sum(1)
    bad = np.lexsort((-np.arange(200), -masks)).astype(np.int32)  # ties in reverse row order
    assert not np.array_equal(good, bad)
    assert np.array_equal(good, _brute(kmap, 0, 8))
