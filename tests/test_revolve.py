"""The binomial checkpoint schedule of the adjoint's tape (host only, no GPU):
its forward-step count equals the minimum of an exhaustive dynamic program over
every split, and the known binomial figure for 20 snapshots over 2000 steps."""
from functools import lru_cache
import sys

import numpy as np

from dgcom_amd import swe2d

sys.setrecursionlimit(10000)


@lru_cache(None)
def T(l, s):
    """Forward steps to reverse l steps from a stored start with s free slots."""
    if l <= 1:
        return 0
    if s == 0:
        return l * (l - 1) // 2
    return min(m + T(l - m, s - 1) + T(m, s) for m in range(1, l))


@lru_cache(None)
def G(l, s):
    """The same when the right-most checkpoints were taken by the first sweep."""
    if l <= 1:
        return 0
    if s == 0:
        return l * (l - 1) // 2
    return min(T(m, s) + G(l - m, s - 1) for m in range(1, l))


def test_schedule_is_optimal_on_small_cases():
    for n in range(1, 45):
        for snaps in range(1, 6):
            assert swe2d.revolve_count(n, snaps) == n + G(n, snaps - 1), (n, snaps)


def _dp(n, S):
    Tt = np.zeros((S + 1, n + 1), dtype=np.int64)
    Gt = np.zeros_like(Tt)
    ll = np.arange(n + 1)
    Tt[0] = ll * (ll - 1) // 2
    Gt[0] = Tt[0]
    for s in range(1, S + 1):
        for L in range(2, n + 1):
            m = np.arange(1, L)
            Tt[s, L] = (m + Tt[s - 1, L - m] + Tt[s, m]).min()
            Gt[s, L] = (Tt[s, m] + Gt[s - 1, L - m]).min()
    return n + int(Gt[S, n])


def test_schedule_is_optimal_on_larger_cases():
    for n, snaps in ((300, 4), (700, 6)):
        assert swe2d.revolve_count(n, snaps) == _dp(n, snaps - 1)


def test_binomial_figures():
    # 20 snapshots, 2000 steps: repetition number 4, C(24,3) = 2024 saved over 4 x 2000
    assert swe2d.revolve_count(2000, 20) == 5980
    assert swe2d.revolve_count(2000, 2001) == 2000      # every state kept
    assert swe2d.revolve_count(10, 1) == 10 + 45        # one slot: restart from the start for every step
