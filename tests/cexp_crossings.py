"""Draws next to every integer crossing of capped_exponential's truncated value.

capped_exponential (chromo/util/bead_selection.pyx:58-65) returns
int(-log10(x / RAND_MAX + 1e-5) * window * 0.45 + 1.0001) for a 31-bit draw x.
The value falls as x grows and crosses the integer j at
x_j = RAND_MAX * (10 ** (-(j - 1.0001) / (0.45 window)) - 1e-5).  Only draws
next to some x_j can truncate differently under two log10 implementations that
differ by an ulp, so comparing every draw within `radius` of every crossing, for
every window up to `wmax`, compares the two truncations on all 2^31 draws.
"""
import numpy as np

RAND_MAX = 2147483647


def crossing_draws(windows, radius=40):
    ds, ws = [], []
    for w in windows:
        jmax = int(np.floor(2.25 * w + 1.0001))
        j = np.arange(2, jmax + 1, dtype=np.float64)
        u = 10.0 ** (-(j - 1.0001) / (0.45 * w)) - 1e-5
        k = np.round(u * RAND_MAX).astype(np.int64)
        d = (k[:, None] + np.arange(-radius, radius + 1)[None, :]).reshape(-1)
        d = np.unique(d[(d >= 0) & (d <= RAND_MAX)])
        ds.append(d)
        ws.append(np.full(d.size, w, dtype=np.int64))
    return np.concatenate(ds).astype(np.int32), np.concatenate(ws)
