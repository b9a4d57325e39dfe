"""Statistics of long runs of chaotic networks, for comparisons with the FP64 oracle.

A run's population rate, mean ISI CV and final distribution of plastic weights are one
realisation: once rounding has made two runs differ by one spike, their spike trains
decorrelate.  Over 2 s of the COBA network, moving every resting potential of the FP64
oracle by one float32 ulp changes the rate by up to 1 % and the ISI CV by up to 2.6 % per
connectivity seed, so single runs cannot be held to 2 %; ensemble means over seeds can.
"""
import numpy as np


def isi_cv(ids, times, n):
    """Mean over neurons with >= 3 spikes of std/mean of their inter-spike intervals."""
    cvs = []
    order = np.lexsort((times, ids))
    ids, times = ids[order], times[order].astype(np.float64)
    starts = np.searchsorted(ids, np.arange(n))
    ends = np.searchsorted(ids, np.arange(n), side="right")
    for a, b in zip(starts, ends):
        if b - a >= 3:
            isi = np.diff(times[a:b])
            cvs.append(isi.std() / isi.mean())
    return float(np.mean(cvs))


def histogram(w, edges):
    """Counts per bin of `edges`; values outside the edges count in the end bins."""
    w = np.clip(np.asarray(w, np.float64), edges[0], edges[-1])
    return np.histogram(w, bins=edges)[0]


def hist_l1(a, b):
    """Half the L1 distance of two histograms as a fraction of their (equal) totals."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.abs(a - b).sum() / (2.0 * a.sum()))


def round_bf16(x):
    """float32 -> bfloat16 (nearest, ties to even) -> float32, for finite values."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)
