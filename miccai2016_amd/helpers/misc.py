"""Host-side helpers (reference: helpers/misc.py)."""
import numpy as np


def _source_range(x, perc_low, perc_up):
    """(lo, hi, from_percentiles): the interval of x that is mapped onto the
    destination, either x's extremes (in x's own dtype) or the two requested
    percentiles, a missing one standing for the 0th or the 100th, taken in a
    single np.percentile call so that x is sorted once."""
    if perc_low is None and perc_up is None:
        return np.min(x), np.max(x), False
    q = np.percentile(x, [0.0 if perc_low is None else perc_low,
                          100.0 if perc_up is None else perc_up])
    return q[0], q[1], True


def normalize(data, dst_min=0.0, dst_max=1.0, perc_low=None, perc_up=None, clip=True):
    """Affine map of data onto [dst_min, dst_max], as a new float64 array
    (the reference's normalize, helpers/misc.py:63-114).

    The interval [lo, hi] that lands on [dst_min, dst_max] is the data's
    minimum and maximum, or its perc_low / perc_up percentiles. Each value x
    goes to unit * (dst_max - dst_min) + dst_min with unit = (1 / (hi - lo)) *
    (x - lo), the width hi - lo taken in the data's dtype and the rest in
    float64 (so the unit position rounds as the reference's does; scaling
    onto [0, 1] is exact). An empty interval gives an array of dst_min. With
    percentiles and clip, the result is clipped to [dst_min, dst_max].
    """
    x = np.asarray(data)
    lo, hi, from_percentiles = _source_range(x, perc_low, perc_up)
    width = hi - lo
    if width == 0:
        return np.full(x.shape, dst_min, dtype=np.float64)
    unit = (1.0 / width) * (x.astype(np.float64) - lo)
    out = unit * (dst_max - dst_min) + dst_min
    if clip and from_percentiles:
        np.clip(out, dst_min, dst_max, out=out)
    return out
