"""Host-side helpers (reference: helpers/misc.py)."""
import numpy as np


def normalize(data, dst_min=0.0, dst_max=1.0, perc_low=None, perc_up=None, clip=True):
    """Map data linearly to [dst_min, dst_max] in float64 (helpers/misc.py:63-114).

    The source range is [min, max] of data, or the given percentiles. If the
    range is empty, an array of dst_min is returned. With percentiles and
    clip, values outside the range are clipped.
    """
    use_percentiles = perc_low is not None or perc_up is not None
    if use_percentiles:
        lo = 0. if perc_low is None else perc_low
        hi = 100. if perc_up is None else perc_up
        src_min, src_max = np.percentile(data, [lo, hi])
    else:
        src_min, src_max = np.min(data), np.max(data)
    if src_max - src_min == 0:
        return np.full(np.shape(data), dst_min, dtype=np.float64)
    result = (1. / (src_max - src_min)) * (np.asarray(data, dtype=np.float64) - src_min)
    if dst_max != 1 and dst_min == 0:
        result = result * dst_max
    elif dst_min != 0:
        result = result * (dst_max - dst_min) + dst_min
    if clip and use_percentiles:
        np.clip(result, dst_min, dst_max, out=result)
    return result
