"""ROI slicing helpers (reference: processing/slicing.py:13-61)."""
import numpy as np


def bounding_box(mask, margin=0):
    """Slices of the bounding box of the non-zeros of mask, widened by margin
    (scalar or per axis, rounded) and clipped at 0; empty mask -> empty slices."""
    mask = np.asarray(mask)
    margin = (np.ones(mask.ndim) * np.round(margin)).astype(int)
    nz = np.nonzero(mask)
    if nz[0].size == 0:
        return tuple(slice(0, 0, 1) for _ in range(mask.ndim))
    lowers = np.maximum(0, np.min(nz, axis=1) - margin)
    uppers = np.maximum(0, np.max(nz, axis=1) + margin + 1)
    return tuple(slice(int(lo), int(up), 1) for lo, up in zip(lowers, uppers))


def add_margin_to_slc(margin, slc):
    """Widen every slice by margin (scalar or per axis, rounded), clipped at 0."""
    d = len(slc)
    margin = (np.ones(d) * np.round(margin)).astype(int)
    out = []
    for s, m in zip(slc, margin):
        out.append(slice(max(0, int(s.start) - int(m)), max(0, int(s.stop) + int(m)), s.step))
    return tuple(out)
