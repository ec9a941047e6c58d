"""Segmentation quality measures of the reference experiment
(phantom_experiment.py:36-103), plus their device form for batched solves.

``fast_dice_for`` thresholds at 0.5 and compares whole volumes;
``select_largest`` keeps the largest face-connected component (scipy's
default structure); ``dice_for`` is the Dice of that component.
``DeviceDice`` counts |S|, |S & T| and |T| for many labellings on the GPU
(exact integers) so a grid search's u fields never leave the device, and
``device_dice_for`` is dice_for with the components labelled on the GPU.
"""
import numpy as np
import scipy.ndimage

from . import _lib
from .device import default_context


def fast_dice_for(segmentation, ground_truth):
    """2 |S & T| / (|S| + |T|), S = segmentation >= 0.5; NaN when both are empty
    (phantom_experiment.py:36-47)."""
    s = np.asarray(segmentation) >= 0.5
    t = np.asarray(ground_truth)
    den = float(np.sum(s)) + float(np.sum(t))
    num = 2 * float(np.sum(np.multiply(s, t)))
    return np.nan if den == 0 else num / den


def select_largest(segmentation, return_size=False, threshold=0.5):
    """The largest connected component of segmentation >= threshold (scipy
    label, face connectivity); the first label wins ties. With no component:
    (None, -1) (phantom_experiment.py:50-80)."""
    labeled, n = scipy.ndimage.label(np.asarray(segmentation) >= threshold)
    best, size = None, -1
    if n:
        sizes = np.bincount(labeled.ravel(), minlength=n + 1)[1:]
        k = int(np.argmax(sizes))            # first of the largest
        best, size = labeled == (k + 1), int(sizes[k])
    return (best, size) if return_size else best


def dice_for(segmentation, ground_truth):
    """Dice of the largest component against ground_truth (phantom_experiment.py:83-103).
    Raises ValueError when the segmentation has no foreground (the reference
    fails there as well)."""
    largest, size = select_largest(segmentation, return_size=True)
    if largest is None:
        raise ValueError("segmentation has no foreground component")
    inter = float(np.sum(np.multiply(largest, ground_truth)))
    return 2 * inter / (size + float(np.sum(ground_truth)))


class DeviceDice:
    """fast_dice_for of S labellings (device tensor S x N, float32) against one
    mask (numpy or device uint8, N), reduced on the GPU."""

    def __init__(self, truth, loader=None):
        self.dev = loader if loader is not None else default_context()
        dev = self.dev
        if isinstance(truth, np.ndarray):
            self.truth = dev.upload(np.asarray(truth) != 0, dtype=np.uint8).reshape(-1)
        else:
            self.truth = truth.reshape(-1)
        self.n = self.truth.numel()

    def counts(self, u, nsolves):
        """(nsolves, 3) int64 array: |S|, |S & T|, |T| per labelling."""
        dev = self.dev
        assert u.numel() == nsolves * self.n
        wb = int(dev.lib.mh_dice_work_bytes(self.n, int(nsolves)))
        work = dev.bytes(wb)
        out = dev.empty(3 * nsolves, dtype=dev.torch.int64)
        dev.check(dev.lib.mh_dice_counts(dev.handle, dev.ptr(u), int(nsolves), self.n,
                                         dev.ptr(self.truth), dev.ptr(work), wb, dev.ptr(out),
                                         dev.stream), "mh_dice_counts")
        return dev.download(out).reshape(nsolves, 3)

    def __call__(self, u, nsolves=1):
        c = self.counts(u, nsolves).astype(np.float64)
        den = c[:, 0] + c[:, 2]
        with np.errstate(invalid="ignore", divide="ignore"):
            return np.where(den > 0, 2 * c[:, 1] / den, np.nan)


def device_dice_for(u, ground_truth, shape, loader=None, labels=None):
    """dice_for on the GPU: u (device tensor, N floats) thresholded at 0.5, its
    largest face-connected component (union-find labelling) against
    ground_truth (numpy or device uint8). Returns (dice, size of the largest
    component). Raises ValueError when u has no foreground, as dice_for does.
    labels: optional int32 device tensor (N) that receives the component of
    every voxel (its smallest voxel index; -1 for background)."""
    dev = loader if loader is not None else default_context()
    if isinstance(ground_truth, np.ndarray):
        t = dev.upload(np.asarray(ground_truth) != 0, dtype=np.uint8).reshape(-1)
    else:
        t = ground_truth.reshape(-1)
    wb = int(dev.lib.mh_largest_component_work_bytes(_lib.dims3(shape)))
    work = dev.bytes(wb)
    out = dev.empty(4, dtype=dev.torch.int64)
    dev.check(dev.lib.mh_largest_component_dice(dev.handle, dev.ptr(u), 0.5, dev.ptr(t),
                                                _lib.dims3(shape), dev.ptr(labels), dev.ptr(work),
                                                wb, dev.ptr(out), dev.stream),
              "mh_largest_component_dice")
    size, inter, tsum, _ = (int(v) for v in dev.download(out))
    if size < 0:
        raise ValueError("segmentation has no foreground component")
    return 2.0 * inter / (size + tsum), size
