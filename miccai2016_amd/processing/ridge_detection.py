"""Ridge detection in the vesselness response (reference: processing/ridge_detection.py)."""
import numpy as np

from .. import _lib
from ..device import default_context


def float_threshold(t):
    """Largest float32 <= t: then float32 r > result  <=>  r > t (double)."""
    f = np.float32(t)
    if float(f) > float(t):
        f = np.nextafter(f, np.float32(-np.inf))
    return float(f)


class RawRidgeDetector:
    """Marks voxels whose response exceeds `threshold` and every trilinear
    sample of the 8-neighbourhood in its cross-section plane."""

    def __init__(self, loader, threshold):
        self.__threshold = threshold
        self._loader = loader if loader is not None else default_context()

    @property
    def threshold(self):
        return self.__threshold

    def execute(self, result_buffer, response_buffer, qeigs_buffer, shape):
        """result_buffer (float32 0/1), response_buffer (float32), qeigs_buffer
        (squeezed frames AoS x3), all device tensors of the volume `shape`."""
        assert len(shape) == 3
        n = int(np.prod(shape))
        assert result_buffer.numel() >= n and response_buffer.numel() >= n
        assert qeigs_buffer.numel() >= 3 * n
        dev = self._loader
        rc = dev.lib.mh_detect_ridges(dev.handle, dev.ptr(result_buffer), dev.ptr(response_buffer),
                                      dev.ptr(qeigs_buffer), _lib.dims3(shape),
                                      float_threshold(self.threshold), dev.stream)
        dev.check(rc, "mh_detect_ridges")


def hysteresis_threshold(cand, sure):
    """Keep candidate voxels 26-connected (through candidates) to a sure
    voxel, like the Canny edge detector. The result is written into `sure`
    (as the reference does) and returned."""
    cand = np.asarray(cand)
    assert cand.shape == np.shape(sure)
    c8 = np.require(cand != 0, np.uint8, ["C"])
    s8 = np.require(np.asarray(sure) != 0, np.uint8, ["C"])
    out = np.empty_like(c8)
    L = _lib.lib()
    rc = L.mh_hysteresis_host(c8.ctypes.data, s8.ctypes.data, out.ctypes.data,
                              _lib.dims3(cand.shape))
    if rc != 0:
        raise RuntimeError(f"mh_hysteresis_host failed (status {rc})")
    if isinstance(sure, np.ndarray):
        sure[...] = out.astype(sure.dtype)
        return sure
    return out.astype(bool)
