"""Frangi vesselness on the GPU with eigenvector frames as squeezed quaternions.

Drop-in for the reference's processing/qfrangi_gpu.py (Frangi et al., MICCAI
1998): same classes, constructors and ``execute`` signatures. Buffer-level
classes take torch device tensors where the reference took cl.Buffer;
``FeatureFilter.execute_for`` keeps numpy in / numpy out. The quaternions
represent the matrix of eigenvectors in columns, squeezed to z-x-z angles.
"""
import ctypes as C
from collections import namedtuple

import numpy as np

from .. import _lib
from ..device import default_context
from ..helpers import parameters
from . import slicing

MAX_RADIUS = 63  # longest kernel arm the HIP passes accept


class GaussianFilter3D:
    """Separable Gaussian (derivative) filter with mirror boundary, the
    equivalent of scipy.ndimage.gaussian_filter(mode='mirror')."""

    def __init__(self, loader, n_sigmas=4, round_up=True):
        self.__n_sigmas = n_sigmas
        self.__round_up = round_up
        self._loader = loader

    @property
    def n_sigmas(self):
        """Standard deviations covered by each kernel arm."""
        return self.__n_sigmas

    @property
    def round_up(self):
        """Arm length ceil(n_sigmas*sigma) if True, else int(n_sigmas*sigma + 0.5) (scipy)."""
        return self.__round_up

    def _gauss_kernel_with(self, sigma, order, scaling_factor=1):
        """Odd-sized taps of the Gaussian derivative of `order` (0..3),
        normalised over the truncated support, times scaling_factor, float32
        (qfrangi_gpu.py:158-185)."""
        s2 = float(sigma) ** 2
        x_max = (int(np.ceil(self.n_sigmas * sigma)) if self.round_up else
                 int(self.n_sigmas * sigma + 0.5))
        x = np.arange(-x_max, x_max + 1).astype(np.float64)
        g = np.exp(-.5 * x ** 2 / s2)
        g = g / np.sum(g)
        if order == 0:
            k = g
        elif order == 1:
            k = g * (-x) / s2
        elif order == 2:
            k = g * (x ** 2 - s2) / (s2 ** 2)
        elif order == 3:
            k = g * x * (3 * s2 - x ** 2) / (s2 ** 3)
        else:
            raise ValueError("order must be in {0, 1, 2, 3}")
        return np.require(scaling_factor * k, np.float32, ["C", "A"])

    def taps_for(self, sigma, order, scale_space_scaling=True, scale_space_gamma=1):
        sigma = np.multiply(sigma, np.ones(3))
        order = np.multiply(order, np.ones(3, dtype=int)).astype(int)
        f = (lambda s, o: s ** (o * scale_space_gamma)) if scale_space_scaling else (lambda s, o: 1)
        return [self._gauss_kernel_with(s, o, f(s, o)) for s, o in zip(sigma, order)]

    def execute(self, in_buffer, swap_buffer, shape, sigma, order, scale_space_scaling=True,
                scale_space_gamma=1):
        """Filter in_buffer (axis 0, 1, 2 in turn). Returns (result, other):
        the result is in swap_buffer; in_buffer's content is altered."""
        assert len(shape) == 3
        n = int(np.prod(shape))
        assert in_buffer.numel() >= n and swap_buffer.numel() >= n
        taps = self.taps_for(sigma, order, scale_space_scaling, scale_space_gamma)
        radii = [len(t) // 2 for t in taps]
        if max(radii) > MAX_RADIUS:
            raise ValueError(f"kernel arm {max(radii)} exceeds {MAX_RADIUS}")
        dev = self._loader
        ptrs = [t.ctypes.data_as(C.c_void_p) for t in taps]
        rc = dev.lib.mh_gauss_filter(dev.handle, dev.ptr(in_buffer), dev.ptr(swap_buffer),
                                     dev.ptr(in_buffer), _lib.dims3(shape), ptrs[0], radii[0],
                                     ptrs[1], radii[1], ptrs[2], radii[2], dev.stream)
        dev.check(rc, "mh_gauss_filter")  # taps are copied into the launch arguments
        return swap_buffer, in_buffer


class Hessian:
    """Scale-normalised Gaussian-derivative Hessians; result AoS x6
    [H00, H11, H22, H01, H02, H12] (qfrangi_gpu.py:195-271). The six entries
    share their axis-0 and axis-1 passes."""

    def __init__(self, gaussian_filter):
        self._gaussian_filter = gaussian_filter
        self._work = None

    def packed_taps(self, scales):
        gf = self._gaussian_filter
        sig = np.multiply(scales, np.ones(3))
        taps = [[gf._gauss_kernel_with(sig[a], o, sig[a] ** o) for o in range(3)] for a in range(3)]
        radii = [len(taps[a][0]) // 2 for a in range(3)]
        rmax = max(radii)
        if rmax > MAX_RADIUS:
            raise ValueError(f"kernel arm {rmax} exceeds {MAX_RADIUS}")
        K = 2 * rmax + 1
        packed = np.zeros(9 * K, np.float32)
        for a in range(3):
            for o in range(3):
                t = taps[a][o]
                packed[(3 * a + o) * K:(3 * a + o) * K + len(t)] = t
        return packed, radii

    def execute(self, in_buffer, swap_buffer1, swap_buffer2, shape, result_buffer, scales):
        """Hessians of in_buffer (unchanged) into result_buffer (>= 6N floats).
        The swap buffers are accepted for compatibility; the shared passes use
        an internal work area."""
        dev = self._gaussian_filter._loader
        n = int(np.prod(shape))
        assert len(shape) == 3 and in_buffer.numel() >= n and result_buffer.numel() >= 6 * n
        wf = int(dev.lib.mh_hessian_work_floats(_lib.dims3(shape)))
        if self._work is None or self._work.numel() < wf:
            self._work = dev.empty(wf)
        packed, radii = self.packed_taps(scales)
        rc = dev.lib.mh_gauss_hessian(dev.handle, dev.ptr(in_buffer), _lib.dims3(shape),
                                      packed.ctypes.data_as(C.c_void_p), _lib.I32x3(*radii),
                                      dev.ptr(result_buffer), dev.ptr(self._work), dev.stream)
        dev.check(rc, "mh_gauss_hessian")
        return result_buffer


class Eigendecomposition:
    """Per-voxel Jacobi eigen-decomposition, eigenvalues sorted by ascending
    |lambda| (qfrangi_gpu.py:274-332)."""

    def __init__(self, loader, maxiter=50):
        self._loader = loader
        self._maxiter = maxiter

    def execute(self, num_voxels, in_buffer, result_buffer=None, norm_max=None):
        """AoS x6 Hessians -> AoS x6 [l1, l2, l3, phi, theta, psi], in place if
        result_buffer is None. With norm_max (1-float device tensor, zeroed by
        the caller) the maximum eigenvalue norm is accumulated too."""
        dev = self._loader
        out = result_buffer if result_buffer is not None else in_buffer
        rc = dev.lib.mh_eigs(dev.handle, dev.ptr(in_buffer), dev.ptr(out), int(num_voxels),
                             int(self._maxiter), dev.ptr(norm_max), dev.stream)
        dev.check(rc, "mh_eigs")
        return out


class VesselnessUpdate:
    """Running maximum of the per-scale Frangi response (qfrangi_gpu.py:335-439)."""

    def __init__(self, loader):
        self._loader = loader
        self._norm = None

    def execute(self, num_voxels, val_buffer, quat_buffer, eigs_buffer, bright_objects, alpha,
                beta, c=None, c_buffer1=None, c_buffer2=None, norm_max=None):
        """Update val/quat with this scale's response. c None: c = half the
        maximum eigenvalue norm (taken from norm_max if given, else reduced
        from eigs_buffer). The c buffers are accepted for compatibility."""
        dev = self._loader
        if c is None and norm_max is None:
            if self._norm is None:
                self._norm = dev.empty(1)
            dev.check(dev.lib.mh_eig_norm_max(dev.handle, dev.ptr(eigs_buffer), int(num_voxels),
                                              dev.ptr(self._norm), dev.stream), "mh_eig_norm_max")
            norm_max = self._norm
        cval = -1.0 if c is None else float(c)
        rc = dev.lib.mh_update_vesselness(dev.handle, dev.ptr(val_buffer), dev.ptr(quat_buffer),
                                          dev.ptr(eigs_buffer), int(num_voxels), float(alpha),
                                          float(beta), cval, dev.ptr(norm_max),
                                          int(bool(bright_objects)), dev.stream)
        dev.check(rc, "mh_update_vesselness")


class VesselnessCalculation:
    """Multi-scale Frangi response (qfrangi_gpu.py:442-631)."""

    class Parameters(parameters.Parameters):

        def __init__(self):
            self.alpha = 0.5   # weight of the lambda2 / lambda3 ratio
            self.beta = 0.5    # weight of the lambda1 / sqrt(lambda2 lambda3) ratio
            self.c = None      # structureness weight; None: half the max eigenvalue norm per scale
            self.bright_objects = None  # bright tubes on dark background?
            self.scales = None  # (start mm, stop mm, number), linspace

    def __init__(self, loader):
        self._loader = loader
        self._scales_vx = None

    def _init_fields(self, params, shape, voxel_size):
        self._params = params
        self._nvoxels = int(np.prod(shape))
        scales_mm = np.linspace(*params.scales).reshape(-1, 1)
        scaling = np.divide(1., voxel_size).reshape(1, -1)
        self._scales_vx = np.dot(scales_mm, scaling)

    def execute(self, params, shape, voxel_size, img_buffer, swap_buffer1=None,
                swap_buffer2=None, hessian_buffer=None, val_buffer=None, quat_buffer=None,
                gaussian_filter_3d=None, eigendecomposition=None, vesselness_update=None):
        """Returns (val_buffer, quat_buffer) device tensors; buffers are created
        when None. img_buffer is unchanged."""
        dev = self._loader
        self._init_fields(params, shape, voxel_size)
        n = self._nvoxels
        hess = hessian_buffer if hessian_buffer is not None else dev.empty(6 * n)
        val = val_buffer if val_buffer is not None else dev.empty(n)
        quat = quat_buffer if quat_buffer is not None else dev.empty(3 * n)
        assert hess.numel() >= 6 * n and val.numel() >= n and quat.numel() >= 3 * n
        dev.fill(val[:n], -np.inf)
        dev.fill(quat[:3 * n], 0.0)
        gf = gaussian_filter_3d if gaussian_filter_3d is not None else GaussianFilter3D(dev)
        hessian = Hessian(gf)
        eig = eigendecomposition if eigendecomposition is not None else Eigendecomposition(dev)
        upd = vesselness_update if vesselness_update is not None else VesselnessUpdate(dev)
        norm = dev.empty(1)
        for scale in self._scales_vx:
            hessian.execute(img_buffer, swap_buffer1, swap_buffer2, shape, hess, scale)
            if params.c is None:
                dev.fill(norm, 0.0)
                eig.execute(n, hess, norm_max=norm)
            else:
                eig.execute(n, hess)
            upd.execute(n, val, quat, hess, params.bright_objects, params.alpha, params.beta,
                        params.c, norm_max=norm if params.c is None else None)
        return val, quat


class FeatureFilter(parameters.TreePrintable):
    """Vesselness of a 3-D image in a region of interest (qfrangi_gpu.py:634-783)."""

    DummyBuffer = namedtuple("DummyBuffer", ["size"])

    def __init__(self, vesselness_params, loader=None, verbose=True):
        self._loader = loader if loader is not None else default_context()
        self.vesselness_params = vesselness_params
        self.verbose = verbose
        # "round_up=False" for compatibility with scipy's truncation
        self._gaussian_filter_3d = GaussianFilter3D(self._loader, round_up=False)
        self._eigendecomposition = Eigendecomposition(self._loader)
        self._vesselness_calculation = VesselnessCalculation(self._loader)
        self._bufs = None
        self._n = None

    def _init_buffers(self, n):
        if self._n != n:
            dev = self._loader
            self._bufs = dict(img=dev.empty(n), hess=dev.empty(6 * n), val=dev.empty(n),
                              quat=dev.empty(3 * n))
            self._n = n

    def execute_for(self, img, roi, voxel_size=None):
        """Returns (vesselness (shape), dirs (shape + (3,)), slc) for the ROI
        (None: whole image; bool/int mask or tuple of slices otherwise, widened
        by the largest kernel arm)."""
        val, quat, slc, shape = self.execute_device(img, roi, voxel_size)
        dev = self._loader
        return (dev.download(val).reshape(shape), dev.download(quat).reshape(shape + (3,)), slc)

    def execute_device(self, img, roi=None, voxel_size=None):
        """As execute_for but returns device tensors (val, quat, slc, shape)."""
        voxel_size = np.ones(3) if voxel_size is None else voxel_size
        cropped, slc = self._determine_slicing_and_crop_voxels(np.asarray(img), roi, voxel_size)
        cropped = np.require(cropped, np.float32, ["C", "A"])
        shape = cropped.shape
        n = int(np.prod(shape))
        self._init_buffers(n)
        b = self._bufs
        b["img"].copy_(self._loader.upload(cropped).reshape(-1))
        val, quat = self._vesselness_calculation.execute(
            self.vesselness_params, shape, voxel_size, b["img"], None, None, b["hess"], b["val"],
            b["quat"], self._gaussian_filter_3d, self._eigendecomposition, None)
        return val[:n], quat[:3 * n], slc, tuple(shape)

    def execute_tensor(self, img_t, shape, voxel_size=None):
        """Whole-volume response of a device tensor (float32, N): returns
        device tensors (val (N), quat (3N)) owned by this filter."""
        voxel_size = np.ones(3) if voxel_size is None else voxel_size
        n = int(np.prod(shape))
        self._init_buffers(n)
        b = self._bufs
        b["img"].copy_(img_t.reshape(-1))
        val, quat = self._vesselness_calculation.execute(
            self.vesselness_params, tuple(shape), voxel_size, b["img"], None, None, b["hess"],
            b["val"], b["quat"], self._gaussian_filter_3d, self._eigendecomposition, None)
        return val[:n], quat[:3 * n]

    def _determine_slicing_and_crop_voxels(self, voxels, roi, voxel_size):
        if roi is None:
            return voxels, None
        max_scales = self.vesselness_params.scales[1] / np.asarray(voxel_size, dtype=float)
        n_sigmas = self._gaussian_filter_3d.n_sigmas
        if self._gaussian_filter_3d.round_up:
            margin = np.ceil(n_sigmas * max_scales)
        else:
            margin = (n_sigmas * max_scales + 0.5).astype(int)
        if isinstance(roi, np.ndarray):
            roi = slicing.bounding_box(roi, margin=0)
        slc = slicing.add_margin_to_slc(margin, roi)
        return voxels[slc], slc
