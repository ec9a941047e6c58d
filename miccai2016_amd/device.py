"""Device context: the object that fills the reference's ``loader`` /
``cl_program_loader`` slot (helpers/cl_helpers.py:84-183 in the reference).

It binds one GPU, owns the C-ABI context (mh_ctx) and hands out torch device
tensors as buffers. Work is enqueued on torch's current stream of that
device; methods that return numpy arrays synchronise first, as the
reference's blocking reads do.
"""
import ctypes as C

import numpy as np

from . import _lib


class DeviceError(RuntimeError):
    pass


class DeviceContext:
    """One GPU and its libmiccai_hip context.

    Raises RuntimeError when no GPU is present (as ClProgramLoader does,
    cl_helpers.py:113-114) or when the HIP library is not built.
    """

    def __init__(self, device=0):
        import torch
        self.torch = torch
        if not torch.cuda.is_available():
            raise RuntimeError("No ROCm GPU device found!")
        self.index = int(device)
        self.device = torch.device("cuda", self.index)
        self.lib = _lib.lib()
        h = C.c_void_p()
        with torch.cuda.device(self.index):
            rc = self.lib.mh_create(self.index, C.byref(h))
        if rc != 0 or not h.value:
            raise RuntimeError(f"mh_create failed on device {self.index} (status {rc})")
        self._h = h

    # ---------------------------------------------------------------- plumbing
    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        """Current torch stream of this device as a raw hipStream_t."""
        return C.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    def check(self, rc, what=""):
        if rc < 0:
            msg = self.lib.mh_last_error(self._h)
            raise DeviceError(f"{what}: {msg.decode() if msg else 'error'} (status {rc})")
        return rc

    def synchronize(self):
        self.torch.cuda.current_stream(self.device).synchronize()

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and self._h.value:
                self.lib.mh_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass

    # ----------------------------------------------------------------- buffers
    def empty(self, shape, dtype=None):
        dtype = self.torch.float32 if dtype is None else dtype
        return self.torch.empty(shape, dtype=dtype, device=self.device)

    def zeros(self, shape, dtype=None):
        dtype = self.torch.float32 if dtype is None else dtype
        return self.torch.zeros(shape, dtype=dtype, device=self.device)

    def bytes(self, nbytes):
        """Raw workspace of nbytes (256-byte aligned by the caching allocator)."""
        return self.torch.empty(int(nbytes), dtype=self.torch.uint8, device=self.device)

    def upload(self, array, dtype=np.float32):
        a = np.require(array, dtype=dtype, requirements=["C", "A"])
        return self.torch.from_numpy(a).to(self.device, non_blocking=False)

    def download(self, tensor):
        self.synchronize()
        return tensor.detach().cpu().numpy()

    @staticmethod
    def ptr(t):
        return C.c_void_p(0 if t is None else t.data_ptr())

    # -------------------------------------------------------------- primitives
    def fill(self, t, value):
        self.check(self.lib.mh_fill(self._h, self.ptr(t), t.numel(), float(value), self.stream),
                   "mh_fill")

    def sum_abs(self, t, out=None):
        """Sum of |t| into a 1-element device tensor (Reducer with sum_abs.cl)."""
        out = self.empty(1) if out is None else out
        self.check(self.lib.mh_reduce_sum_abs(self._h, self.ptr(t), t.numel(), self.ptr(out),
                                              self.stream), "mh_reduce_sum_abs")
        return out

    def max(self, t, out=None):
        """NaN-neutral maximum of t into a 1-element device tensor (Reducer with max.cl)."""
        out = self.empty(1) if out is None else out
        self.check(self.lib.mh_reduce_max(self._h, self.ptr(t), t.numel(), self.ptr(out),
                                          self.stream), "mh_reduce_max")
        return out


_default = {}


def default_context(device=0):
    """Process-wide context per device (created on first use)."""
    if device not in _default:
        _default[device] = DeviceContext(device)
    return _default[device]
