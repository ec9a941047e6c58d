"""Gradient vector flow (Xu and Prince) on the GPU.

Drop-in for the reference's processing/gradient_vector_flow.py: same
``GvfParameters``, ``time_step_for`` and ``GvfGpu`` constructor/``execute``.
The iteration runs in libmiccai_hip (``mh_gvf``): one fused ping-pong stencil
launch per step. The convergence check (mean |delta| every ``update`` steps,
callback, ``tol``) keeps the reference's order: step i, then the check if
i % update == 0.
"""
import ctypes as C

import numpy as np

from .. import _lib
from ..device import default_context
from ..helpers.parameters import Parameters


class GvfParameters(Parameters):
    """mu: regularisation (the higher, the smoother); tol: stop when the mean
    update magnitude falls below it; maxiter; callback(v, i, mag) called at
    every check with the current field; update: check period; normalize:
    unit-length result vectors."""

    def __init__(self):
        self.mu = None
        self.tol = None
        self.maxiter = None
        self.callback = None
        self.update = None
        self.normalize = None


def time_step_for(voxel_size, mu, ndim):
    """dt = min(0.99, min(voxel_size)^ndim / (2.01^ndim mu)) (gradient_vector_flow.py:58-65)."""
    return np.minimum(0.99, (np.min(voxel_size) ** ndim) / (2.01 ** ndim * mu))


class GvfGpu:
    """3-D only. vfield: (D0, D1, D2, 3) vector field; response: optional
    fidelity weight (default |v|^2)."""

    def __init__(self, vfield, params, voxel_size=None, response=None, cl_program_loader=None,
                 shape=None):
        """vfield: numpy (D0, D1, D2, 3), or a device tensor holding the field
        SoA x3 (3N floats) together with `shape`."""
        self._device_input = shape is not None
        if not self._device_input:
            vfield = np.asarray(vfield)
            assert vfield.ndim == 4 and vfield.shape[-1] == 3
            shape = vfield.shape[:3]
        self.params = params
        self.shape = tuple(shape)
        self.n_voxels = int(np.prod(self.shape))
        if min(self.shape) < 2:
            raise ValueError("GvfGpu needs at least 2 voxels along every axis")
        self.voxel_size = (np.ones(3) if voxel_size is None else np.multiply(voxel_size, np.ones(3)))
        self.dt = time_step_for(self.voxel_size, params.mu, ndim=3)
        self.loader = cl_program_loader if cl_program_loader is not None else default_context()
        dev = self.loader
        # SoA x3 on the device (gradient_vector_flow.py:77-79)
        if self._device_input:
            assert vfield.numel() == 3 * self.n_voxels
            self.v = vfield.reshape(-1)
        else:
            soa = np.stack([np.require(vfield[..., k], np.float32, ["C"]) for k in range(3)])
            self.v = dev.upload(soa)
        self.g = dev.zeros(3 * self.n_voxels)
        self.swap = dev.empty(3 * self.n_voxels)
        self.r = dev.empty(self.n_voxels)
        self.work = dev.empty(int(dev.lib.mh_gvf_work_floats(_lib.dims3(self.shape))))
        self.mag = dev.empty(1)
        if response is not None:
            assert np.shape(response) == tuple(self.shape)
            self.r.copy_(dev.upload(response))
        else:
            dev.check(dev.lib.mh_gvf_response(dev.handle, dev.ptr(self.r), dev.ptr(self.v),
                                              self.n_voxels, dev.stream), "mh_gvf_response")

    def _gvf_params(self):
        P = _lib.GvfParams()
        P.mu = float(self.params.mu)
        P.dt = float(self.dt)
        rss = 1. / (self.voxel_size ** 2)
        for k in range(3):
            P.rss[k] = float(rss[k])
        return P

    def _run(self, iters, want_mag):
        dev = self.loader
        P = self._gvf_params()
        rc = dev.lib.mh_gvf(dev.handle, dev.ptr(self.g), dev.ptr(self.swap), dev.ptr(self.v),
                            dev.ptr(self.r), _lib.dims3(self.shape), C.byref(P), int(iters),
                            dev.ptr(self.mag) if want_mag else C.c_void_p(0),
                            dev.ptr(self.work) if want_mag else C.c_void_p(0), dev.stream)
        dev.check(rc, "mh_gvf")

    def execute(self):
        """Run GVF from g = 0; returns the field as a (D0, D1, D2, 3) float32 view."""
        self.execute_device()
        return self._gvf_from_buffers()

    def execute_device(self):
        """As execute, leaving the result on the device (self.g, SoA x3)."""
        p = self.params
        dev = self.loader
        self.g.zero_()
        maxiter = int(p.maxiter)
        update = int(p.update) if p.update else 0
        tol = p.tol
        checks = update > 0 and (p.callback is not None or (tol is not None and tol > 0))
        if not checks:
            self._run(maxiter, False)
        else:
            i = 0
            while i < maxiter:
                # advance to the next check iteration (i % update == 0), inclusive
                nxt = i if i % update == 0 else (i // update + 1) * update
                nxt = min(nxt, maxiter - 1)
                self._run(nxt - i + 1, nxt % update == 0)
                i = nxt + 1
                if nxt % update == 0:
                    mag = float(dev.download(self.mag)[0]) / self.n_voxels
                    if p.callback is not None:
                        p.callback(self._gvf_from_buffers(), nxt, mag)
                    if tol is not None and mag < tol:
                        break
        if p.normalize:
            dev.check(dev.lib.mh_gvf_normalize(dev.handle, dev.ptr(self.g), self.n_voxels,
                                               dev.stream), "mh_gvf_normalize")
        return self.g

    def _gvf_from_buffers(self):
        g = self.loader.download(self.g).reshape((3,) + tuple(self.shape))
        return np.rollaxis(g, 0, g.ndim)
