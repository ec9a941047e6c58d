"""Continuous max-flow with anisotropic total variation (ATV) on the GPU.

Drop-in for the reference's processing/continuous_cut_anisotropic_gpu.py:
same class, parameter bag, constructor and ``execute`` signature; numpy in,
numpy out, blocking. The solver runs in libmiccai_hip (``mh_cmf_solve``).
Extensions beside it: ``execute_device`` (device-resident torch tensors),
``execute_batched`` (many solves sharing one volume's fields, the grid search
of phantom_experiment.py:350-368) and ``execute_ragged`` (packed volumes of
different shapes).
"""
import ctypes as C
import math

import numpy as np

from .. import _lib
from ..device import default_context
from ..helpers.parameters import Parameters, TreePrintable


def derived_constants(params, scaling):
    """Step constants of continuous_cut_anisotropic_gpu.py:104-111:
    rs = 1/s, ss = steps * min(s)^2 / s, rcc = 1 / (cc * min(s)^2)."""
    scaling = np.asarray(scaling, dtype=np.float64)
    rs = np.divide(1., scaling)
    min_s = np.min(scaling)
    ss = np.divide(params.steps, scaling) * min_s ** 2
    rcc = 1. / (params.cc * min_s ** 2)
    return rcc, rs, ss


def initial_u(img):
    """u0 = img / (max(img) + tiny), float32 (continuous_cut_anisotropic_gpu.py:282-283)."""
    img = np.require(img, dtype=np.float32, requirements=["C", "A"])
    return np.require(np.divide(img, np.max(img) + np.finfo(img.dtype).tiny), np.float32,
                      ["C", "A"])


class CmfCutQAnisotropicGpu(TreePrintable):

    class CutParameters(Parameters):

        def __init__(self):
            self.cc = 0.2             # step size of the augmented Lagrangian
            self.error_bound = 1e-4   # stop when the criterion falls below it
            self.num_iter = 1000
            self.update = 1           # criterion every `update` iterations
            self.steps = 0.11         # gradient-ascent step of the p update
            self.lsize_3d = (64, 2, 1)  # accepted for compatibility; launch shapes are internal
            self.lsize_1d = None

    def __init__(self, params, cl_program_loader=None):
        self.float_type = np.float32
        self.params = params
        self.loader = cl_program_loader if cl_program_loader is not None else default_context()
        self.last_result = None
        self.last_criteria = None

    # ------------------------------------------------------------ helpers ---
    def _cmf_params(self, scaling, const_sqevals, n_voxels):
        p = self.params
        if int(p.update) < 1:
            raise ValueError("CutParameters.update must be >= 1")
        rcc, rs, ss = derived_constants(p, scaling)
        P = _lib.CmfParams()
        P.rcc = float(rcc)
        for k in range(3):
            P.rs[k] = float(rs[k])
            P.ss[k] = float(ss[k])
        if const_sqevals is not None:
            P.const_evals = 1
            for k in range(3):
                P.sqev[k] = float(const_sqevals[k])
        else:
            P.const_evals = 0
        P.num_iter = int(p.num_iter)
        P.update = int(p.update)
        P.error_bound = float(p.error_bound)
        return P

    def _report(self, crit, verbose):
        if not verbose:
            return
        upd = int(self.params.update)
        for k, c in enumerate(crit):
            print(k * upd + 1, float(c))
        res = self.last_result
        if res is not None and res.stopped == 1:
            print("Convergence criterion is NaN. Stop iterating.")
        elif res is not None and res.stopped == 2:
            print("Converged.")

    # ------------------------------------------------------------ public ---
    def execute(self, img, quats, eigs, alpha, cs, ct, labels=None, scaling=None, verbose=True):
        """Segment img; returns u (float32, img.shape), not thresholded.

        quats: squeezed frames (D0, D1, D2, 3); eigs: 3 constant eigenvalues
        or (D0, D1, D2, 3) pointwise ones; alpha: scalar or array (TV weight);
        cs, ct: source/sink capacities; labels: uint8 seeds (1 foreground,
        2 background) for the supervised cut; scaling: voxel size (mm/voxel).
        """
        dev = self.loader
        img = np.require(img, dtype=np.float32, requirements=["C", "A"])
        scaling = (1, 1, 1) if scaling is None else scaling
        sqevals = np.sqrt(np.asanyarray(eigs, dtype=np.float64))
        if not isinstance(alpha, np.ndarray):
            alpha = np.multiply(alpha, np.ones(img.shape), dtype=np.float32)
        if sqevals.ndim == 1:
            eig_buf = np.require(quats, dtype=np.float32, requirements=["C", "A"])
            const = sqevals
        else:
            eig_buf = np.empty(img.shape + (6,), dtype=np.float32)
            eig_buf[..., :3] = sqevals
            eig_buf[..., 3:] = quats
            const = None
        u0 = initial_u(img)
        t = dict(u=dev.upload(u0), eigs=dev.upload(eig_buf), alpha=dev.upload(alpha),
                 cs=dev.upload(cs), ct=dev.upload(ct))
        lab = None if labels is None else dev.upload(labels, dtype=np.uint8)
        u = self.execute_device(t["u"], t["eigs"], t["alpha"], t["cs"], t["ct"], labels=lab,
                                scaling=scaling, const_sqevals=const, shape=img.shape,
                                verbose=verbose)
        return dev.download(u).reshape(img.shape)

    def execute_device(self, u, eigs, alpha, cs, ct, labels=None, scaling=None,
                       const_sqevals=None, shape=None, verbose=False):
        """Solve on device tensors in place of u (u0 on entry). eigs is the
        squeezed frame field (AoS x3) when const_sqevals is given, else
        AoS x6 [sqrt(l) x3, squeezed x3]. Returns u."""
        dev = self.loader
        shape = tuple(u.shape) if shape is None else tuple(shape)
        n = int(np.prod(shape))
        scaling = (1, 1, 1) if scaling is None else scaling
        P = self._cmf_params(scaling, const_sqevals, n)
        wbytes = dev.lib.mh_cmf_workspace_bytes(n)
        work = dev.bytes(wbytes)
        ncheck = max(1, math.ceil(P.num_iter / P.update))
        crit = np.zeros(ncheck, np.float32)
        res = _lib.CmfResult()
        rc = dev.lib.mh_cmf_solve(dev.handle, dev.ptr(u), dev.ptr(eigs), dev.ptr(alpha),
                                  dev.ptr(cs), dev.ptr(ct), dev.ptr(labels), _lib.dims3(shape),
                                  C.byref(P), dev.ptr(work), wbytes, C.byref(res),
                                  crit.ctypes.data_as(C.c_void_p), dev.stream)
        dev.check(rc, "mh_cmf_solve")
        self.last_result = res
        self.last_criteria = crit[:res.ncrit].copy()
        self._report(self.last_criteria, verbose)
        return u

    def execute_batched(self, img, quats, eigs, nt_weightings, nt_decays, grad_mag_sq, cs, ct,
                        scaling=None, device_out=False):
        """S unsupervised solves sharing img (u0), quats, cs, ct and |grad I|^2;
        solve s uses alpha = nt_weightings[s] * exp(-grad_mag_sq / nt_decays[s]^2)
        (phantom_experiment.py:359). Returns u of shape (S,) + img.shape."""
        dev = self.loader
        img = np.require(img, dtype=np.float32, requirements=["C", "A"])
        w = np.ascontiguousarray(nt_weightings, dtype=np.float32).ravel()
        d = np.ascontiguousarray(nt_decays, dtype=np.float32).ravel()
        if w.shape != d.shape:
            raise ValueError("nt_weightings and nt_decays must have the same length")
        S = w.size
        u0 = dev.upload(initial_u(img))
        u = u0.reshape(1, -1).repeat(S, 1).contiguous()
        out = self.execute_batched_device(u, S, w, d, dev.upload(quats), dev.upload(grad_mag_sq),
                                          dev.upload(cs), dev.upload(ct), img.shape, eigs,
                                          scaling)
        if device_out:
            return out.reshape((S,) + img.shape)
        return dev.download(out).reshape((S,) + img.shape)

    def execute_batched_device(self, u, S, w, decay, quats, gms, cs, ct, shape, eigs, scaling=None):
        """Device form of execute_batched: u is (S*N) float32 holding u0 per
        solve; returns u. Per-solve results in self.last_result (list)."""
        dev = self.loader
        n = int(np.prod(shape))
        scaling = (1, 1, 1) if scaling is None else scaling
        sq = np.sqrt(np.asarray(eigs, dtype=np.float64))
        if sq.ndim != 1:
            raise ValueError("batched solves need 3 constant eigenvalues")
        P = self._cmf_params(scaling, sq, n)
        wbytes = dev.lib.mh_cmf_batched_workspace_bytes(n, int(S))
        work = dev.bytes(wbytes)
        res = (_lib.CmfResult * int(S))()
        w = np.ascontiguousarray(w, np.float32)
        decay = np.ascontiguousarray(decay, np.float32)
        rc = dev.lib.mh_cmf_solve_batched(dev.handle, dev.ptr(u), int(S),
                                          w.ctypes.data_as(C.c_void_p),
                                          decay.ctypes.data_as(C.c_void_p), dev.ptr(quats),
                                          dev.ptr(gms), dev.ptr(cs), dev.ptr(ct),
                                          _lib.dims3(shape), C.byref(P), dev.ptr(work), wbytes,
                                          res, dev.stream)
        dev.check(rc, "mh_cmf_solve_batched")
        self.last_result = list(res)
        return u

    def execute_ragged(self, imgs, quats, eigs, alphas, css, cts, scaling=None,
                       sweep_min_voxels=1 << 19):
        """One unsupervised solve per volume, volumes of different shapes packed
        into one launch sequence (volumes of sweep_min_voxels or more on their
        own, see execute_ragged_device). Lists in, list of u arrays out."""
        dev = self.loader
        shapes = [np.shape(a) for a in imgs]
        sizes = [int(np.prod(s)) for s in shapes]
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        pack = lambda arrs: np.concatenate([np.require(a, np.float32, ["C"]).ravel() for a in arrs])
        u0 = pack([initial_u(a) for a in imgs])
        alph = [a if isinstance(a, np.ndarray) else np.full(s, a, np.float32)
                for a, s in zip(alphas, shapes)]
        t_u = dev.upload(u0)
        self.execute_ragged_device(t_u, dev.upload(pack(quats)), dev.upload(pack(alph)),
                                   dev.upload(pack(css)), dev.upload(pack(cts)), shapes, offs,
                                   eigs, scaling, sweep_min_voxels)
        u = dev.download(t_u)
        return [u[o:o + sz].reshape(s) for o, sz, s in zip(offs, sizes, shapes)]

    def execute_ragged_device(self, u, quats, alpha, cs, ct, shapes, offsets, eigs, scaling=None,
                              sweep_min_voxels=1 << 19):
        """Volumes of at least sweep_min_voxels voxels are solved one at a
        time with the tuned single-solve method (the multi-level sweep where it
        wins), the others together in one packed launch sequence; both give
        the same iterations bit for bit (None: all packed)."""
        dev = self.loader
        shapes = [tuple(int(v) for v in s) for s in shapes]
        offsets = [int(o) for o in offsets]
        sizes = [int(np.prod(s)) for s in shapes]
        big = [v for v in range(len(shapes))
               if sweep_min_voxels is not None and sizes[v] >= sweep_min_voxels]
        if big:
            sq = np.sqrt(np.asarray(eigs, dtype=np.float64))
            results = [None] * len(shapes)
            for v in big:
                o, m = offsets[v], sizes[v]
                self.execute_device(u[o:o + m], quats[3 * o:3 * (o + m)], alpha[o:o + m],
                                    cs[o:o + m], ct[o:o + m], scaling=scaling, const_sqevals=sq,
                                    shape=shapes[v])
                results[v] = self.last_result
            small = [v for v in range(len(shapes)) if v not in set(big)]
            if small:
                self.execute_ragged_device(u, quats, alpha, cs, ct, [shapes[v] for v in small],
                                           [offsets[v] for v in small], eigs, scaling, None)
                for v, r in zip(small, self.last_result):
                    results[v] = r
            self.last_result = results
            return u
        scaling = (1, 1, 1) if scaling is None else scaling
        sq = np.sqrt(np.asarray(eigs, dtype=np.float64))
        total = int(max(int(o) + int(np.prod(s)) for o, s in zip(offsets, shapes)))
        P = self._cmf_params(scaling, sq, total)
        nv = len(shapes)
        vd = np.ascontiguousarray(np.asarray(shapes, np.int64).reshape(nv, 3))
        vo = np.ascontiguousarray(np.asarray(offsets, np.int64))
        wbytes = dev.lib.mh_cmf_ragged_workspace_bytes(total, nv)
        work = dev.bytes(wbytes)
        res = (_lib.CmfResult * nv)()
        rc = dev.lib.mh_cmf_solve_ragged(dev.handle, dev.ptr(u), dev.ptr(quats), dev.ptr(alpha),
                                         dev.ptr(cs), dev.ptr(ct), nv,
                                         vd.ctypes.data_as(C.c_void_p),
                                         vo.ctypes.data_as(C.c_void_p), C.byref(P),
                                         dev.ptr(work), wbytes, res, dev.stream)
        dev.check(rc, "mh_cmf_solve_ragged")
        self.last_result = list(res)
        return u
