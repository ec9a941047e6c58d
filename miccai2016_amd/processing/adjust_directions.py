"""Consistent signs for the vesselness main directions, and frame completion.

Drop-in for the reference's processing/adjust_directions.py: ridge voxels are
detected in the response (GPU), hysteresis-thresholded (host C++), their main
directions aligned along a minimum spanning tree (host, sparse Euclidean MST in
ridge_graph), the ridge
optionally cleaned, and every voxel's direction flipped to agree with its
nearest ridge voxel (GPU). ``TransformationCompleter`` turns directions back
into squeezed frames (GPU).
"""
import numpy as np
from scipy.ndimage import label

from .. import _lib
from ..device import default_context
from ..helpers.parameters import TreePrintable
from . import ridge_graph
from .ridge_detection import RawRidgeDetector, hysteresis_threshold


class AdvancedAdjuster(TreePrintable):
    """See the module docstring. threshold: scalar, or (lower, upper) for
    hysteresis; qeigs: squeezed frames (D0, D1, D2, 3); response: vesselness."""

    def __init__(self, threshold, qeigs, response, voxel_size=1, ridge_detection_only=False,
                 clean_ridge=True, loader=None):
        self.__threshold = None
        self.__qeigs = None
        self.__response = None
        self.__voxel_size = None
        self.__ridge = None
        self.__ridge_coords = None
        self.__ridge_device = None
        self.__main_dirs_unadjusted = None
        # execute_tensor: the device form of the ridge steps
        # (mh_adjust_ridge_device) where it applies, else the host construction
        self.device_adjust = True
        self.device_stats = None
        self.device_used = False
        self._ridge_detector = None
        self._hysteresis = None
        # (ridge points, their neighbour forest): the cleaning's forest, then
        # its restriction to the kept points, the final tree's first phase
        self._forest = None
        self._loader = loader if loader is not None else default_context()
        if threshold is not None:
            self.threshold = threshold
        if qeigs is not None:
            self.qeigs = qeigs
        if response is not None:
            self.response = response
        if voxel_size is not None:
            self.voxel_size = voxel_size
        self.ridge_detection_only = ridge_detection_only
        self.clean_ridge = clean_ridge

    @property
    def threshold(self):
        return self.__threshold

    @threshold.setter
    def threshold(self, value):
        try:
            len(value)
            value = tuple(sorted(value))
            self._hysteresis = True
            lower = value[0]
        except TypeError:
            self._hysteresis = False
            lower = value
        if value != self.__threshold:
            self.__threshold = value
            self._ridge_detector = RawRidgeDetector(self._loader, lower)

    @property
    def qeigs(self):
        return self.__qeigs

    @qeigs.setter
    def qeigs(self, value):
        self.__qeigs = np.require(value, dtype=np.float32, requirements=["C", "A"])

    @property
    def response(self):
        return self.__response

    @response.setter
    def response(self, value):
        self.__response = np.require(value, dtype=np.float32, requirements=["C", "A"])

    @property
    def voxel_size(self):
        return self.__voxel_size

    @voxel_size.setter
    def voxel_size(self, value):
        self.__voxel_size = tuple(np.multiply(np.ones(3), value))

    @property
    def ridge(self):
        """The detected (and, if requested, cleaned) ridge after execute()."""
        if self.__ridge is None and self.__ridge_device is not None:
            mask_t, shape, _ = self.__ridge_device
            self.__ridge = self._loader.download(mask_t).reshape(shape)
        if self.__ridge is None and self.__ridge_coords is not None:
            coords, shape = self.__ridge_coords
            mask = np.zeros(shape, np.float32)
            if len(coords):
                mask[tuple(np.asarray(coords).T)] = 1
            self.__ridge = mask
        return self.__ridge

    @property
    def main_dirs_unadjusted(self):
        """The extracted main directions before any sign adjustment."""
        return self.__main_dirs_unadjusted

    def execute(self, verbose=True):
        """Adjusted directions (D0, D1, D2, 3), or None if ridge_detection_only."""
        assert self.response.shape == self.qeigs.shape[:3] and self.qeigs.shape[3] == 3
        dev = self._loader
        self._verbose = verbose
        shape = self.response.shape
        self._d_qeigs = dev.upload(self.qeigs)
        self._d_resp = dev.upload(self.response)
        labeled = self._label_ridge()
        adjusted = None
        self._forest = None
        if (not self.ridge_detection_only) or self.clean_ridge:
            main_dirs = self._qeigs_to_main_dirs()
            if self.clean_ridge:
                coords, rdirs, parts = self._adjust_ridge_directions(labeled, main_dirs,
                                                                     for_cleaning=True)
                labeled = self._do_clean_ridge(labeled, coords, rdirs, parts)
            if not self.ridge_detection_only:
                coords, rdirs = self._adjust_ridge_directions(labeled, main_dirs)
                adjusted = self._adjust_all_directions(coords, rdirs, shape)
        self._forest = None
        return adjusted

    def execute_tensor(self, qeigs_t, resp_t, shape, verbose=False, candidates_t=None,
                       main_dirs_t=None):
        """Device form of execute(): qeigs_t (squeezed frames, 3N) and resp_t
        (N) device tensors; returns the adjusted directions (AoS x3 device
        tensor). The ridge steps run on the device (mh_adjust_ridge_device)
        where that form applies; otherwise the ridge's hysteresis and
        segments run on the device (mh_ridge_segments) and only the ridge
        voxels' indices, segments and directions visit the host (spanning
        tree, cleaning). candidates_t: ridge candidates (N floats, nonzero =
        candidate) detected elsewhere, in place of this adjuster's own ridge
        test; main_dirs_t: the frames' first columns (AoS x3) computed
        elsewhere, in place of mh_main_dirs (qeigs_t may then be None when
        candidates_t is given)."""
        dev = self._loader
        torch = dev.torch
        self._verbose = verbose
        shape = tuple(shape)
        n = int(np.prod(shape))
        src = main_dirs_t if main_dirs_t is not None else qeigs_t
        self._d_qeigs = src.reshape(-1).clone()
        self._d_resp = resp_t.reshape(-1)
        d_ridge = dev.empty(n)
        if candidates_t is None:
            self._ridge_detector.execute(d_ridge, self._d_resp, qeigs_t.reshape(-1), shape)
        else:
            if candidates_t.numel() != n:
                raise ValueError("candidates_t must hold one value per voxel")
            d_ridge.copy_(candidates_t.reshape(-1))
        labels = dev.empty(n, dtype=torch.int32)
        wbytes = int(dev.lib.mh_ridge_segments_work_bytes(_lib.dims3(shape)))
        work = dev.bytes(wbytes)
        upper = float(self.threshold[1]) if self._hysteresis else 0.0
        dev.check(dev.lib.mh_ridge_segments(dev.handle, dev.ptr(d_ridge), dev.ptr(self._d_resp),
                                            upper, int(self._hysteresis), _lib.dims3(shape),
                                            dev.ptr(d_ridge), dev.ptr(labels), dev.ptr(work),
                                            wbytes, dev.stream), "mh_ridge_segments")
        if main_dirs_t is None:
            dev.check(dev.lib.mh_main_dirs(dev.handle, dev.ptr(self._d_qeigs),
                                           dev.ptr(self._d_qeigs), n, dev.stream), "mh_main_dirs")
        if self.device_adjust:
            kept = dev.empty(n)
            stats = np.zeros(8, np.int64)
            vs = np.require(self.voxel_size, np.float64, ["C"])
            rc = dev.lib.mh_adjust_ridge_device(dev.handle, dev.ptr(self._d_qeigs), dev.ptr(d_ridge),
                                                dev.ptr(labels), _lib.dims3(shape), vs.ctypes.data,
                                                int(bool(self.clean_ridge)), dev.ptr(kept),
                                                stats.ctypes.data, dev.stream)
            self.device_stats = stats
            self.device_used = rc == _lib.MH_OK
            if rc == _lib.MH_OK:
                if getattr(self, "_verbose", False) and self.clean_ridge:
                    print("Keeping %s of %s labels." % (stats[2], stats[1]))
                self._set_ridge_device(kept, shape, int(stats[3]))
                return self._d_qeigs
            if rc != _lib.MH_DEFER:
                dev.check(rc, "mh_adjust_ridge_device")
        # ridge voxels in C order, their segments and main directions
        lin_t = torch.nonzero(d_ridge).reshape(-1)
        lin = dev.download(lin_t)
        # each ridge point's segment as an index into the sorted segment labels
        roots_t, lab_t = torch.unique(labels.index_select(0, lin_t), sorted=True,
                                      return_inverse=True)
        seg = (len(roots_t), dev.download(lab_t))
        dirs = dev.download(self._d_qeigs.reshape(n, 3).index_select(0, lin_t))
        del labels, work, d_ridge
        coords = np.stack(np.unravel_index(lin, shape), 1) if len(lin) else np.zeros((0, 3), np.int64)
        self._set_ridge(coords, shape)
        forest = None
        if self.clean_ridge:
            self._forest = None
            _, adjusted, parts = self._cleaning_signs(coords, dirs, shape)
            keep = self._segments_to_keep(seg, adjusted, parts)
            if keep is None:          # the forest's bounds did not decide: the full tree
                _, adjusted = self._tree_signs(coords, dirs, shape)
                keep = self._segments_to_keep(seg, adjusted)
            coords, dirs = coords[keep], dirs[keep]
            if self._forest is not None:
                forest = ridge_graph.restrict_forest(self._forest[1], keep)
            self._forest = None
            self._set_ridge(coords, shape)
        coords, adjusted = self._tree_signs(coords, dirs, shape, forest)
        self._flip_nonridge(coords, adjusted, shape)
        return self._d_qeigs

    def _set_ridge(self, coords, shape):
        """The ridge as voxel coordinates; the mask is built when asked for."""
        self.__ridge = None
        self.__ridge_coords = (coords, tuple(shape))
        self.__ridge_device = None

    def _set_ridge_device(self, mask_t, shape, count):
        """The ridge as a device mask (N floats); downloaded when asked for."""
        self.__ridge = None
        self.__ridge_coords = None
        self.__ridge_device = (mask_t, tuple(shape), count)

    @property
    def ridge_voxels(self):
        """Number of voxels of the (cleaned) ridge after execute()."""
        if self.__ridge is None and self.__ridge_device is not None:
            return self.__ridge_device[2]
        if self.__ridge is None and self.__ridge_coords is not None:
            return len(self.__ridge_coords[0])
        return int(np.count_nonzero(self.ridge)) if self.ridge is not None else 0

    def _segments_to_keep(self, seg, ridge_dirs, parts=None):
        """_do_clean_ridge on segment labels: seg holds each ridge point's
        segment (the segment's smallest voxel index, so increasing labels are
        scipy's label numbering), or is (number of segments, each point's
        index into the sorted labels). Returns the points of the kept
        segments, or None when `parts` leaves the choice open (see
        _kept_segments)."""
        if isinstance(seg, tuple):
            ncomp, lab_of = seg
        else:
            roots, lab_of = np.unique(seg, return_inverse=True)
            ncomp = len(roots)
        if ncomp == 0:
            return np.zeros(0, bool)
        kept = self._kept_segments(lab_of, ncomp, ridge_dirs, parts)
        return None if kept is None else kept[lab_of]

    def _kept_segments(self, lab_of, ncomp, ridge_dirs, parts=None):
        """Mask over the ncomp segments: the ceil(3 ln(e + n)) whose summed
        (voxel-size scaled) directions are longest, the lower label first on
        ties (adjust_directions.py:290-339). With `parts` (each point's piece
        of its segment between zero-product forest edges, labelled 0 .. P-1,
        whose relative
        signs the forest leaves open) a segment's length is only known to lie
        in [2 max|P| - sum|P|, sum|P|] over its pieces' sums P (triangle
        inequality); the mask is returned when those bounds settle every
        segment in or out, else None."""
        scaled = ridge_dirs * np.asarray(self.voxel_size)
        sums = np.zeros((ncomp, 3))
        for k in range(3):
            sums[:, k] = np.bincount(lab_of, weights=scaled[:, k], minlength=ncomp)
        lengths = np.linalg.norm(sums, axis=1)
        keep = int(np.ceil(3 * np.log(np.e + ncomp)))
        if getattr(self, "_verbose", False):
            print("Keeping %s of %s labels." % (min(keep, ncomp), ncomp))
        if parts is None:
            order = np.argsort(-lengths, kind="stable")
            kept = np.zeros(ncomp, bool)
            kept[order[:keep]] = True
            return kept
        p_of = np.asarray(parts, np.int64)           # piece labels 0 .. P-1
        npiece = int(p_of.max()) + 1                 # (an unused label adds a zero piece)
        psum = np.zeros((npiece, 3))
        for k in range(3):
            psum[:, k] = np.bincount(p_of, weights=scaled[:, k], minlength=npiece)
        plen = np.linalg.norm(psum, axis=1)
        pseg = np.zeros(npiece, np.int64)
        pseg[p_of] = lab_of                          # a piece lies within one segment
        tot = np.bincount(pseg, weights=plen, minlength=ncomp)
        big = np.zeros(ncomp)
        np.maximum.at(big, pseg, plen)
        multi = np.bincount(pseg, minlength=ncomp) > 1
        margin = 1e-9 * (tot + 1.0)                  # summation-order rounding
        lo = np.where(multi, np.maximum(0.0, 2 * big - tot) - margin, lengths)
        hi = np.where(multi, tot + margin, lengths)
        # others that may reach x's lower bound; others sure to beat x's upper one
        reach = ncomp - np.searchsorted(np.sort(hi), lo, "left") - 1
        beat = ncomp - np.searchsorted(np.sort(lo), hi, "right")
        sure_in, sure_out = reach < keep, beat >= keep
        if not (sure_in | sure_out).all() or sure_in.sum() != min(keep, ncomp):
            return None
        return sure_in

    def _cleaning_signs(self, coords, dirs, shape):
        """Ridge directions good for the cleaning scores, which only need
        signs that agree with the spanning tree within each 26-connected
        segment, up to one flip per segment (the score is the norm of the
        segment's summed directions). With an isotropic voxel size the tree
        restricted to a segment is the forest of its neighbour edges, and a
        sign decision flips exactly with its parent's sign (negating a float
        vector negates its dot products exactly), so the forest's signs
        suffice, unless a forest edge's dot product is zero or NaN (such a
        decision keeps the child whatever its parent's sign, so it depends on
        the direction the tree is walked). Returns (coords, signed directions,
        parts): parts is None when the signs are the tree's up to one flip
        per segment; with zero products it labels each point's piece between
        them (the pieces' relative signs stay open, see _kept_segments). A
        NaN product, or an anisotropic voxel size, takes the full tree. The
        forest's signs and pieces come from one union-find pass
        (ridge_graph.forest_pieces), whose products are formed as
        propagate_signs forms them: float32, ((a0 b0 + a1 b1) + a2 b2)."""
        vs = self.voxel_size
        if len(coords) > ridge_graph.DENSE_LIMIT and vs[0] == vs[1] == vs[2]:
            edges = ridge_graph.ridge_forest_edges(coords, shape, vs)
            # the forest's trees are the 26-connected segments: the kept
            # segments' share is the first phase of the final tree
            self._forest = (coords, edges)
            signs, pieces, nzero, nnan = ridge_graph.forest_pieces(dirs, edges)
            if not nnan:
                return (np.require(coords, np.uint16, ["C"]), signs,
                        pieces if nzero else None)
        return self._tree_signs(coords, dirs, shape) + (None,)

    def _tree_signs(self, coords, dirs, shape, forest=None):
        """Tree signs from ridge point 0; forest: the first phase of the
        tree (ridge_graph.ridge_tree_from_forest), when already known."""
        if len(coords) == 0:
            return coords.astype(np.uint16), np.zeros((0, 3), np.float32)
        if forest is not None:
            edges = ridge_graph.ridge_tree_from_forest(coords, shape, self.voxel_size, forest)
        else:
            edges = ridge_graph.ridge_tree_edges(coords, shape, self.voxel_size)
        return np.require(coords, np.uint16, ["C"]), ridge_graph.propagate_signs(dirs, edges, 0)

    def _label_ridge(self):
        dev = self._loader
        shape = self.response.shape
        d_ridge = dev.empty(int(np.prod(shape)))
        self._ridge_detector.execute(d_ridge, self._d_resp, self._d_qeigs, shape)
        ridge = dev.download(d_ridge).reshape(shape)
        if self._hysteresis:
            sure = np.logical_and(ridge, self.response >= self.threshold[1])
            ridge[:] = hysteresis_threshold(ridge, sure.astype(np.float32))
        self.__ridge = ridge
        return ridge

    def _qeigs_to_main_dirs(self):
        """Column 0 of every frame; also kept on the device for the final pass."""
        dev = self._loader
        n = self.response.size
        rc = dev.lib.mh_main_dirs(dev.handle, dev.ptr(self._d_qeigs), dev.ptr(self._d_qeigs), n,
                                  dev.stream)
        dev.check(rc, "mh_main_dirs")
        main_dirs = dev.download(self._d_qeigs).reshape(self.qeigs.shape)
        self.__main_dirs_unadjusted = main_dirs.copy()
        return main_dirs

    def _adjust_ridge_directions(self, labeled_ridge, main_dirs, for_cleaning=False):
        """Minimum spanning tree over the ridge voxels (scaled Euclidean
        distances, adjust_directions.py:365-417); signs propagated from ridge
        voxel 0 so that every tree edge has a non-negative dot product.
        Returns (uint16 coords R x 3 in C order, float32 adjusted dirs).
        for_cleaning: the directions only feed the cleaning scores, and the
        pieces of _cleaning_signs come third."""
        coords = np.argwhere(labeled_ridge != 0)
        dirs = main_dirs[tuple(coords.T)] if len(coords) else None
        if for_cleaning:
            if len(coords):
                return self._cleaning_signs(coords, dirs, labeled_ridge.shape)
            return self._tree_signs(coords, dirs, labeled_ridge.shape) + (None,)
        # the cleaning's forest, restricted by _do_clean_ridge to this ridge
        stash, self._forest = self._forest, None
        forest = None
        if stash is not None and np.array_equal(stash[0], coords):
            forest = stash[1]
        return self._tree_signs(coords, dirs, labeled_ridge.shape, forest)

    def _do_clean_ridge(self, labeled_ridge, ridge_coords, ridge_dirs, parts=None):
        """Keep the ceil(3 ln(e + n)) 26-connected ridge segments whose summed
        (voxel-size scaled) directions are longest (parts: see
        _cleaning_signs)."""
        comps, ncomp = label(labeled_ridge, structure=np.ones((3, 3, 3), dtype=int))
        if ncomp == 0:
            self.__ridge = labeled_ridge
            return labeled_ridge
        lab_of = comps[tuple(np.asarray(ridge_coords, dtype=np.intp).T)] - 1
        kept = self._kept_segments(lab_of, ncomp, ridge_dirs, parts)
        if kept is None:              # the forest's bounds did not decide: the full tree
            _, tree_dirs = self._tree_signs(np.asarray(ridge_coords, np.int64), ridge_dirs,
                                            labeled_ridge.shape)
            kept = self._kept_segments(lab_of, ncomp, tree_dirs)
        cleaned = np.isin(comps, np.nonzero(kept)[0] + 1).astype(labeled_ridge.dtype)
        stash, self._forest = self._forest, None
        if stash is not None and np.array_equal(stash[0], ridge_coords):
            # the kept points, in order: the cleaned ridge's argwhere
            keep = kept[lab_of]
            self._forest = (stash[0][keep], ridge_graph.restrict_forest(stash[1], keep))
        self.__ridge = cleaned
        return cleaned

    def _flip_nonridge(self, ridge_coords, ridge_dirs, shape):
        """Every voxel's direction takes the sign that agrees with its nearest
        (scaled Euclidean) ridge voxel, on the device."""
        dev = self._loader
        n_r = len(ridge_coords)
        if n_r == 0:
            return
        # uint16 coordinates travel as their int16 bit pattern (torch has no uint16 copy)
        raw = np.ascontiguousarray(ridge_coords, dtype=np.uint16).view(np.int16)
        d_coords = dev.torch.from_numpy(raw).to(dev.device)
        d_rdirs = dev.upload(ridge_dirs)
        scaling = _lib.F32x3(*[float(v) for v in self.voxel_size])
        wbytes = int(dev.lib.mh_nonridge_work_bytes(_lib.dims3(shape)))
        work = dev.bytes(wbytes)
        rc = dev.lib.mh_nonridge_signs(dev.handle, dev.ptr(self._d_qeigs), _lib.dims3(shape),
                                       dev.ptr(d_coords), dev.ptr(d_rdirs), n_r, scaling,
                                       dev.ptr(work), wbytes, dev.stream)
        dev.check(rc, "mh_nonridge_signs")
        # the temporaries above are released when this returns; keep them alive
        # until the kernels that read them are done
        dev.synchronize()

    def _adjust_all_directions(self, ridge_coords, ridge_dirs, shape):
        self._flip_nonridge(ridge_coords, ridge_dirs, shape)
        return self._loader.download(self._d_qeigs).reshape(self.qeigs.shape)


class TransformationCompleter(TreePrintable):
    """Directions -> orthonormal frames (det +1) -> squeezed quaternions."""

    def __init__(self, dirs, loader=None):
        self.__dirs = None
        self._loader = loader if loader is not None else default_context()
        if dirs is not None:
            self.dirs = dirs

    @property
    def dirs(self):
        return self.__dirs

    @dirs.setter
    def dirs(self, value):
        self.__dirs = np.require(value, dtype=np.float32, requirements=["C", "A"])

    def complete(self):
        """Squeezed frames with the same shape as dirs."""
        dev = self._loader
        d = dev.upload(self.dirs)
        out = self.complete_device(d)
        return dev.download(out).reshape(self.dirs.shape)

    def complete_device(self, d_dirs, out=None):
        dev = self._loader
        out = d_dirs if out is None else out
        rc = dev.lib.mh_complete_frames(dev.handle, dev.ptr(out), dev.ptr(d_dirs),
                                        d_dirs.numel() // 3, dev.stream)
        dev.check(rc, "mh_complete_frames")
        return out
