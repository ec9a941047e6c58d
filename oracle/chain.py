"""The reference experiment's per-volume chain built from the oracle's stages:
Frangi vesselness -> normalise -> direction adjustment -> GVF -> frame
completion -> capacities -> ATV max-flow cut (phantom_experiment.py:319-368).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): it is the checker of the
GPU pipeline end to end, never part of the product path.

Stage by stage:
  vesselness      ref32.vesselness (qfrangi_gpu.py:442-631)
  normalise       fp64.normalize, float64 (misc.py:63-114)
  ridges          ref32.detect_ridges on the float32 response
                  (ridge_detection.py:12-77, detect_ridges.cl:34-102)
  hysteresis      scipy 26-connected labelling (ridge_detection.py:80-131)
  ridge signs     a minimum spanning tree of the ridge voxels, signs
                  propagated from ridge point 0 by a scipy breadth-first
                  traversal (adjust_directions.py:365-417)
  cleaning        scipy 26-connected segments, scores |sum of signed
                  directions * voxel size|, the ceil(3 ln(e + n)) best kept
                  (adjust_directions.py:290-339), then the ridge signs again
  other voxels    ref32.adjust_nonridge_signs, brute force
                  (adjust_directions.cl:182-246)
  GVF             ref32.gvf of float32(dirs * response) (phantom_experiment.py:
                  335-336, gvf.cl)
  frames          ref32.complete_frames (adjust_directions.cl:68-100)
  capacities      fp64.capacities, fp64.grad_mag_sq, fp64.nonterminal_weights
                  (phantom_experiment.py:280-301, 345, 359)
  cut             ref32.cmf_solve (continuous_cut_anisotropic_gpu.py:198-344)

The one step it shares with the product is the spanning tree's edge set
(ridge_graph.ridge_tree_edges, host C++): the reference's dense pdist + MST
is O(R^2) in memory (tens of thousands of ridge voxels at 64^3), and the
host tree is pinned against scipy's dense minimum spanning tree by the CPU
tests (edge for edge with distinct lengths, equal total length with lattice
ties). The signs along it are computed here independently.
"""
from dataclasses import dataclass

import numpy as np
import scipy.ndimage as ndi
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import breadth_first_order

from . import fp64, ref32

A_ATV = 10 ** -1.5     # phantom_experiment.py:154


def tree_signs(dirs, edges, root=0):
    """Signs along a spanning tree from `root` (adjust_directions.py:384-417):
    the root keeps its direction; a child keeps its direction when its dot
    product with its adjusted parent is >= 0, else flips (a NaN product
    flips). dirs: R x 3 float32; edges: (R - 1) x 2."""
    n = len(dirs)
    out = np.array(dirs, np.float32, copy=True)
    if n < 2:
        return out
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    g = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(n, n)).tocsr()
    order, pred = breadth_first_order(g, root, directed=False, return_predecessors=True)
    if len(order) != n:
        raise ValueError("tree_signs: the edges do not span the points")
    for j in order[1:]:
        p = pred[j]
        d = dirs[j]
        dot = float(np.dot(d.astype(np.float32), out[p]))
        out[j] = d if dot >= 0 else -d
    return out


def ridge_signs(ridge, main_dirs, voxel_size=(1.0, 1.0, 1.0)):
    """(coords uint16 R x 3 in C order, signed directions float32 R x 3)."""
    from miccai2016_amd.processing import ridge_graph   # the shared host tree
    coords = np.argwhere(ridge)
    d = main_dirs[tuple(coords.T)].astype(np.float32)
    edges = ridge_graph.ridge_tree_edges(coords, ridge.shape, voxel_size)
    return coords.astype(np.uint16), tree_signs(d, edges)


def clean_ridge(ridge, coords, rdirs, voxel_size=(1.0, 1.0, 1.0)):
    """Keep the ceil(3 ln(e + n)) 26-connected segments whose summed signed
    directions (times the voxel size) are longest (adjust_directions.py:
    290-339). Returns (cleaned ridge, segment count, kept count)."""
    lab, nseg = ndi.label(ridge, structure=np.ones((3, 3, 3), int))
    seg = lab[tuple(coords.astype(np.int64).T)]
    scaled = rdirs.astype(np.float64) * np.asarray(voxel_size, np.float64)
    sums = np.zeros((nseg + 1, 3))
    np.add.at(sums, seg, scaled)
    lengths = np.linalg.norm(sums[1:], axis=1)
    order = sorted(range(1, nseg + 1), key=lambda k: lengths[k - 1], reverse=True)
    keep = int(np.ceil(3 * np.log(np.e + nseg)))
    kept = np.zeros(nseg + 1, bool)
    kept[order[:keep]] = True
    return kept[lab] & (lab > 0), nseg, min(keep, nseg)


def adjust_directions(qeig3, response, thresholds=(0.01, 0.3), voxel_size=(1.0, 1.0, 1.0),
                      clean=True):
    """AdvancedAdjuster.execute (adjust_directions.py:213-256) with hysteresis
    and cleaning. Returns (directions float32 (D0, D1, D2, 3), final ridge
    bool, info dict)."""
    resp32 = np.asarray(response, np.float32)
    lo, hi = sorted(thresholds)
    cand = ref32.detect_ridges(resp32, qeig3, lo) != 0
    sure = cand & (resp32 >= hi)
    ridge = fp64.hysteresis(cand, sure)
    md = ref32.main_dirs(qeig3)
    info = {"candidates": int(cand.sum()), "hysteresis": int(ridge.sum())}
    coords, rdirs = ridge_signs(ridge, md, voxel_size)
    if clean:
        ridge, nseg, nkeep = clean_ridge(ridge, coords, rdirs, voxel_size)
        info.update(segments=nseg, kept_segments=nkeep, cleaned=int(ridge.sum()))
        coords, rdirs = ridge_signs(ridge, md, voxel_size)
    out = ref32.adjust_nonridge_signs(md, coords, rdirs, np.asarray(voxel_size, np.float32))
    return out, ridge, info


@dataclass
class ChainPrepared:
    vesselness: np.ndarray   # normalised, float64
    qeig: np.ndarray         # squeezed frames of the winning scale, float32 (D0, D1, D2, 3)
    ridge: np.ndarray        # final (cleaned) ridge, bool
    dirs: np.ndarray         # adjusted main directions, float32 (D0, D1, D2, 3)
    quats: np.ndarray        # completed GVF frames, float32 (D0, D1, D2, 3)
    cs: np.ndarray           # float32
    ct: np.ndarray           # float32
    gms: np.ndarray          # |grad I|^2, float64
    info: dict


def prepare(img, noise_level, scales=(1.0, 2.0, 3), gvf_mu=0.1, gvf_iters=200,
            thresholds=(0.01, 0.3), voxel_size=(1.0, 1.0, 1.0), clean=True):
    """Everything the cuts of one noisy volume share (phantom_experiment.py:
    323-345)."""
    img = np.asarray(img)
    val, qeig = ref32.vesselness(img.astype(np.float32), fp64.scales_voxels(scales, voxel_size))
    ves = fp64.normalize(val)
    dirs, ridge, info = adjust_directions(qeig, ves, thresholds, voxel_size, clean)
    v = (dirs.astype(np.float64) * ves[..., None]).astype(np.float32)
    g, _ = ref32.gvf(v, gvf_mu, gvf_iters, voxel_size)
    quats = ref32.complete_frames(g)
    ct, cs = fp64.capacities(img, noise_level)
    return ChainPrepared(ves, qeig, ridge, dirs, quats, cs.astype(np.float32),
                         ct.astype(np.float32), fp64.grad_mag_sq(img), info)


def cut(img, prep, weighting, decay, num_iter=500, update=30, a=A_ATV, steps=0.16):
    """One cut (phantom_experiment.py:358-362): returns (u, iterations)."""
    alpha = fp64.nonterminal_weights(prep.gms, weighting, decay).astype(np.float32)
    u, done, _ = ref32.cmf_solve(np.asarray(img, np.float32), prep.quats, (1.0, a, a), alpha,
                                 prep.cs, prep.ct, steps=steps, num_iter=num_iter, update=update,
                                 error_bound=0)
    return u, done
