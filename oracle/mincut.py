"""Brute-force minimum cuts of 1-D chains: small max-flow problems whose
converged continuous cut is known exactly.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). On a chain along one axis
the ATV energy of a labelling u is sum (1 - u) cs + u ct + sum_i w_i
|u_{i+1} - u_i|, with w_i = alpha_i |H_i^T e_axis| and H = R(q) diag(sqrt l)
(the support function of the flow bound |H^-1 hp| <= alpha of update_qp.cl:
150-232 along the chain's only difference), so by the coarea formula the
thresholded minimiser of the relaxation is a binary minimum cut. Every
labelling of up to 12 voxels is enumerated; a case is kept only if its best
labelling beats the next by a margin, so a converged solver must reproduce
it exactly.
"""
import itertools

import numpy as np
from scipy.spatial.transform import Rotation

from . import fp64, ref32


def _squeezed(rot):
    return ref32.qsqueeze(rot.as_quat()[..., [3, 0, 1, 2]].astype(np.float32))


def _chain_problem(seed, n, axis, atv):
    """Random 1-D chain along `axis` whose binary minimum cut beats every other
    labelling by a margin. Energy of a labelling b (u = 1 pays ct, u = 0 pays
    cs; edge (i, i+1) pays the flow bound of voxel i, alpha_i |H_i^T e_axis|):
    sum (1 - b) cs + b ct + sum_i w_i |b_{i+1} - b_i|."""
    rng = np.random.default_rng(seed)
    shape = [1, 1, 1]
    shape[axis] = n
    shape = tuple(shape)
    cs = rng.uniform(0.0, 1.0, n)
    ct = rng.uniform(0.0, 1.0, n)
    alpha = rng.uniform(0.05, 0.6, n)
    eigs = (1.0, 10 ** -1.5, 10 ** -1.5) if atv else (1.0, 1.0, 1.0)
    rot = Rotation.random(n, random_state=seed)
    q = _squeezed(rot)
    R = fp64.squeezed_to_mat(q)
    w = alpha * np.sqrt((np.asarray(eigs) * R[:, axis, :] ** 2).sum(1))
    cs32, ct32, a32 = (x.astype(np.float32) for x in (cs, ct, alpha))
    best = []
    for bits in itertools.product((0, 1), repeat=n):
        b = np.asarray(bits, np.float64)
        e = ((1 - b) * cs32 + b * ct32).sum() + (w[:-1] * np.abs(np.diff(b))).sum()
        best.append((e, bits))
    best.sort()
    margin = best[1][0] - best[0][0]
    return (shape, q.reshape(shape + (3,)), eigs, cs32.reshape(shape), ct32.reshape(shape),
            a32.reshape(shape), np.asarray(best[0][1], bool).reshape(shape), margin)


def chain_cases():
    """(shape, frames, eigs, cs, ct, alpha, min-cut labelling) of 1-D chains
    along each axis, TV and ATV, with a margin of at least 0.03."""
    cases = []
    seed = 0
    for axis in range(3):
        for atv in (False, True):
            for n in (6, 9, 12):
                while True:
                    seed += 1
                    case = _chain_problem(seed, n, axis, atv)
                    if case[-1] >= 0.03:
                        cases.append(case[:-1])
                        break
    return cases


def case_id(case):
    return f"{case[0]}-{'atv' if case[2][1] < 1 else 'tv'}"
