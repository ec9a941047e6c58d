"""mh_cmf_solve_noalloc (the max-flow solve whose workspace is a static device
array) against mh_cmf_solve with an allocated workspace, through the C-ABI:
two calls in a row, the second on a smaller volume, each bit for bit."""
import ctypes as C

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from miccai2016_amd import _lib
from miccai2016_amd import phantom as ph
from oracle import ref32

pytestmark = pytest.mark.gpu


def _inputs(shape, seed, pointwise=False, supervised=False):
    spec = ph.HelixSpec(shape, 3.0, min(shape[1:]) * 0.25, 14.0, 0.8, seed=seed)
    img, _ = ph.make_helix(spec)
    n = img.size
    q = ref32.qsqueeze(Rotation.random(n, random_state=seed).as_quat()[:, [3, 0, 1, 2]]
                       .astype(np.float32)).reshape(shape + (3,))
    cs, ct = ph.terminal_capacities(img, spec.sigma)
    alpha = ph.nonterminal_weight(ph.grad_mag_sq(img), 10 ** -1.625, 10 ** 1.5).astype(np.float32)
    if pointwise:
        rng = np.random.default_rng(seed)
        ev = rng.uniform(0.05, 1.0, size=shape + (3,))
        eigs = np.concatenate([np.sqrt(ev), q], axis=-1).astype(np.float32)
    else:
        eigs = q
    labels = None
    if supervised:
        labels = np.zeros(shape, np.uint8)
        labels[img > 1.5] = 1
        labels[img < -1.5] = 2
    u0 = ref32.cmf_initial_u(img)
    return u0, eigs, alpha, cs.astype(np.float32), ct.astype(np.float32), labels


def _params(num_iter, update, const_evals):
    P = _lib.CmfParams()
    P.rcc = 5.0
    for k in range(3):
        P.rs[k] = 1.0
        P.ss[k] = 0.16
    sq = np.sqrt([1.0, 10 ** -1.5, 10 ** -1.5])
    for k in range(3):
        P.sqev[k] = float(sq[k])
    P.const_evals = int(const_evals)
    P.num_iter = num_iter
    P.update = update
    P.error_bound = 0.0
    return P


def _solve(dev, fields, shape, P, noalloc):
    u0, eigs, alpha, cs, ct, labels = fields
    t = [dev.upload(a) for a in (u0, eigs, alpha, cs, ct)]
    lab = None if labels is None else dev.upload(labels, dtype=np.uint8)
    ncheck = max(1, -(-P.num_iter // P.update))
    crit = np.full(ncheck, -1.0, np.float32)
    res = _lib.CmfResult()
    args = [dev.handle, dev.ptr(t[0]), dev.ptr(t[1]), dev.ptr(t[2]), dev.ptr(t[3]), dev.ptr(t[4]),
            dev.ptr(lab), _lib.dims3(shape), C.byref(P)]
    if noalloc:
        rc = dev.lib.mh_cmf_solve_noalloc(*args, C.byref(res), crit.ctypes.data_as(C.c_void_p),
                                          dev.stream)
        dev.check(rc, "mh_cmf_solve_noalloc")
    else:
        wbytes = dev.lib.mh_cmf_workspace_bytes(int(np.prod(shape)))
        work = dev.bytes(wbytes)
        rc = dev.lib.mh_cmf_solve(*args, dev.ptr(work), wbytes, C.byref(res),
                                  crit.ctypes.data_as(C.c_void_p), dev.stream)
        dev.check(rc, "mh_cmf_solve")
    u = dev.download(t[0])
    return u, (res.iterations, res.stopped, res.ncrit), crit[:res.ncrit].copy()


CASES = [
    # (shape, iterations, update, pointwise eigenvalues, supervised): the
    # first reaches the multi-level sweep (>= 800,000 voxels), the second is
    # smaller and runs the per-iteration kernels with pointwise eigenvalues
    # and seeds
    ((100, 100, 100), 61, 30, False, False),
    ((40, 33, 50), 45, 7, True, True),
]


def test_noalloc_equals_allocating_solve_twice_in_a_row(dev):
    fields = [_inputs(c[0], 11 + k, c[3], c[4]) for k, c in enumerate(CASES)]
    want = []
    for (shape, iters, upd, pw, _), f in zip(CASES, fields):
        print(f"allocating solve {shape} x {iters}", flush=True)
        want.append(_solve(dev, f, shape, _params(iters, upd, not pw), noalloc=False))
    for (shape, iters, upd, pw, _), f, w in zip(CASES, fields, want):
        print(f"no-allocation solve {shape} x {iters}", flush=True)
        got = _solve(dev, f, shape, _params(iters, upd, not pw), noalloc=True)
        np.testing.assert_array_equal(got[0], w[0])
        assert got[1] == w[1]
        np.testing.assert_array_equal(got[2], w[2])
        assert np.isfinite(got[0]).all()


def test_noalloc_rejects_volumes_beyond_its_limit(dev):
    # checked on the host before anything is launched
    shape = (1, 1, (1 << 20) + 1)
    P = _params(10, 5, True)
    res = _lib.CmfResult()
    null = C.c_void_p(0)
    one = dev.zeros(1)
    p = dev.ptr(one)
    rc = dev.lib.mh_cmf_solve_noalloc(dev.handle, p, p, p, p, p, null, _lib.dims3(shape),
                                      C.byref(P), C.byref(res), null, dev.stream)
    assert rc == _lib.MH_ERR_ARG
    assert b"no-allocation limit" in dev.lib.mh_last_error(dev.handle)
