"""Pins the oracle's max-flow and GVF stencils by mathematics, not by a second
reading of the reference: numpy.diff for the forward difference, the adjoint
identity and an explicit sparse transpose for the divergence, scipy's mirrored
Laplacian for GVF, the linear part of the p-step against R diag(l) R^T, and the
converged cut of 1-D chains against a brute-force minimum cut. CPU only."""
import numpy as np
import pytest
import scipy.ndimage as ndi
import scipy.sparse as sp
from scipy.spatial.transform import Rotation

from oracle import fp64, mincut, ref32

SHAPES = [(5, 7, 9), (1, 6, 4), (3, 1, 1), (2, 2, 2), (1, 1, 11), (4, 3, 2)]
WEIGHTS = [(1.0, 1.0, 1.0), (0.3, 1.7, 2.5)]


def _squeezed(rot):
    return ref32.qsqueeze(rot.as_quat()[..., [3, 0, 1, 2]].astype(np.float32))


def _fwd_matrix(shape, axis):
    """Sparse forward difference along `axis` of a C-order volume, zero rows on
    the axis's last index: (D x)[i] = x[i + e_axis] - x[i]."""
    mats = []
    for a, n in enumerate(shape):
        if a == axis:
            d = sp.diags([-np.ones(n), np.ones(n - 1)], [0, 1], shape=(n, n), format="lil")
            d[n - 1, :] = 0
            mats.append(sp.csr_matrix(d))
        else:
            mats.append(sp.identity(n, format="csr"))
    return sp.kron(sp.kron(mats[0], mats[1]), mats[2], format="csr")


# ----------------------------------------------------- forward difference ----

@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("w", WEIGHTS)
def test_grad_fwd_is_weighted_numpy_diff(rng, shape, w):
    x = rng.normal(size=shape).astype(np.float32)
    g = ref32.grad_fwd(x, w)
    for k in range(3):
        want = np.zeros(shape, np.float32)
        if shape[k] > 1:
            d = np.diff(x, axis=k)                       # float32 differences
            sl = [slice(None)] * 3
            sl[k] = slice(0, shape[k] - 1)
            want[tuple(sl)] = (d.astype(np.float64) * w[k]).astype(np.float32)
        np.testing.assert_array_equal(g[k], want)


# ------------------------------------------------------------- divergence ----

@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("w", WEIGHTS)
def test_divergence_is_negative_adjoint_of_grad(rng, shape, w):
    # <grad+ x, y>_w = -<x, div_w y> for the same weights: a wrong index, a
    # dropped face term or a swapped axis in either stencil breaks it
    x = rng.normal(size=shape).astype(np.float32)
    y = rng.normal(size=(3,) + shape).astype(np.float32)
    lhs = float((ref32.grad_fwd(x, w).astype(np.float64) * y).sum())
    rhs = -float((x.astype(np.float64) * ref32.divergence(y, w)).sum())
    scale = float(np.abs(x).sum() * np.abs(y).sum(0).max() * max(w))
    assert abs(lhs - rhs) <= 1e-6 * scale


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("w", WEIGHTS)
def test_divergence_is_minus_sparse_transpose(rng, shape, w):
    y = rng.normal(size=(3,) + shape).astype(np.float32)
    want = np.zeros(int(np.prod(shape)))
    for k in range(3):
        want -= w[k] * (_fwd_matrix(shape, k).T @ y[k].astype(np.float64).ravel())
    got = ref32.divergence(y, w).ravel()
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-6 * max(1.0, np.abs(want).max()))


def test_sparse_forward_matrix_matches_numpy_diff(rng):
    # the sparse construction used above is itself checked against numpy
    x = rng.normal(size=(4, 5, 6))
    for k in range(3):
        want = np.zeros_like(x)
        sl = [slice(None)] * 3
        sl[k] = slice(0, x.shape[k] - 1)
        want[tuple(sl)] = np.diff(x, axis=k)
        np.testing.assert_allclose((_fwd_matrix(x.shape, k) @ x.ravel()).reshape(x.shape), want,
                                   atol=1e-12)


# ------------------------------------------------- the stencils in the steps ----

@pytest.mark.parametrize("shape", [(5, 7, 9), (1, 1, 8), (3, 4, 1)])
def test_update_p_linear_part_is_r_lambda_rt(rng, shape):
    # without the projection (alpha huge), hp' = H (H^-1 hp + H^T g) = hp + A g
    # with A = R diag(l) R^T and g = SS * grad+ tmp. A transposed frame gives
    # R^T diag(l) R instead and fails.
    n = int(np.prod(shape))
    eigs = (1.0, 10 ** -1.5, 0.3)
    steps = 0.16
    P = ref32.cut_params((1, 1, 1), steps, 0.2, eigs)
    q = _squeezed(Rotation.random(n, random_state=5)).reshape(shape + (3,))
    hp = rng.normal(size=(3,) + shape).astype(np.float32)
    tmp = rng.normal(size=shape).astype(np.float32)
    alpha = np.full(shape, 1e30, np.float32)
    got = ref32.update_p(hp, tmp, q, alpha, P).astype(np.float64)
    R = Rotation.from_quat(fp64.qexpand(q.reshape(-1, 3))[:, [1, 2, 3, 0]]).as_matrix()
    A = R @ np.diag(eigs) @ np.transpose(R, (0, 2, 1))
    g = np.zeros((3, n))
    for k in range(3):
        g[k] = steps * (_fwd_matrix(shape, k) @ tmp.astype(np.float64).ravel())
    want = hp.reshape(3, n).astype(np.float64) + np.einsum("nij,jn->in", A, g)
    np.testing.assert_allclose(got.reshape(3, n), want, atol=2e-5)


def test_update_p_projection_keeps_direction(rng):
    shape = (6, 5, 4)
    n = int(np.prod(shape))
    eigs = (1.0, 0.1, 0.1)
    P = ref32.cut_params((1, 1, 1), 0.16, 0.2, eigs)
    q = _squeezed(Rotation.random(n, random_state=7)).reshape(shape + (3,))
    hp = rng.normal(size=(3,) + shape).astype(np.float32)
    tmp = rng.normal(size=shape).astype(np.float32)
    big = ref32.update_p(hp, tmp, q, np.full(shape, 1e30, np.float32), P).astype(np.float64)
    small = ref32.update_p(hp, tmp, q, np.full(shape, 0.05, np.float32), P).astype(np.float64)
    R = fp64.squeezed_to_mat(q).reshape(n, 3, 3)
    to_p = lambda h: np.einsum("nji,jn->ni", R, h.reshape(3, n)) / np.sqrt(eigs)   # H^-1 h
    pb, ps = to_p(big), to_p(small)
    nb = np.linalg.norm(pb, axis=1)
    proj = nb > 0.05
    assert proj.sum() > n // 2
    # projected: on the sphere of radius alpha, same direction as unprojected
    np.testing.assert_allclose(np.linalg.norm(ps[proj], axis=1), 0.05, rtol=1e-5)
    cos = (pb[proj] * ps[proj]).sum(1) / (nb[proj] * 0.05)
    np.testing.assert_allclose(cos, 1.0, atol=1e-5)
    np.testing.assert_allclose(ps[~proj], pb[~proj], atol=1e-6)


@pytest.mark.parametrize("shape", [(5, 7, 9), (1, 9, 1), (2, 2, 3)])
def test_update_u_residual_is_the_divergence(rng, shape):
    # cs = ct = pt = 0 and u inside (div/rcc, div/rcc + 1) keep both clamps
    # active, so err = div hp: tmp' = div hp and u' = u - div hp / rcc, with
    # div = -sum_k RS_k D_k^T hp_k from the sparse forward difference
    scaling = (1.0, 1.5, 0.8)
    P = ref32.cut_params(scaling, 0.16, 0.2, (1, 1, 1))
    n = int(np.prod(shape))
    hp = rng.normal(size=(3,) + shape).astype(np.float32)
    div = np.zeros(n)
    for k in range(3):
        div -= (_fwd_matrix(shape, k).T @ hp[k].astype(np.float64).ravel()) / scaling[k]
    div = div.reshape(shape)
    zero = np.zeros(shape, np.float32)
    u = (div / P.rcc + 0.5).astype(np.float32)
    tmp, pt, u2 = ref32.update_u(hp, zero, zero, zero, zero, u, P)
    np.testing.assert_allclose(tmp, div, atol=3e-6 * max(1, np.abs(div).max()))
    np.testing.assert_array_equal(pt, 0)
    np.testing.assert_allclose(u2, u - div / P.rcc, atol=2e-6)


def test_supervised_labels_fix_the_seeds(rng):
    # label bit 0 = foreground seed, bit 1 = background seed
    # (update_qp.cl:278-284): with alpha = 0 every voxel is its own cut, and a
    # seed takes its label against the capacities that would decide it
    # otherwise
    shape = (4, 5, 6)
    img = rng.uniform(0.1, 1.0, size=shape).astype(np.float32)
    q = np.zeros(shape + (3,), np.float32)
    cs = rng.uniform(0.0, 1.0, size=shape).astype(np.float32)
    ct = rng.uniform(0.0, 1.0, size=shape).astype(np.float32)
    labels = rng.integers(0, 3, size=shape).astype(np.uint8)
    u, _, _ = ref32.cmf_solve(img, q, (1, 1, 1), 0.0, cs, ct, labels=labels, steps=0.16,
                              num_iter=600, update=600, error_bound=0)
    free, _, _ = ref32.cmf_solve(img, q, (1, 1, 1), 0.0, cs, ct, steps=0.16, num_iter=600,
                                 update=600, error_bound=0)
    sure = np.abs(cs - ct) > 0.05
    assert ((u >= 0.5) == (labels == 1))[labels > 0].all()
    np.testing.assert_array_equal((u >= 0.5)[(labels == 0) & sure], (free >= 0.5)[(labels == 0) & sure])
    np.testing.assert_array_equal((free >= 0.5)[sure], (cs > ct)[sure])


# ------------------------------------------------------------------- GVF ----

GVF_SHAPES = [(5, 7, 9), (2, 3, 2), (2, 2, 2), (6, 2, 5)]


@pytest.mark.parametrize("shape", GVF_SHAPES)
def test_gvf_laplacian_is_scipy_mirror_laplace(rng, shape):
    g = rng.normal(size=shape).astype(np.float32)
    want = ndi.laplace(g.astype(np.float64), mode="mirror")
    np.testing.assert_allclose(ref32.gvf_laplacian(g), want, atol=4e-6 * np.abs(want).max())
    rss = (0.25, 1.0, 4.0)
    want = sum(rss[k] * ndi.correlate1d(g.astype(np.float64), [1.0, -2.0, 1.0], axis=k,
                                        mode="mirror") for k in range(3))
    np.testing.assert_allclose(ref32.gvf_laplacian(g, rss), want, atol=4e-6 * np.abs(want).max())


@pytest.mark.parametrize("shape", GVF_SHAPES)
def test_gvf_second_step_applies_the_mirrored_laplacian(rng, shape):
    # r = 1: g1 = dt v, g2 = dt (2 - dt) v + mu dt^2 lap(v), lap = scipy's
    # mirrored Laplacian weighted by 1 / voxel_size^2
    vs = (1.0, 0.8, 1.25)
    mu = 0.2
    v = rng.normal(size=shape + (3,)).astype(np.float32)
    g, _ = ref32.gvf(v, mu=mu, maxiter=2, voxel_size=vs, response=np.ones(shape, np.float32))
    dt = ref32.gvf_time_step(np.asarray(vs), mu)
    for c in range(3):
        vc = v[..., c].astype(np.float64)
        lap = sum(ndi.correlate1d(vc, [1.0, -2.0, 1.0], axis=k, mode="mirror") / vs[k] ** 2
                  for k in range(3))
        want = dt * (2 - dt) * vc + mu * dt * dt * lap
        np.testing.assert_allclose(g[..., c], want, atol=3e-6 * np.abs(want).max())


# ------------------------------------------------- brute-force 1-D min cut ----

@pytest.mark.parametrize("case", mincut.chain_cases(), ids=mincut.case_id)
def test_converged_cut_is_the_brute_force_min_cut(case):
    shape, q, eigs, cs, ct, alpha, want = case
    img = np.linspace(0.2, 1.0, int(np.prod(shape))).reshape(shape).astype(np.float32)
    u, done, _ = ref32.cmf_solve(img, q, eigs, alpha, cs, ct, steps=0.16, num_iter=4000,
                                 update=4000, error_bound=0)
    assert done == 4000
    np.testing.assert_array_equal(u >= 0.5, want)


def test_ridge_margin_decides_as_the_ridge_test(rng):
    """ref32.ridge_margin (resp - the ridge test's sample maximum) is positive
    exactly where a voxel above the threshold is a ridge candidate, away from
    the voxels it marks as ties (0)."""
    r = rng.random((12, 13, 14)).astype(np.float32)
    q = (rng.random((12, 13, 14, 3)) * 3).astype(np.float32)
    cand = ref32.detect_ridges(r, q, 0.01) != 0
    m = ref32.ridge_margin(r, q)
    decided = m != 0
    assert decided.mean() > 0.99
    assert np.array_equal(cand[decided], ((r > 0.01) & (m > 0))[decided])


def test_ridge_margin_marks_the_truncation_jump(rng):
    """Identity frames sample the in-plane neighbours at pos +- e1 +- e2: on
    the faces i1 = 0 and i2 = 0 a sample coordinate is exactly -1, where the
    reference's truncation toward zero jumps, so those decisions are ties;
    interior voxels are not."""
    r = rng.random((6, 6, 6)).astype(np.float32)
    m = ref32.ridge_margin(r, np.zeros((6, 6, 6, 3), np.float32))
    assert np.all(m[:, 0, :] == 0) and np.all(m[:, :, 0] == 0)
    assert np.all(m[:, 1:, 1:] != 0)
