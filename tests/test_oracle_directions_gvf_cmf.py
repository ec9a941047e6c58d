"""Pins the oracle's direction, GVF and max-flow stages against closed forms,
brute force, library routines and invariants. CPU only."""
import numpy as np
import pytest
import scipy.ndimage as ndi
from scipy.spatial.transform import Rotation

from oracle import fp64, ref32


def _squeezed(rot):
    return ref32.qsqueeze(rot.as_quat()[..., [3, 0, 1, 2]].astype(np.float32))


# --------------------------------------------------------------- ridges ----

def test_ridge_of_straight_tube_is_its_centre_line():
    # a Gaussian tube along axis 0 with the frame (e0 along the tube)
    z = np.arange(15.0)
    _, y, x = np.meshgrid(z, z, z, indexing="ij")
    resp = np.exp(-((y - 7) ** 2 + (x - 7) ** 2) / 4.0).astype(np.float32)
    q = _squeezed(Rotation.identity(15 ** 3)).reshape(15, 15, 15, 3)
    r = ref32.detect_ridges(resp, q, 0.01)
    assert r[:, 7, 7].all()
    assert r.sum() == 15


def test_ridge_sampling_truncates_toward_zero():
    # response decreasing along axis 1; frame rotated 30 degrees about axis 0,
    # so at i1 = 0 some samples fall at fractional negative axis-1 positions.
    # Truncation toward zero keeps both corners inside ({0, 1}) and
    # extrapolates the ramp to a larger value: not a ridge voxel. Flooring
    # would mix in outside corners (value -1) and make it one.
    z = np.arange(6.0)
    _, b, _ = np.meshgrid(z, z, z, indexing="ij")
    resp = (0.5 - 0.01 * b).astype(np.float32)
    rot = Rotation.from_euler("x", np.full(6 ** 3, 30.0), degrees=True)
    q = _squeezed(rot).reshape(6, 6, 6, 3)
    r = ref32.detect_ridges(resp, q, -10.0)
    assert r[2, 0, 2] == 0
    # the same voxel with an outside value below it along axis 1 is a ridge
    # voxel under the exact (non-extrapolating) reading: check the sampled
    # maximum the oracle must have used, f at the most negative position
    c, s_ = np.cos(np.pi / 6), np.sin(np.pi / 6)
    assert 0.5 + 0.01 * (c + s_) > 0.5


def test_ridge_threshold_is_a_double_comparison():
    resp = np.zeros((3, 3, 3), np.float32)
    resp[1, 1, 1] = np.float32(0.01)   # below 0.01 in double
    q = _squeezed(Rotation.identity(27)).reshape(3, 3, 3, 3)
    assert ref32.detect_ridges(resp, q, 0.01)[1, 1, 1] == 0
    resp[1, 1, 1] = np.nextafter(np.float32(0.01), np.float32(1))
    assert ref32.detect_ridges(resp, q, 0.01)[1, 1, 1] == 1


# ------------------------------------------------------------ main dirs ----

def test_main_dirs_are_first_frame_column(rng):
    rot = Rotation.random(500, random_state=11)
    v = _squeezed(rot)
    d = ref32.main_dirs(v)
    np.testing.assert_allclose(d, rot.as_matrix()[:, :, 0], atol=2e-6)


# ------------------------------------------------------ nearest ridge signs ----

def test_nonridge_signs_match_brute_force_argmin(rng):
    shape = (7, 8, 9)
    dirs = rng.normal(size=shape + (3,)).astype(np.float32)
    coords = np.stack([rng.integers(0, s, 25) for s in shape], -1).astype(np.uint16)
    coords = np.unique(coords, axis=0)
    rdirs = rng.normal(size=(len(coords), 3)).astype(np.float32)
    for scaling in [(1, 1, 1), (0.5, 1.0, 2.0)]:
        got = ref32.adjust_nonridge_signs(dirs, coords, rdirs, scaling)
        want = fp64.nearest_ridge_signs(dirs, coords, rdirs.astype(np.float64), scaling)
        np.testing.assert_array_equal(got, want)


def test_nonridge_ties_pick_lowest_index():
    dirs = np.zeros((1, 1, 3, 3), np.float32)
    dirs[..., 0] = 1
    coords = np.array([[0, 0, 0], [0, 0, 2]], np.uint16)      # voxel 1 is equidistant
    rdirs = np.array([[-1, 0, 0], [1, 0, 0]], np.float32)
    got = ref32.adjust_nonridge_signs(dirs, coords, rdirs)
    assert got[0, 0, 1, 0] == -1      # aligned with ridge point 0


# ------------------------------------------------------- frame completion ----

def test_complete_frames_invariants(rng):
    v = rng.normal(size=(2000, 3)).astype(np.float32)
    q = ref32.complete_frames(v)
    R = fp64.squeezed_to_mat(q)
    np.testing.assert_allclose(R[:, :, 0], v / np.linalg.norm(v, axis=1, keepdims=True), atol=3e-6)
    np.testing.assert_allclose(np.einsum("nji,njk->nik", R, R), np.broadcast_to(np.eye(3), R.shape),
                               atol=3e-6)
    np.testing.assert_allclose(np.linalg.det(R), 1.0, atol=3e-6)


def test_complete_frames_zero_vector_is_identity():
    np.testing.assert_array_equal(ref32.complete_frames(np.zeros((1, 3), np.float32)), 0.0)


# ------------------------------------------------------------------ GVF ----

def test_gvf_mu_zero_closed_form(rng):
    # |v|^2 <= 0.75 keeps dt * r below 2 (stable explicit Euler)
    v = rng.uniform(-0.5, 0.5, size=(5, 6, 7, 3)).astype(np.float32)
    g, _ = ref32.gvf(v, mu=1e-30, maxiter=15)
    r = (v.astype(np.float64) ** 2).sum(-1, keepdims=True)
    dt = ref32.gvf_time_step((1, 1, 1), 1e-30)
    want = v * (1 - (1 - dt * r) ** 15)
    np.testing.assert_allclose(g, want, atol=2e-6)


def test_gvf_constant_unit_field_is_fixed_point():
    v = np.zeros((6, 6, 6, 3), np.float32)
    v[..., 1] = 1.0
    g, _ = ref32.gvf(v, mu=0.1, maxiter=40)
    np.testing.assert_allclose(g[..., 1], 1 - (1 - 0.99) ** 40, atol=1e-6)
    np.testing.assert_array_equal(g[..., 0], 0.0)


def test_gvf_matches_fp64(rng):
    v = ndi.gaussian_filter(rng.normal(size=(8, 9, 10, 3)), (1, 1, 1, 0)).astype(np.float32)
    g, mags = ref32.gvf(v, mu=0.3, maxiter=25, update=5)
    want = fp64.gvf(v, 0.3, 25)
    np.testing.assert_allclose(g, want, atol=2e-5 * np.abs(want).max())
    assert len(mags) == 5 and np.all(np.isfinite(mags))


# ------------------------------------------------------------- max-flow ----

def _cut_inputs(rng, shape=(8, 9, 10)):
    img = (ndi.gaussian_filter(rng.normal(size=shape), 1.0) * 3 + 0.5).astype(np.float32)
    quats = _squeezed(Rotation.random(int(np.prod(shape)), random_state=2)).reshape(shape + (3,))
    ct, cs = fp64.capacities(img, 0.8)
    return img, quats, cs.astype(np.float32), ct.astype(np.float32)


def test_cmf_matches_fp64(rng):
    img, q, cs, ct = _cut_inputs(rng)
    eigs = (1.0, 10 ** -1.5, 10 ** -1.5)
    u32, done, _ = ref32.cmf_solve(img, q, eigs, 0.05, cs, ct, steps=0.16, num_iter=120, update=30,
                                   error_bound=0)
    u64 = fp64.cmf(img, q, eigs, 0.05, cs, ct, 120, steps=0.16)
    assert done == 120
    assert np.abs(u32 - u64).max() < 1e-4


def test_cmf_alpha_zero_is_pointwise_cut(rng):
    img, q, cs, ct = _cut_inputs(rng)
    u, _, _ = ref32.cmf_solve(img, q, (1, 1, 1), 0.0, cs, ct, steps=0.16, num_iter=400, update=50,
                              error_bound=0)
    sure = np.abs(cs - ct) > 0.05
    np.testing.assert_array_equal((u >= 0.5)[sure], (cs > ct)[sure])


def test_cmf_tv_ignores_the_frame_field(rng):
    img, q, cs, ct = _cut_inputs(rng)
    q2 = _squeezed(Rotation.random(q.size // 3, random_state=99)).reshape(q.shape)
    a, _, _ = ref32.cmf_solve(img, q, (1, 1, 1), 0.05, cs, ct, steps=0.16, num_iter=60, update=60,
                              error_bound=0)
    b, _, _ = ref32.cmf_solve(img, q2, (1, 1, 1), 0.05, cs, ct, steps=0.16, num_iter=60, update=60,
                              error_bound=0)
    assert np.abs(a - b).max() < 2e-5


def test_cmf_step_invariants(rng):
    img, q, cs, ct = _cut_inputs(rng)
    eigs = (1.0, 0.1, 0.1)
    P = ref32.cut_params((1, 1, 1), 0.16, 0.2, eigs)
    shape = img.shape
    hp = rng.normal(size=(3,) + shape).astype(np.float32)
    tmp = rng.normal(size=shape).astype(np.float32)
    alpha = np.full(shape, 0.07, np.float32)
    hp2 = ref32.update_p(hp, tmp, q, alpha, P)
    # |H^-1 hp| <= alpha after the projection
    R = fp64.squeezed_to_mat(q)
    p = np.einsum("...ji,j...->...i", R, hp2.astype(np.float64)) / np.sqrt(eigs)
    assert (np.linalg.norm(p, axis=-1) <= 0.07 * (1 + 1e-5)).all()
    pt = np.minimum(cs, ct)
    u = ref32.cmf_initial_u(img)
    tmp2, pt2, u2 = ref32.update_u(hp2, tmp, cs, ct, pt, u, P)
    assert (pt2 <= ct).all()


def test_cmf_nan_frame_stops_at_first_check(rng):
    img, q, cs, ct = _cut_inputs(rng)
    q = q.copy()
    q[2, 3, 4, 0] = np.nan          # NaN residual -> NaN criterion
    u, done, crit = ref32.cmf_solve(img, q, (1, 1, 1), 0.05, cs, ct, num_iter=50, update=10,
                                    error_bound=0)
    assert done == 1 and len(crit) == 1 and np.isnan(crit[0])


def test_cmf_nan_capacity_is_dropped_by_min(rng):
    # min(x, NaN) = x in the kernel language: a NaN capacity does not spread
    img, q, cs, ct = _cut_inputs(rng)
    cs = cs.copy()
    cs[2, 3, 4] = np.nan
    u, done, crit = ref32.cmf_solve(img, q, (1, 1, 1), 0.05, cs, ct, num_iter=20, update=10,
                                    error_bound=0)
    assert done == 20 and np.isfinite(u).all()


def test_cmf_converged_stop_iteration_count(rng):
    img, q, cs, ct = _cut_inputs(rng)
    u, done, crit = ref32.cmf_solve(img, q, (1, 1, 1), 0.05, cs, ct, num_iter=2000, update=7,
                                    error_bound=1e-3)
    assert done < 2000 and (done - 1) % 7 == 0 and crit[-1] < 1e-3 and (crit[:-1] >= 1e-3).all()


def test_sum_abs_reduce_matches_exact_sum(rng):
    for n in [5, 1000, 1025, 4096, 100003]:
        x = rng.normal(size=n).astype(np.float32)
        exact = np.sum(np.abs(x.astype(np.float64)))
        assert abs(ref32.sum_abs_reduce(x) - exact) <= 1e-6 * exact
