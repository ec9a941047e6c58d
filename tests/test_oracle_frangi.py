"""Pins the oracle's Frangi stages (oracle/miccai_oracle.c) against library
routines (scipy.ndimage, numpy.linalg.eigh, scipy.spatial.transform),
closed forms and invariants. CPU only."""
import numpy as np
import pytest
import scipy.ndimage as ndi
from scipy.spatial.transform import Rotation

from oracle import fp64, ref32


# ------------------------------------------------------------------ taps ----

@pytest.mark.parametrize("sigma", [1.0, 1.5, 2.0, 3.3])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_taps_match_scipy_impulse_response(sigma, order):
    # the impulse response of scipy's gaussian_filter1d (truncate=4, mirror) is
    # the convolution kernel itself; sigma^order is the scale-space factor
    R = int(4.0 * sigma + 0.5)
    n = 4 * R + 9
    imp = np.zeros(n)
    imp[n // 2] = 1.0
    resp = ndi.gaussian_filter1d(imp, sigma, order=order, mode="mirror", truncate=4.0)
    scipy_taps = resp[n // 2 - R:n // 2 + R + 1] * sigma ** order
    taps = ref32.gauss_taps(sigma, order, sigma ** order, 4.0, round_up=False)
    assert len(taps) == 2 * R + 1
    np.testing.assert_allclose(taps, scipy_taps, rtol=2e-6, atol=2e-7 * np.abs(scipy_taps).max())


def test_tap_arm_rounding_modes():
    assert len(ref32.gauss_taps(1.1, 0, 1.0, 4.0, round_up=True)) == 2 * 5 + 1   # ceil(4.4)
    assert len(ref32.gauss_taps(1.1, 0, 1.0, 4.0, round_up=False)) == 2 * 4 + 1  # int(4.9)
    t0 = ref32.gauss_taps(2.0, 0)
    t1 = ref32.gauss_taps(2.0, 1)
    assert abs(t0.sum() - 1.0) < 1e-6
    np.testing.assert_array_equal(t1, -t1[::-1])  # odd order: antisymmetric


# ------------------------------------------------------------ convolution ----

@pytest.mark.parametrize("axis", [0, 1, 2])
def test_convolution_matches_scipy_mirror(axis, rng):
    img = rng.normal(size=(9, 11, 13)).astype(np.float32)
    sigma, order = 1.7, 1
    taps = ref32.gauss_taps(sigma, order)
    got = ref32.convolve_axis(img, axis, taps)
    want = ndi.gaussian_filter1d(img.astype(np.float64), sigma, axis=axis, order=order,
                                 mode="mirror", truncate=4.0)
    np.testing.assert_allclose(got, want, atol=2e-6 * np.abs(want).max())


def test_convolution_kernel_longer_than_axis(rng):
    # the mirror indexing must wrap several periods (kernel arm > axis length)
    img = rng.normal(size=(3, 4, 5)).astype(np.float32)
    taps = ref32.gauss_taps(3.0, 2, 9.0)
    got = ref32.convolve_axis(img, 2, taps)
    want = fp64.convolve_axis(img, 2, taps.astype(np.float64))
    np.testing.assert_allclose(got, want, atol=1e-5 * np.abs(want).max())


def test_convolution_is_true_convolution_not_correlation():
    # a linear ramp along axis 2: the order-1 derivative must be +1 (a
    # correlation would give -1)
    img = np.broadcast_to(np.arange(40, dtype=np.float32), (2, 3, 40)).copy()
    d = ref32.convolve_axis(img, 2, ref32.gauss_taps(1.0, 1))
    np.testing.assert_allclose(d[:, :, 10:30], 1.0, rtol=1e-4)


def test_hessian_matches_scipy(rng):
    img = ndi.gaussian_filter(rng.normal(size=(14, 12, 16)), 1.0).astype(np.float32)
    sig = (1.5, 1.2, 1.8)
    got = ref32.hessian(img, sig)
    want = fp64.hessian_scipy(img, sig)
    np.testing.assert_allclose(got, want, atol=3e-6 * np.abs(want).max())
    # the float64 restatement agrees as well
    np.testing.assert_allclose(fp64.hessian(img, sig), want, atol=1e-12 * np.abs(want).max())


def test_hessian_layout_on_quadratic_field():
    # f = u^2 + 2 v^2 + 3 u w (u, v, w centred coordinates): H is
    # [[2,0,3],[0,4,0],[3,0,0]] up to the sampling of the discrete kernels,
    # which is small at sigma = 3; pins the AoS layout [H00,H11,H22,H01,H02,H12]
    z = np.arange(48.0) - 24
    u, v, w = np.meshgrid(z, z, z, indexing="ij")
    f = (u ** 2 + 2 * v ** 2 + 3 * u * w).astype(np.float32) / 100
    h = ref32.hessian(f, 3.0)[24, 24, 24] * 100 / 9.0
    np.testing.assert_allclose(h, [2, 4, 0, 0, 3, 0], rtol=0.01, atol=0.01)


# ------------------------------------------------------------ eigensolver ----

def _random_sym(rng, n):
    a = rng.normal(size=(n, 3, 3))
    m = (a + np.swapaxes(a, 1, 2)) / 2
    return np.stack([m[:, 0, 0], m[:, 1, 1], m[:, 2, 2], m[:, 0, 1], m[:, 0, 2], m[:, 1, 2]], -1)


def test_eigs_match_numpy_eigh(rng):
    h6 = _random_sym(rng, 4000).astype(np.float32)
    e = ref32.eigs(h6)
    lam, vec = fp64.eig_sorted(h6)
    np.testing.assert_allclose(e[:, :3], lam, atol=2e-6 * np.abs(lam).max())
    R = fp64.squeezed_to_mat(e[:, 3:])
    gap = np.min(np.abs(np.diff(np.sort(np.abs(lam), -1), axis=-1)), -1)
    ok = gap > 1e-3
    for k in range(3):
        cosang = np.abs(np.sum(R[ok, :, k] * vec[ok, :, k], -1))
        assert np.arccos(np.clip(cosang, -1, 1)).max() < 1e-3
    # reconstruction R diag(l) R^T = H
    H = fp64.h6_to_mat(h6)
    rec = np.einsum("nij,nj,nkj->nik", R, e[:, :3].astype(np.float64), R)
    np.testing.assert_allclose(rec, H, atol=1e-5 * np.abs(H).max())


def test_eigs_zero_matrix_takes_no_sweep_and_sorts_ties_by_swapping():
    # tolerance 0 and off-diagonal sum 0: converged before the first sweep;
    # every compare-exchange of equal |lambda| swaps, giving order (2, 1, 0)
    e = ref32.eigs(np.zeros((1, 6), np.float32))[0]
    np.testing.assert_array_equal(e[:3], 0.0)
    R = fp64.squeezed_to_mat(e[3:])
    np.testing.assert_allclose(np.abs(R), np.eye(3)[:, [2, 1, 0]], atol=1e-7)
    np.testing.assert_allclose(np.linalg.det(R), 1.0, atol=1e-7)


@pytest.mark.parametrize("diag", [(1, 2, 3), (1, 3, 2), (2, 1, 3), (2, 3, 1), (3, 1, 2), (3, 2, 1),
                                  (-3, 2, -1)])
def test_eigs_diagonal_exercises_every_sorter(diag):
    h6 = np.array([[*diag, 0, 0, 0]], np.float32)
    e = ref32.eigs(h6)[0]
    order = np.argsort(np.abs(diag), kind="stable")
    np.testing.assert_array_equal(e[:3], np.asarray(diag, np.float32)[order])
    R = fp64.squeezed_to_mat(e[3:])
    np.testing.assert_allclose(np.abs(R), np.eye(3)[:, order], atol=1e-6)
    np.testing.assert_allclose(np.linalg.det(R), 1.0, atol=1e-6)


def test_sorter_table_is_the_signed_permutation_rotation():
    tab = ref32.sorter_table()  # scalar last
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    for p in perms:
        idx = (p[0] << 1) | (p[1] > p[2])
        inv = sum(p[i] > p[j] for i in range(3) for j in range(i + 1, 3))
        P = np.zeros((3, 3))
        P[np.arange(3), p] = -1 if inv % 2 else 1
        M = Rotation.from_quat(tab[idx].astype(np.float64)).as_matrix()
        np.testing.assert_allclose(M, P, atol=1e-7)


def test_eigs_nonconvergence_gives_nan(rng):
    h6 = _random_sym(rng, 10).astype(np.float32)
    e = ref32.eigs(h6, maxiter=0)
    assert np.isnan(e).all()


# ------------------------------------------------------------ quaternions ----

def test_squeeze_is_intrinsic_zxz_euler(rng):
    R = Rotation.random(500, random_state=7)
    q = R.as_quat()[:, [3, 0, 1, 2]].astype(np.float32)  # scalar first
    v = ref32.qsqueeze(q)
    np.testing.assert_allclose(fp64.squeezed_to_mat(v), R.as_matrix(), atol=2e-6)
    # away from gimbal lock the angles are scipy's intrinsic ZXZ Euler angles
    eul = R.as_euler("ZXZ")
    ok = (eul[:, 1] > 1e-2) & (eul[:, 1] < np.pi - 1e-2)
    d = np.angle(np.exp(1j * (v[ok] - eul[ok])))
    assert np.abs(d).max() < 1e-4


def test_expand_squeeze_round_trip_up_to_sign(rng):
    q = Rotation.random(300, random_state=3).as_quat()[:, [3, 0, 1, 2]].astype(np.float32)
    back = ref32.qexpand(ref32.qsqueeze(q))
    s = np.sign(np.sum(back * q, -1, keepdims=True))
    np.testing.assert_allclose(back * s, q, atol=2e-6)


def test_gimbal_branches():
    for q in ([1, 0, 0, 0], [0.6, 0, 0, 0.8], [0, 1, 0, 0], [0, 0.6, 0.8, 0]):
        q = np.asarray([q], np.float32)
        v = ref32.qsqueeze(q)
        M = fp64.squeezed_to_mat(v)[0]
        want = Rotation.from_quat(np.asarray(q[0])[[1, 2, 3, 0]].astype(np.float64)).as_matrix()
        np.testing.assert_allclose(M, want, atol=1e-6)


def test_mat2quat_quat2mat_round_trip():
    q = Rotation.random(300, random_state=5).as_quat()[:, [3, 0, 1, 2]].astype(np.float32)
    back = ref32.mat2quat(ref32.quat2mat(q))
    s = np.sign(np.sum(back * q, -1, keepdims=True))
    np.testing.assert_allclose(back * s, q, atol=3e-6)


def test_qrot_matches_matrix(rng):
    q = Rotation.random(200, random_state=9).as_quat()[:, [3, 0, 1, 2]].astype(np.float32)
    v = rng.normal(size=(200, 3)).astype(np.float32)
    got = ref32.qrot(q, v)
    want = np.einsum("nij,nj->ni", fp64.quat2mat(q), v)
    np.testing.assert_allclose(got, want, atol=3e-6)


# ------------------------------------------------------------- vesselness ----

def test_vesselness_formula_matches_fp64(rng):
    lam = np.sort(rng.normal(size=(3000, 3)), -1)
    lam = np.take_along_axis(lam, np.argsort(np.abs(lam), -1), -1)
    e6 = np.zeros((3000, 6), np.float32)
    e6[:, :3] = lam
    e6[:, 3:] = rng.normal(size=(3000, 3))
    val = np.full(3000, -np.inf, np.float32)
    quat = np.zeros((3000, 3), np.float32)
    c = ref32.half_max_norm(e6)
    assert abs(c - 0.5 * np.sqrt((lam.astype(np.float32).astype(np.float64) ** 2).sum(-1)).max()) < 1e-6
    ref32.update_vesselness(val, quat, e6, 0.5, 0.5, c, True)
    want = fp64.vesselness_from_eigs(e6[:, :3], 0.5, 0.5, c, True)
    np.testing.assert_allclose(val, want, atol=2e-6)
    # the winning frame is copied exactly
    np.testing.assert_array_equal(quat, e6[:, 3:])
    # dark objects: sign flip
    val2 = np.full(3000, -np.inf, np.float32)
    ref32.update_vesselness(val2, quat, e6, 0.5, 0.5, c, False)
    np.testing.assert_allclose(val2, fp64.vesselness_from_eigs(e6[:, :3], 0.5, 0.5, c, False),
                               atol=2e-6)


def test_running_max_uses_strict_greater():
    e6 = np.array([[0, -1, -2, 0.1, 0.2, 0.3]], np.float32)
    val = np.full(1, -np.inf, np.float32)
    quat = np.zeros((1, 3), np.float32)
    ref32.update_vesselness(val, quat, e6, 0.5, 0.5, 1.0, True)
    first = val.copy()
    e6b = e6.copy()
    e6b[0, 3:] = 9
    ref32.update_vesselness(val, quat, e6b, 0.5, 0.5, 1.0, True)  # equal response
    assert val[0] == first[0]
    np.testing.assert_array_equal(quat[0], e6[0, 3:])


def test_zero_image_keeps_minus_infinity():
    # H = 0 everywhere -> c = 0 -> S^2 / c^2 = 0/0 = NaN -> 'v > val' fails
    img = np.zeros((6, 6, 6), np.float32)
    val, _ = ref32.vesselness(img, [[1.0, 1.0, 1.0]])
    assert np.isneginf(val).all()


def test_constant_image_is_not_flat_for_discrete_kernels():
    # the sampled order-2 kernel has a small DC response, so a constant image
    # gives the same small non-zero Hessian everywhere (no NaN, finite values)
    img = np.ones((6, 6, 6), np.float32)
    val, _ = ref32.vesselness(img, [[1.0, 1.0, 1.0]])
    assert np.isfinite(val).all() and np.ptp(val) == 0
