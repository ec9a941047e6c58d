"""CPU tests: the C-ABI library loads and exports every declared symbol, and
the host-side logic of the drop-in modules (no GPU compute calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import fp64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "miccai_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mh_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from miccai2016_amd import _lib
    L = _lib.lib()
    names = _declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), f"{name} declared but not exported"
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(names)
    assert L.mh_version() == 100


def test_library_rejects_bad_arguments_without_a_gpu():
    from miccai2016_amd import _lib
    L = _lib.lib()
    assert L.mh_cmf_workspace_bytes(0) == -1
    assert L.mh_cmf_workspace_bytes(1 << 31) == -1
    assert L.mh_cmf_workspace_bytes(64 ** 3) > 9 * 4 * 64 ** 3
    assert L.mh_hessian_work_floats(_lib.dims3((4, 5, 6))) == 9 * 120
    assert L.mh_hessian_work_floats(_lib.dims3((0, 5, 6))) == -1
    assert L.mh_cmf_batched_workspace_bytes(1000, 0) == -1
    assert L.mh_cmf_ragged_workspace_bytes(1000, 3) > 0
    assert L.mh_last_error(None) == b"null context"
    out = C.c_void_p()
    assert L.mh_create(-1, C.byref(out)) != 0


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from miccai2016_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        _lib.lib()


def test_hysteresis_host_matches_labelling(rng):
    from miccai2016_amd.processing.ridge_detection import hysteresis_threshold
    cand = ndi.gaussian_filter(rng.normal(size=(20, 21, 22)), 1.2) > 0.05
    sure = cand & (rng.random(cand.shape) < 0.01)
    got = hysteresis_threshold(cand.astype(np.float32), sure.astype(np.float32))
    want = fp64.hysteresis(cand, sure)
    np.testing.assert_array_equal(got.astype(bool), want)
    # no seed -> empty; every candidate a seed -> all candidates
    assert not hysteresis_threshold(cand.astype(np.float32), np.zeros_like(cand, np.float32)).any()
    full = hysteresis_threshold(cand.astype(np.float32), cand.astype(np.float32))
    np.testing.assert_array_equal(full.astype(bool), cand)


def test_float_threshold_reproduces_double_comparison():
    from miccai2016_amd.processing.ridge_detection import float_threshold
    for t in [0.01, 0.3, 1 / 3, 0.5, -0.1, 1e-8]:
        f = np.float32(float_threshold(t))
        vals = np.array([np.nextafter(f, -1), f, np.nextafter(f, 2),
                         np.nextafter(np.nextafter(f, 2), 2)], np.float32)
        np.testing.assert_array_equal(vals > f, vals.astype(np.float64) > t)


def test_normalize_and_parameters():
    from miccai2016_amd.helpers.misc import normalize
    from miccai2016_amd.helpers.parameters import Parameters
    d = np.array([1.0, 3.0, 2.0], np.float32)
    np.testing.assert_allclose(normalize(d), [0, 1, 0.5])
    np.testing.assert_array_equal(normalize(np.ones(4)), np.zeros(4))
    np.testing.assert_allclose(normalize(d, 2, 4), [2, 4, 3])
    np.testing.assert_allclose(normalize(np.arange(101.), perc_low=10, perc_up=90)[[0, 50, 100]],
                               [0, 0.5, 1])
    np.testing.assert_allclose(normalize(d), fp64.normalize(d))

    class P(Parameters):
        def __init__(self):
            self.a = 1
            self.b = None
    assert P().check() == ["b"]
    with pytest.raises(RuntimeError):
        P().check(raise_e=True)


def test_slicing():
    from miccai2016_amd.processing.slicing import add_margin_to_slc, bounding_box
    m = np.zeros((10, 10, 10), bool)
    m[3:5, 2, 7:9] = True
    bb = bounding_box(m)
    assert bb == (slice(3, 5, 1), slice(2, 3, 1), slice(7, 9, 1))
    assert bounding_box(m, 4)[0] == slice(0, 9, 1)
    assert add_margin_to_slc(2, bb) == (slice(1, 7, 1), slice(0, 5, 1), slice(5, 11, 1))
    assert bounding_box(np.zeros((3, 3)))[0] == slice(0, 0, 1)


def test_derived_cut_constants_and_time_step():
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import (
        CmfCutQAnisotropicGpu, derived_constants)
    from miccai2016_amd.processing.gradient_vector_flow import time_step_for
    p = CmfCutQAnisotropicGpu.CutParameters()
    rcc, rs, ss = derived_constants(p, (1, 1, 1))
    assert rcc == 5.0 and np.allclose(rs, 1) and np.allclose(ss, 0.11)
    rcc, rs, ss = derived_constants(p, (0.5, 1, 2))
    assert np.isclose(rcc, 1 / (0.2 * 0.25)) and np.allclose(ss, 0.11 * 0.25 / np.array([0.5, 1, 2]))
    assert time_step_for((1, 1, 1), 0.1, 3) == 0.99
    assert np.isclose(time_step_for((1, 1, 1), 10 ** 0.5, 3), 1 / (2.01 ** 3 * 10 ** 0.5))


def test_phantom_and_capacities():
    from miccai2016_amd import phantom as ph
    img, truth = ph.make_helix(ph.HelixSpec((32, 32, 32), 3, 8, 16, 0.5, seed=3))
    L = 32 * np.sqrt(1 + (2 * np.pi * 8 / 16) ** 2)
    assert abs(truth.sum() - np.pi * 9 * L) < 0.1 * np.pi * 9 * L
    img2, _ = ph.make_helix(ph.HelixSpec((32, 32, 32), 3, 8, 16, 0.5, seed=3))
    np.testing.assert_array_equal(img, img2)          # seeded
    cs, ct = ph.terminal_capacities(img, 0.5)
    bg, fg = fp64.capacities(img, 0.5)
    np.testing.assert_allclose(cs, fg)
    np.testing.assert_allclose(ct, bg)
    assert ((cs == 0) | (ct == 0)).all()
    np.testing.assert_allclose(ph.grad_mag_sq(img), fp64.grad_mag_sq(img))
    w, d = ph.grid_pairs()
    assert len(w) == 425 and np.isclose(w[0], 10 ** -2.5) and np.isclose(d[1], 10 ** -0.75)
    assert ph.fast_dice(truth, truth) == 1.0


class _NoDevice:
    """Stand-in loader for host-only steps of the adjuster."""


def test_adjuster_mst_sign_propagation_and_cleaning(rng):
    from miccai2016_amd.processing.adjust_directions import AdvancedAdjuster
    adj = AdvancedAdjuster((0.01, 0.3), None, None, voxel_size=1, loader=_NoDevice())
    shape = (12, 12, 12)
    ridge = np.zeros(shape, np.float32)
    ridge[2:10, 3, 3] = 1            # a straight segment
    ridge[5, 8, 2:9] = 1             # a second one
    ridge[10, 10, 10] = 1            # an isolated voxel
    base = np.zeros(shape + (3,), np.float32)
    base[..., 0] = 1
    flip = np.where(rng.random(shape) < 0.5, -1, 1).astype(np.float32)
    main_dirs = base * flip[..., None]
    coords, dirs = adj._adjust_ridge_directions(ridge, main_dirs)
    assert coords.dtype == np.uint16 and len(coords) == ridge.sum()
    # every output is +- its input and all agree in sign with ridge point 0
    np.testing.assert_array_equal(np.abs(dirs), np.abs(main_dirs[tuple(coords.T.astype(int))]))
    assert (dirs[:, 0] == dirs[0, 0]).all()
    # the reference procedure (scipy MST + traversal) gives the same signs
    want = fp64.mst_signs(coords.astype(int), main_dirs[tuple(coords.T.astype(int))])
    np.testing.assert_array_equal(dirs, want)
    cleaned = adj._do_clean_ridge(ridge, coords, dirs)
    # ceil(3 ln(e + 3)) = 6 >= 3 segments: all kept
    np.testing.assert_array_equal(cleaned, ridge)


def test_adjuster_cleaning_keeps_longest_segments():
    from miccai2016_amd.processing.adjust_directions import AdvancedAdjuster
    adj = AdvancedAdjuster(0.01, None, None, voxel_size=1, loader=_NoDevice())
    shape = (40, 40, 40)
    ridge = np.zeros(shape, np.float32)
    lengths = list(range(1, 12))                    # 11 segments -> keep ceil(3 ln(e + 11)) = 8
    for k, L in enumerate(lengths):
        ridge[3 * k + 2, 1:1 + L, 5] = 1
    dirs_field = np.zeros(shape + (3,), np.float32)
    dirs_field[..., 1] = 1
    coords, dirs = adj._adjust_ridge_directions(ridge, dirs_field)
    cleaned = adj._do_clean_ridge(ridge, coords, dirs)
    kept = [cleaned[3 * k + 2].sum() for k in range(len(lengths))]
    assert kept == [0, 0, 0] + lengths[3:]
