"""The reference driver's imports (phantom_experiment.py:12-20) resolve
unchanged to this package's modules, and the host helpers behind them keep
their documented behaviour. CPU only."""
import numpy as np
import pytest
from scipy.integrate import quad


def test_reference_driver_imports_resolve():
    from helpers.cl_helpers import ClProgramLoader
    from helpers.helix_phantom import create_toric_helix
    from helpers.misc import normalize
    from helpers.parameters import Parameters
    from helpers.status import Status
    from processing.adjust_directions import AdvancedAdjuster, TransformationCompleter
    from processing.continuous_cut_anisotropic_gpu import CmfCutQAnisotropicGpu
    from processing.gradient_vector_flow import GvfGpu, GvfParameters
    from processing.qfrangi_gpu import FeatureFilter, VesselnessCalculation

    import miccai2016_amd.helpers.cl_helpers as ch
    import miccai2016_amd.processing.adjust_directions as ad
    import miccai2016_amd.processing.continuous_cut_anisotropic_gpu as cc
    import miccai2016_amd.processing.gradient_vector_flow as gv
    import miccai2016_amd.processing.qfrangi_gpu as qf
    assert ClProgramLoader is ch.ClProgramLoader
    assert (AdvancedAdjuster, TransformationCompleter) == (ad.AdvancedAdjuster,
                                                           ad.TransformationCompleter)
    assert CmfCutQAnisotropicGpu is cc.CmfCutQAnisotropicGpu
    assert (GvfGpu, GvfParameters) == (gv.GvfGpu, gv.GvfParameters)
    assert (FeatureFilter, VesselnessCalculation) == (qf.FeatureFilter, qf.VesselnessCalculation)
    assert callable(create_toric_helix) and callable(normalize) and callable(Status)
    assert issubclass(CmfCutQAnisotropicGpu.CutParameters, Parameters)
    import processing
    import processing.slicing
    assert processing.qfrangi_gpu is qf


def test_normalize_is_an_affine_map_onto_the_destination():
    from helpers.misc import normalize
    rng = np.random.default_rng(3)
    for dtype in (np.float32, np.float64):
        x = (rng.normal(size=5000) * 7 - 2).astype(dtype)
        y = normalize(x)
        assert y.dtype == np.float64
        # the width max - min is taken in the data's dtype (as the reference
        # does), so the maximum lands on 1 within that dtype's rounding
        assert y[np.argmin(x)] == 0.0 and abs(y[np.argmax(x)] - 1.0) <= np.finfo(dtype).eps
        # affine: a least-squares line through (x, y) fits exactly, slope > 0
        A = np.stack([x.astype(np.float64), np.ones(x.size)], 1)
        coef, res, *_ = np.linalg.lstsq(A, y, rcond=None)
        assert coef[0] > 0 and np.abs(A @ coef - y).max() <= 1e-12
        z = normalize(x, -3.0, 5.0)
        np.testing.assert_allclose(z, y * 8.0 - 3.0, rtol=0, atol=1e-12)
    assert np.array_equal(normalize(np.full(4, 2.5)), np.zeros(4))
    assert np.array_equal(normalize(np.full(4, 2.5), 7.0, 9.0), np.full(4, 7.0))
    # percentiles: the 10th and 90th land on the ends, clipped beyond them
    x = np.arange(101.0)
    y = normalize(x, perc_low=10, perc_up=90)
    np.testing.assert_allclose(y[[10, 50, 90]], [0.0, 0.5, 1.0], atol=1e-12)
    assert y.min() == 0.0 and y.max() == 1.0
    y = normalize(x, perc_low=10, perc_up=90, clip=False)
    np.testing.assert_allclose(y[[0, 100]], [-0.125, 1.125], atol=1e-12)


def _pappus_volume(r_helix, r_tube, n_windings, factor_helix, factor_tube):
    """Volume swept by the cross-section disc about axis 2: at every azimuth
    the half-plane section is a disc of area pi r_t^2 centred at radius r_h,
    so V = int 2 pi (pi r_t(t)^2) r_h(t) dt over the windings (Pappus; the
    ascent along axis 2 does not change it while windings do not overlap)."""
    rh = lambda t: r_helix * ((factor_helix - 1) * t / n_windings + 1)
    rt = lambda t: r_tube * ((factor_tube - 1) * t / n_windings + 1)
    return quad(lambda t: 2 * np.pi * np.pi * rt(t) ** 2 * rh(t), 0, n_windings)[0]


@pytest.mark.parametrize("args", [(25.0, 6.0, 0.1, 3.0, 0.5, 0.5, 1.0),   # the experiment's
                                  (20.0, 4.0, 0.15, 2.0, 1.0, 1.0, 1.0)])
def test_toric_helix_volume_and_centre_line(args):
    from helpers.helix_phantom import create_toric_helix
    r_helix, r_tube, slope, nw, fh, ft, fs = args
    h = create_toric_helix(*args)
    assert h.dtype == np.uint8 and set(np.unique(h)) == {0, 100}
    want = _pappus_volume(r_helix, r_tube, nw, fh, ft)
    assert abs((h > 0).sum() - want) <= 0.01 * want
    # with margins of 2 voxels every face of the volume is background and
    # the tube keeps its size
    h2 = create_toric_helix(*args, margins=2.0)
    assert np.abs(np.array(h2.shape) - np.array(h.shape) - 4).max() <= 1
    assert not h2[[0, -1]].any() and not h2[:, [0, -1]].any() and not h2[:, :, [0, -1]].any()
    assert abs(int((h2 > 0).sum()) - int((h > 0).sum())) <= 0.002 * want


def test_status_prints_start_and_end(capsys):
    from helpers.status import Status, StatusDecorator
    s = Status("step one")
    s.update("step two")
    s.fail()
    del s
    out = capsys.readouterr().out.splitlines()
    assert out[0] == "step one ..." and out[1].startswith("step one ... done (")
    assert out[2] == "step two ..." and out[3].startswith("step two ... FAILED (")
    quiet = Status("hidden", verbose=False)
    del quiet

    @StatusDecorator("work")
    def f(x):
        return 2 * x
    assert f(3) == 6
    out = capsys.readouterr().out.splitlines()
    assert out[0] == "work ..." and out[1].startswith("work ... done (")
