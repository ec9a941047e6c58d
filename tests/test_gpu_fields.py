"""GPU parity of the per-voxel fields around the max-flow (capacities,
|grad I|^2, nonterminal weight, min-max normalisation, layouts, initial u,
Dice counts, the helix generator) and of the device-resident preparation.

Tolerances: FP32 device results against float64 host formulas within
2e-6 relative to the field's maximum (capacities, weights, normalisation),
exact for layouts, integer counts and the initial u (one correctly rounded
division on both sides).
"""
import numpy as np
import pytest

from miccai2016_amd import _lib
from miccai2016_amd import phantom as ph

pytestmark = pytest.mark.gpu


def _img(rng, shape=(17, 23, 30), sigma=0.7):
    spec = ph.HelixSpec(shape, 3.0, min(shape[1:]) / 4, 14.0, sigma, seed=5)
    return ph.make_helix(spec)


def test_terminal_capacities(dev, rng):
    img, _ = _img(rng)
    x = img.astype(np.float32)
    n = x.size
    cs, ct, x_t = dev.empty(n), dev.empty(n), dev.upload(x)
    dev.check(dev.lib.mh_terminal_capacities(dev.handle, dev.ptr(x_t), n, 0.7, 1e-9,
                                             dev.ptr(cs), dev.ptr(ct), dev.stream))
    want_s, want_t = ph.terminal_capacities(x.astype(np.float64), 0.7)
    got_s, got_t = dev.download(cs).reshape(x.shape), dev.download(ct).reshape(x.shape)
    assert np.abs(got_s - want_s).max() <= 2e-6 * want_s.max()
    assert np.abs(got_t - want_t).max() <= 2e-6 * want_t.max()
    # exactly one of the two is nonzero, as max(r, 0) / max(-r, 0) make them
    assert not np.any((got_s > 0) & (got_t > 0))


def test_grad_mag_sq_and_weight(dev, rng):
    for shape in [(17, 23, 30), (2, 9, 5), (1, 12, 7)]:
        x = rng.normal(size=shape).astype(np.float32)
        g, x_t = dev.empty(x.size), dev.upload(x)
        dev.check(dev.lib.mh_grad_mag_sq(dev.handle, dev.ptr(x_t), _lib.dims3(shape),
                                         dev.ptr(g), dev.stream))
        got = dev.download(g).reshape(shape)
        if shape[0] == 1:
            # numpy.gradient refuses an axis of length 1; that axis contributes 0 here
            parts = np.gradient(x.astype(np.float64)[0])
            want = parts[0] ** 2 + parts[1] ** 2
            want = want[None]
        else:
            want = ph.grad_mag_sq(x.astype(np.float64))
        assert np.abs(got - want).max() <= 2e-6 * want.max()
        a = dev.empty(x.size)
        dev.check(dev.lib.mh_nonterminal_weight(dev.handle, dev.ptr(g), x.size, 0.3, 2.5,
                                                dev.ptr(a), dev.stream))
        wa = ph.nonterminal_weight(got.astype(np.float64), 0.3, 2.5)
        assert np.abs(dev.download(a).reshape(shape) - wa).max() <= 2e-6 * 0.3


def test_minmax_normalize(dev, rng):
    from miccai2016_amd.helpers.misc import normalize
    for n in [1, 1000, 3 * 2 ** 20 + 17]:
        x = (rng.normal(size=n) * 5 + 2).astype(np.float32)
        out, work = dev.empty(n), dev.empty(int(dev.lib.mh_normalize_work_floats(n)))
        x_t = dev.upload(x)
        dev.check(dev.lib.mh_minmax_normalize(dev.handle, dev.ptr(x_t), n, dev.ptr(out),
                                              dev.ptr(work), dev.stream))
        got = dev.download(out)
        want = normalize(x)
        assert np.abs(got - want).max() <= 2e-6
        if n > 1:
            assert got.min() == 0.0 and got[np.argmax(x)] == pytest.approx(1.0, abs=1e-6)
    x = np.full(77, 3.0, np.float32)      # empty range -> zeros, as normalize() gives
    out, work = dev.empty(77), dev.empty(int(dev.lib.mh_normalize_work_floats(77)))
    x_t = dev.upload(x)
    dev.check(dev.lib.mh_minmax_normalize(dev.handle, dev.ptr(x_t), 77, dev.ptr(out),
                                          dev.ptr(work), dev.stream))
    assert (dev.download(out) == 0).all()


def test_layouts(dev, rng):
    n = 5003
    d = rng.normal(size=(n, 3)).astype(np.float32)
    r = rng.uniform(size=n).astype(np.float32)
    v3 = dev.empty(3 * n)
    d_t, r_t = dev.upload(d), dev.upload(r)     # keep the inputs alive across the launch
    dev.check(dev.lib.mh_weighted_dirs(dev.handle, dev.ptr(d_t), dev.ptr(r_t), n, dev.ptr(v3),
                                       dev.stream))
    want = np.ascontiguousarray((d * r[:, None]).T)
    assert np.array_equal(dev.download(v3).reshape(3, n), want)
    back = dev.empty(3 * n)
    dev.check(dev.lib.mh_soa_to_aos3(dev.handle, dev.ptr(v3), n, dev.ptr(back), dev.stream))
    assert np.array_equal(dev.download(back).reshape(n, 3), want.T)


def test_initial_u_bit_exact(dev, rng):
    from miccai2016_amd.processing.continuous_cut_anisotropic_gpu import initial_u
    for n in [1, 4097, 2 ** 21 + 3]:
        x = rng.normal(size=n).astype(np.float32)
        u, work = dev.empty(n), dev.empty(int(dev.lib.mh_initial_u_work_floats(n)))
        x_t = dev.upload(x)
        dev.check(dev.lib.mh_initial_u(dev.handle, dev.ptr(x_t), n, dev.ptr(u),
                                       dev.ptr(work), dev.stream))
        assert np.array_equal(dev.download(u), initial_u(x))
    x[123] = np.nan                       # numpy's max propagates NaN: u is NaN everywhere
    x_t = dev.upload(x)
    dev.check(dev.lib.mh_initial_u(dev.handle, dev.ptr(x_t), x.size, dev.ptr(u),
                                   dev.ptr(work), dev.stream))
    assert np.isnan(dev.download(u)).all()


def test_dice_counts(dev, rng):
    from miccai2016_amd.evaluation import DeviceDice, fast_dice_for
    n, S = 70001, 5
    truth = (rng.uniform(size=n) < 0.1).astype(np.uint8)
    u = rng.uniform(size=(S, n)).astype(np.float32)
    u[1] = 0.0                            # empty labelling
    u[2] = np.where(truth, 0.5, 0.4999999)   # exactly the mask (0.5 counts as foreground)
    dd = DeviceDice(truth, dev)
    u_t = dev.upload(u)
    c = dd.counts(u_t, S)
    for s in range(S):
        b = u[s] >= 0.5
        assert tuple(c[s]) == (b.sum(), (b & (truth != 0)).sum(), truth.sum())
    got = dd(u_t, S)
    np.testing.assert_array_equal(got, [fast_dice_for(u[s], truth) for s in range(S)])
    assert got[2] == 1.0
    empty = DeviceDice(np.zeros(n, np.uint8), dev)(dev.upload(np.zeros(n, np.float32)), 1)
    assert np.isnan(empty[0])


def test_helix_phantom(dev):
    spec = ph.HelixSpec((48, 40, 44), 4.0, 10.0, 24.0, 0.8, seed=3)
    n = int(np.prod(spec.shape))
    img, truth, tan = dev.empty(n), dev.empty(n, dtype=dev.torch.uint8), dev.empty(3 * n)
    dev.check(dev.lib.mh_helix_phantom(dev.handle, _lib.dims3(spec.shape), spec.r_tube,
                                       spec.r_helix, spec.pitch, spec.sigma, 99, dev.ptr(img),
                                       dev.ptr(truth), dev.ptr(tan), dev.stream))
    t = dev.download(truth).reshape(spec.shape)
    dist = ph.helix_distance(spec)
    want = dist <= spec.r_tube
    diff = t.astype(bool) != want
    # masks agree except where the distance is within FP32 rounding of r_tube
    assert np.all(np.abs(dist[diff] - spec.r_tube) < 1e-3)
    x = dev.download(img).reshape(spec.shape) - t
    assert abs(x.mean()) < 0.01 and abs(x.std() - spec.sigma) < 0.01
    v = dev.download(tan).reshape(-1, 3)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)
    # a helix's tangent makes a constant angle with its axis: cos = 1 / sqrt(1 + (Rh w)^2)
    rw = spec.r_helix * 2 * np.pi / spec.pitch
    np.testing.assert_allclose(v[:, 0], 1 / np.sqrt(1 + rw ** 2), atol=1e-6)


def test_prepare_device_matches_host(dev):
    """The device-resident preparation against the numpy-API one: same
    capacities, |grad I|^2 and nearly all frames (the device normalisation is
    FP32, the host one float64, which may move a ridge decision at a
    threshold)."""
    from miccai2016_amd.pipeline import PipelineConfig, prepare, prepare_device
    from oracle import ref32
    spec = ph.HelixSpec((40, 40, 40), 3.0, 10.0, 20.0, 0.6, seed=11)
    img, _ = ph.make_helix(spec)
    cfg = PipelineConfig()
    cfg.gvf_iters = 60
    host = prepare(img, spec.sigma, cfg, dev)
    x = img.astype(np.float32)
    devp = prepare_device(dev.upload(x).reshape(-1), x.shape, spec.sigma, cfg, dev)
    cs, ct = dev.download(devp.cs).reshape(x.shape), dev.download(devp.ct).reshape(x.shape)
    assert np.abs(cs - host.cs).max() <= 2e-6 * host.cs.max()
    assert np.abs(ct - host.ct).max() <= 2e-6 * host.ct.max()
    gms = dev.download(devp.gms).reshape(x.shape)
    # the host |grad I|^2 is float64 of the float64 image; the device's is of the FP32 image
    assert np.abs(gms - host.gms).max() <= 1e-5 * host.gms.max()
    qd = dev.download(devp.quats).reshape(-1, 3)
    qh = host.quats.reshape(-1, 3)
    vd = ref32.quat2mat(ref32.qexpand(qd))[:, :, 0]
    vh = ref32.quat2mat(ref32.qexpand(qh))[:, :, 0]
    cosang = np.abs(np.sum(vd * vh, axis=1))
    assert np.mean(cosang > np.cos(1e-2)) > 0.99


def test_largest_component_dice(dev, rng):
    """dice_for with the components labelled on the GPU: the same labels as
    scipy (face connectivity, smallest voxel index per component), the same
    largest component (first one on ties) and the same Dice."""
    import scipy.ndimage as ndi
    from miccai2016_amd.evaluation import device_dice_for, dice_for, select_largest
    cases = []
    for shape, p in [((30, 41, 37), 0.45), ((64, 64, 64), 0.3), ((5, 200, 3), 0.6), ((1, 1, 9), 0.5)]:
        u = (rng.uniform(size=shape) < p).astype(np.float32) * 0.8 + 0.1
        u.flat[u.size // 2] = 0.9          # at least one foreground voxel
        cases.append((u, shape))
    spec = ph.HelixSpec((48, 52, 50), 4.0, 12.0, 24.0, 0.6, seed=2)
    img, truth_h = ph.make_helix(spec)
    cases.append((img.astype(np.float32), spec.shape))
    tie = np.zeros((4, 5, 6), np.float32)     # two components of 3 voxels: the first wins
    tie[0, 0, :3] = 1
    tie[3, 4, 3:] = 1
    cases.append((tie, tie.shape))
    for u, shape in cases:
        truth = truth_h if shape == spec.shape else (rng.uniform(size=shape) < 0.3).astype(np.uint8)
        lab = dev.empty(int(np.prod(shape)), dtype=dev.torch.int32)
        u_t = dev.upload(u).reshape(-1)
        d, size = device_dice_for(u_t, truth, shape, dev, labels=lab)
        want_comp, want_size = select_largest(u, return_size=True)
        assert size == want_size
        assert d == pytest.approx(dice_for(u, truth), abs=1e-12)
        # labels: scipy's components, each named by its smallest voxel index
        sl, nl = ndi.label(u >= 0.5)
        got = dev.download(lab).reshape(shape)
        assert np.array_equal(got < 0, sl == 0)
        flat = sl.ravel()
        first = np.full(nl + 1, -1, np.int64)
        order = np.argsort(flat, kind="stable")
        starts = np.searchsorted(flat[order], np.arange(1, nl + 1))
        first[1:] = order[starts]
        assert np.array_equal(got.ravel()[flat > 0], first[flat[flat > 0]])
    with pytest.raises(ValueError):
        device_dice_for(dev.zeros(27), np.ones((3, 3, 3), np.uint8), (3, 3, 3), dev)
