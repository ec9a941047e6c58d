"""The device form of the direction adjustment (mh_adjust_ridge_device)
against the host construction it replaces (spanning forest, cleaning,
segment tree and signs in host C++, adjust_directions.py's execute_tensor
with device_adjust off), bit for bit, on the workloads' own preparations:
the 64^3 helix at three settings, the 256^3 helix, a ragged 48 x 80 x 61
volume, without cleaning, and the cases the device form hands back."""
import numpy as np
import pytest

from miccai2016_amd import _lib
from miccai2016_amd import phantom as ph

pytestmark = pytest.mark.gpu


def _prepared(dev, spec, scales):
    from miccai2016_amd.helpers.misc import normalize
    from miccai2016_amd.processing.qfrangi_gpu import FeatureFilter, VesselnessCalculation
    img, _ = ph.make_helix(spec)
    vp = VesselnessCalculation.Parameters()
    vp.scales, vp.bright_objects = scales, True
    val_t, quat_t = FeatureFilter(vp, dev).execute_tensor(
        dev.upload(img.astype(np.float32)).reshape(-1), img.shape, (1.0, 1.0, 1.0))
    ves = normalize(dev.download(val_t).reshape(img.shape)).astype(np.float32)
    return quat_t, dev.upload(ves).reshape(-1), img.shape


def _both(dev, quat_t, ves_t, shape, clean=True, voxel_size=(1.0, 1.0, 1.0)):
    from miccai2016_amd.processing.adjust_directions import AdvancedAdjuster
    out = []
    for device in (True, False):
        adj = AdvancedAdjuster((0.01, 0.3), None, None, voxel_size, clean_ridge=clean, loader=dev)
        adj.device_adjust = device
        d = dev.download(adj.execute_tensor(quat_t, ves_t, shape)).reshape(shape + (3,))
        out.append((d, np.asarray(adj.ridge, bool), adj))
    return out


CASES = [
    ("64-workload", ph.HelixSpec((64, 64, 64), 4.0, 16.0, 24.0, 1.5), (1.0, 2.0, 3)),
    ("64-noise-0.5", ph.HelixSpec((64, 64, 64), 4.0, 16.0, 24.0, 0.5), (1.0, 2.0, 3)),
    ("64-matched-scales", ph.HelixSpec((64, 64, 64), 4.0, 16.0, 24.0, 1.5), (4 / 1.2, 4 * 1.2, 16)),
    ("ragged-48x80x61", ph.HelixSpec((48, 80, 61), 3.0, 14.0, 20.0, 1.0), (1.0, 2.0, 3)),
    ("256-workload", ph.PLAIN_256, (1.0, 2.0, 3)),
]


@pytest.mark.parametrize("name,spec,scales", CASES, ids=[c[0] for c in CASES])
def test_device_adjustment_equals_host(dev, name, spec, scales):
    quat_t, ves_t, shape = _prepared(dev, spec, scales)
    (dd, rd, ad), (dh, rh, _) = _both(dev, quat_t, ves_t, shape)
    st = ad.device_stats
    print(f"{name}: device form {'ran' if ad.device_used else 'deferred'}; ridge points {st[0]}, "
          f"segments {st[1]}, kept {st[2]} with {st[3]} points, zero products {st[4]}, "
          f"forest rounds {st[7]}, full-tree walk {st[5] / 1e3:.2f} ms, call {st[6] / 1e3:.2f} ms",
          flush=True)
    assert ad.device_used
    assert np.array_equal(rd, rh)
    assert np.array_equal(dd, dh)


@pytest.mark.parametrize("scales", [(4 / 1.2, 4 * 1.2, 16), (1.0, 2.0, 3)],
                         ids=["matched-scales", "workload"])
def test_device_adjustment_without_cleaning(dev, scales):
    """Every segment kept: at the workload's scales the tree joins 161
    segments, through the segment tree's Boruvka rounds on the device."""
    spec = ph.HelixSpec((64, 64, 64), 4.0, 16.0, 24.0, 1.5)
    quat_t, ves_t, shape = _prepared(dev, spec, scales)
    (dd, rd, ad), (dh, rh, _) = _both(dev, quat_t, ves_t, shape, clean=False)
    print("without cleaning: device form", "ran" if ad.device_used else "deferred",
          ad.device_stats, flush=True)
    assert ad.device_used
    assert np.array_equal(rd, rh)
    assert np.array_equal(dd, dh)


def test_device_adjustment_defers_other_voxel_sizes(dev):
    spec = ph.HelixSpec((40, 48, 44), 3.0, 12.0, 20.0, 1.0)
    quat_t, ves_t, shape = _prepared(dev, spec, (1.0, 2.0, 3))
    for vs in ((1.0, 1.0, 2.0), (0.3, 0.3, 0.3)):
        (dd, rd, ad), (dh, rh, _) = _both(dev, quat_t, ves_t, shape, voxel_size=vs)
        assert not ad.device_used
        assert np.array_equal(rd, rh) and np.array_equal(dd, dh)
    (dd, rd, ad), (dh, rh, _) = _both(dev, quat_t, ves_t, shape, voxel_size=(0.5, 0.5, 0.5))
    assert ad.device_used
    assert np.array_equal(rd, rh) and np.array_equal(dd, dh)


def test_device_adjustment_small_ridge_defers(dev):
    """A ridge of at most 256 points takes the host's dense tree."""
    from miccai2016_amd.processing.adjust_directions import AdvancedAdjuster
    shape = (24, 24, 24)
    n = int(np.prod(shape))
    rng = np.random.default_rng(5)
    resp = np.zeros(shape, np.float32)
    resp[12, 12, 4:20] = 1.0 + 0.01 * rng.random(16).astype(np.float32)
    q = np.zeros(shape + (3,), np.float32)
    adj = AdvancedAdjuster((0.01, 0.3), None, None, 1.0, loader=dev)
    d = dev.download(adj.execute_tensor(dev.upload(q).reshape(-1), dev.upload(resp).reshape(-1),
                                        shape)).reshape(shape + (3,))
    assert not adj.device_used and adj.device_stats[0] <= 256
    assert d.shape == shape + (3,) and n > 0


def _synthetic_ridge(seed, shape=(48, 52, 56), nan_dirs=0):
    """A ridge of straight runs and blobs whose main directions are half exact
    axis vectors, so that many spanning-tree products are exactly zero (cuts
    of the walk) or exactly +-1, and half random unit directions."""
    rng = np.random.default_rng(seed)
    cand = np.zeros(shape, np.float32)
    for _ in range(60):   # straight runs along random axes
        ax = rng.integers(0, 3)
        p = [rng.integers(0, s) for s in shape]
        ln = rng.integers(6, 40)
        for t in range(ln):
            q = list(p)
            q[ax] = min(shape[ax] - 1, p[ax] + t)
            cand[tuple(q)] = 1
    for _ in range(25):   # small blobs
        c = np.array([rng.integers(2, s - 2) for s in shape])
        cand[c[0] - 1:c[0] + 2, c[1] - 1:c[1] + 1, c[2]:c[2] + 2] = 1
    resp = (cand * (0.2 + 0.8 * rng.random(shape))).astype(np.float32)
    axes = np.eye(3, dtype=np.float32)
    pick = rng.integers(0, 6, shape)   # half exact axes, half random (distinct segment lengths)
    sign = np.where(rng.random(shape) < 0.5, -1.0, 1.0).astype(np.float32)
    rnd = rng.normal(size=shape + (3,)).astype(np.float32)
    rnd /= np.linalg.norm(rnd, axis=-1, keepdims=True)
    dirs = np.where((pick < 3)[..., None], axes[np.minimum(pick, 2)] * sign[..., None], rnd)
    if nan_dirs:
        idx = np.argwhere(cand > 0)[:nan_dirs]
        dirs[tuple(idx.T)] = np.nan
    return cand, resp, dirs.astype(np.float32)


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("clean", [True, False], ids=["clean", "all-segments"])
def test_device_adjustment_with_many_zero_products(dev, seed, clean):
    from miccai2016_amd.processing.adjust_directions import AdvancedAdjuster
    cand, resp, dirs = _synthetic_ridge(seed)
    shape = cand.shape
    out = []
    for device in (True, False):
        adj = AdvancedAdjuster((0.01, 0.3), None, None, 1.0, clean_ridge=clean, loader=dev)
        adj.device_adjust = device
        d = adj.execute_tensor(None, dev.upload(resp).reshape(-1), shape,
                               candidates_t=dev.upload(cand).reshape(-1),
                               main_dirs_t=dev.upload(dirs).reshape(-1))
        out.append((dev.download(d).reshape(shape + (3,)), np.asarray(adj.ridge, bool), adj))
    (dd, rd, ad), (dh, rh, _) = out
    st = ad.device_stats
    print(f"seed {seed} {'clean' if clean else 'all segments'}: device form "
          f"{'ran' if ad.device_used else 'deferred'}; ridge points {st[0]}, segments {st[1]}, "
          f"kept {st[2]}, zero products {st[4]}, full-tree walk {st[5] / 1e3:.2f} ms", flush=True)
    assert np.array_equal(rd, rh)
    assert np.array_equal(dd, dh)
    if ad.device_used:
        assert st[4] > 10   # the case exercises the walk's cuts


def test_device_adjustment_defers_nan_products(dev):
    from miccai2016_amd.processing.adjust_directions import AdvancedAdjuster
    cand, resp, dirs = _synthetic_ridge(4, nan_dirs=3)
    shape = cand.shape
    adj = AdvancedAdjuster((0.01, 0.3), None, None, 1.0, loader=dev)
    d = adj.execute_tensor(None, dev.upload(resp).reshape(-1), shape,
                           candidates_t=dev.upload(cand).reshape(-1),
                           main_dirs_t=dev.upload(dirs).reshape(-1))
    assert not adj.device_used
    assert dev.download(d).shape == (int(np.prod(shape)) * 3,)
