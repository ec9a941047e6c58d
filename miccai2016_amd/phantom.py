"""Synthetic helix phantoms and the capacity / weight fields of the grid search.

The tube follows the helix x = cx + Rh cos(2 pi z / P), y = cy + Rh sin(2 pi z / P)
with z along axis 0 and (y, x) on axes (1, 2), centred in-plane. A voxel is
inside when its centre lies within r_tube of the centre line. Intensity is
1 inside, 0 outside, plus i.i.d. N(0, sigma^2) noise drawn in float64 from
numpy's RandomState(seed). The noisy volume stays float64 (the capacities are
computed from it); every GPU stage casts its input to float32.

Capacities follow the reference experiment (phantom_experiment.py:29-35,
280-301, 345, 359): Gaussian log-likelihood ratio weights with eps = 1e-9,
|grad I|^2 from central differences, and the nonterminal weight
C = w exp(-|grad I|^2 / decay^2).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class HelixSpec:
    shape: tuple = (64, 64, 64)
    r_tube: float = 4.0
    r_helix: float = 16.0
    pitch: float = 24.0
    sigma: float = 1.5
    seed: int = 42


# the five workloads of the benchmark description
PLAIN_64 = HelixSpec((64, 64, 64), 4.0, 16.0, 24.0, 1.5)
PLAIN_256 = HelixSpec((256, 256, 256), 8.0, 64.0, 96.0, 1.0)
GRID_128 = HelixSpec((128, 128, 128), 6.0, 32.0, 48.0, 1.5)


def _centre(t, spec, cy, cx):
    w = 2 * np.pi / spec.pitch
    return cx + spec.r_helix * np.cos(w * t), cy + spec.r_helix * np.sin(w * t)


def helix_distance(spec):
    """Distance (float64) from every voxel centre to the helix centre line,
    within r_tube + 1 of it (larger distances are reported as +inf)."""
    D0, D1, D2 = spec.shape
    cy, cx = (D1 - 1) / 2.0, (D2 - 1) / 2.0
    w = 2 * np.pi / spec.pitch
    reach = spec.r_tube + 1.0
    dist = np.full(spec.shape, np.inf)
    yy, xx = np.mgrid[0:D1, 0:D2].astype(np.float64)
    # centre-line samples cover every t within `reach` of the slice
    step = 0.25
    for z in range(D0):
        ts = np.arange(z - reach - 1, z + reach + 1 + step, step)
        hx, hy = _centre(ts, spec, cy, cx)
        lo_y = max(int(np.floor(hy.min() - reach)), 0)
        hi_y = min(int(np.ceil(hy.max() + reach)) + 1, D1)
        lo_x = max(int(np.floor(hx.min() - reach)), 0)
        hi_x = min(int(np.ceil(hx.max() + reach)) + 1, D2)
        if lo_y >= hi_y or lo_x >= hi_x:
            continue
        py = yy[lo_y:hi_y, lo_x:hi_x][..., None]
        px = xx[lo_y:hi_y, lo_x:hi_x][..., None]
        d2 = (px - hx) ** 2 + (py - hy) ** 2 + (z - ts) ** 2
        k = np.argmin(d2, axis=-1)
        t = ts[k]
        # Newton refinement of the closest parameter t
        for _ in range(4):
            ex, ey = _centre(t, spec, cy, cx)
            dxp, dyp = -spec.r_helix * w * np.sin(w * t), spec.r_helix * w * np.cos(w * t)
            rx, ry, rz = px[..., 0] - ex, py[..., 0] - ey, z - t
            g = -(rx * dxp + ry * dyp + rz)
            h = dxp ** 2 + dyp ** 2 + 1 + rx * (spec.r_helix * w * w * np.cos(w * t)) + \
                ry * (spec.r_helix * w * w * np.sin(w * t))
            t = t - g / np.where(np.abs(h) > 1e-9, h, 1.0)
        ex, ey = _centre(t, spec, cy, cx)
        dd = np.sqrt((px[..., 0] - ex) ** 2 + (py[..., 0] - ey) ** 2 + (z - t) ** 2)
        dd = np.minimum(dd, np.sqrt(d2.min(-1)))
        sl = dist[z, lo_y:hi_y, lo_x:hi_x]
        dist[z, lo_y:hi_y, lo_x:hi_x] = np.where(dd <= reach, dd, sl)
    return dist


def make_helix(spec):
    """Returns (noisy float64, truth uint8) volumes of spec.shape."""
    truth = (helix_distance(spec) <= spec.r_tube).astype(np.uint8)
    rng = np.random.RandomState(spec.seed)
    noise = rng.normal(size=spec.shape, scale=spec.sigma)
    return truth + noise, truth


def normpdf(x, mean, var):
    """Normal density with the given mean and variance (phantom_experiment.py:29-35)."""
    return np.exp((x - mean) ** 2 * (1 / (-2 * var))) * (1 / np.sqrt(2 * np.pi * var))


def terminal_capacities(image, noise_level, eps=1e-9):
    """(cs, ct) = (foreground, background) weights from the log-likelihood
    ratio of N(1, s^2) vs N(0, s^2), scaled by 1 / log((p_hat + eps) / eps),
    float64 (phantom_experiment.py:280-301)."""
    var = noise_level ** 2
    image = np.asarray(image, dtype=np.float64)
    bgp = normpdf(image, 0, var)
    fgp = normpdf(image, 1, var)
    p_hat = np.maximum(normpdf(0, 0, var), normpdf(1, 1, var))
    r = np.log((fgp + eps) / (bgp + eps))
    rq = 1 / np.log((p_hat + eps) / eps)
    return rq * np.maximum(r, 0), rq * np.maximum(-r, 0)


def grad_mag_sq(image):
    """|grad I|^2 with numpy's central/one-sided differences (phantom_experiment.py:345)."""
    return np.linalg.norm(np.gradient(np.asarray(image, dtype=np.float64)), axis=0) ** 2


def nonterminal_weight(gms, weighting, decay):
    """C = w exp(-|grad I|^2 / decay^2) (phantom_experiment.py:359)."""
    return weighting * np.exp(((-1.) / (decay ** 2)) * gms)


# the 25 x 17 grid of phantom_experiment.py:167-168
GRID_WEIGHTINGS = np.logspace(-2.5, 0.5, 24 + 1)
GRID_DECAYS = np.logspace(-1, 3, 16 + 1)
A_ATV = 10 ** (-1.5)


def ragged_set(count, seed=0, side=(48, 192)):
    """Specs and (weighting, decay) of a ragged set: dims uniform per axis in
    `side`, tube radius U[3, 8], helix radius a quarter of the smallest side,
    pitch U[20, 80], noise U[0.1, 1.5], (w, decay) drawn from the grid; every
    volume's noise seed is drawn from the master seed."""
    rng = np.random.default_rng(seed)
    specs, w, d = [], [], []
    for _ in range(int(count)):
        dims = tuple(int(v) for v in rng.integers(side[0], side[1] + 1, size=3))
        specs.append(HelixSpec(dims, float(rng.uniform(3, 8)), 0.25 * min(dims),
                               float(rng.uniform(20, 80)), float(rng.uniform(0.1, 1.5)),
                               seed=int(rng.integers(1 << 31))))
        w.append(float(rng.choice(GRID_WEIGHTINGS)))
        d.append(float(rng.choice(GRID_DECAYS)))
    return specs, np.array(w), np.array(d)


def grid_pairs():
    """All (weighting, decay) pairs in the reference's itertools.product order."""
    w, d = np.meshgrid(GRID_WEIGHTINGS, GRID_DECAYS, indexing="ij")
    return w.ravel(), d.ravel()


def fast_dice(seg, truth, threshold=0.5):
    """2 |S & T| / (|S| + |T|) with S = seg >= threshold (phantom_experiment.py:36-47);
    NaN when both are empty, as there."""
    b = np.asarray(seg) >= threshold
    t = np.asarray(truth) != 0
    den = int(b.sum()) + int(t.sum())
    return np.nan if den == 0 else 2.0 * int(np.logical_and(b, t).sum()) / den


def dice_between(a, b, threshold=0.5):
    """Dice overlap of two thresholded segmentations (1 when both are empty)."""
    d = fast_dice(np.asarray(a), np.asarray(b) >= threshold, threshold)
    return 1.0 if np.isnan(d) else d
